// Kernel instantiations of the J/K build for bra class (30| and all
// smaller ket classes.
#include "jk_v2.h"

void bqc_register_bra_30(JKKernelTable &t1, RowKernelTable &t2) { reg_bra<3, 0>(t1, t2); }
