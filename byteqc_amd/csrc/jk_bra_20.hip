// Kernel instantiations of the J/K build for bra class (20| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_20(JKKernelTable &tab) { reg_bra<2, 0>(tab); }
