// Kernel instantiations of the J/K build for bra class (10| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_10(JKKernelTable &tab) { reg_bra<1, 0>(tab); }
