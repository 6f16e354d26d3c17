// Kernel instantiations of the J/K build for bra class (00| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_00(JKKernelTable &tab) { reg_bra<0, 0>(tab); }
