// Kernel instantiations of the J/K build for bra class (21| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_21(JKKernelTable &tab) { reg_bra<2, 1>(tab); }
