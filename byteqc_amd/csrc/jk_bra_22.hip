// Kernel instantiations of the J/K build for bra class (22| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_22(JKKernelTable &tab) { reg_bra<2, 2>(tab); }
