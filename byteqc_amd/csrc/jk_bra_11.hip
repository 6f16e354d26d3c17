// Kernel instantiations of the J/K build for bra class (11| and all
// smaller ket classes.
#include "jk_kernel.h"

void bqc_register_bra_11(JKKernelTable &tab) { reg_bra<1, 1>(tab); }
