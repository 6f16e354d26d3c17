// Kernel instantiations of the J/K build for bra class (31| and all
// smaller ket classes.
#include "jk_wave.h"

void bqc_register_bra_31(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3) {
    reg_bra<3, 1>(t1, t2, t3);
}
