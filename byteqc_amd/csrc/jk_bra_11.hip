// Kernel instantiations of the J/K build for bra class (11| and all
// smaller ket classes.
#include "jk_wave.h"

void bqc_register_bra_11(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3) {
    reg_bra<1, 1>(t1, t2, t3);
}
