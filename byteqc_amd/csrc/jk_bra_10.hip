// Kernel instantiations of the J/K build for bra class (10| and all
// smaller ket classes.
#include "jk_wave.h"

void bqc_register_bra_10(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3) {
    reg_bra<1, 0>(t1, t2, t3);
}
