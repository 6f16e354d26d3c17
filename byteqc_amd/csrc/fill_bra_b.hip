// Instantiations of the four-index fill kernels for bra classes 20 21 22.
#include "fill_kernel.h"

void bqc_register_fill_b(FillKernelTable &t) {
    reg_fill_bra<2, 0>(t);
    reg_fill_bra<2, 1>(t);
    reg_fill_bra<2, 2>(t);
}
