// Instantiations of the four-index fill kernels for bra classes 00 10 11.
#include "fill_kernel.h"

void bqc_register_fill_a(FillKernelTable &t) {
    reg_fill_bra<0, 0>(t);
    reg_fill_bra<1, 0>(t);
    reg_fill_bra<1, 1>(t);
}
