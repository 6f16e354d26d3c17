// Instantiations of the four-index fill kernels for bra classes 30 31.
#include "fill_kernel.h"

void bqc_register_fill_c(FillKernelTable &t) {
    reg_fill_bra<3, 0>(t);
    reg_fill_bra<3, 1>(t);
}
