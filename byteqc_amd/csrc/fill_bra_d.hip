// Instantiations of the four-index fill kernels for bra classes 32 33.
#include "fill_kernel.h"

void bqc_register_fill_d(FillKernelTable &t) {
    reg_fill_bra<3, 2>(t);
    reg_fill_bra<3, 3>(t);
}
