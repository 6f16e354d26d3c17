// Host side of the four-index ERI fill (GINTfill_int2e).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/byteqc_jk.h"
#include "fill_kernel.h"

void bqc_register_fill_a(FillKernelTable &t);
void bqc_register_fill_b(FillKernelTable &t);
void bqc_register_fill_c(FillKernelTable &t);
void bqc_register_fill_d(FillKernelTable &t);

static FillKernelTable g_fill;
static std::once_flag g_fill_once;

static void register_fill() {
    memset(g_fill, 0, sizeof(g_fill));
    bqc_register_fill_a(g_fill);
    bqc_register_fill_b(g_fill);
    bqc_register_fill_c(g_fill);
    bqc_register_fill_d(g_fill);
}

// Largest private (scratch) segment a fill kernel may use per lane.
#define BQC_FILL_MAX_SCRATCH 32768

extern "C" int GINTfill_int2e(BasisProdCache *bp, double *eri, int nao, const int *strides,
                              const int *ao_offsets, const int *bins_locs_ij,
                              const int *bins_locs_kl, int nbins, int cp_ij_id, int cp_kl_id,
                              hipStream_t stream) {
    (void)nao;
    std::call_once(g_fill_once, register_fill);
    if (nbins < 1 || nbins > BQC_MAX_BINS || cp_ij_id < 0 || cp_kl_id < 0 ||
        cp_ij_id >= bp->ncptype || cp_kl_id >= bp->ncptype)
        return 1;
    const ContractionProdType &tij = bp->cptype[cp_ij_id], &tkl = bp->cptype[cp_kl_id];
    const int li = tij.l_bra, lj = tij.l_ket, lk = tkl.l_bra, ll = tkl.l_ket;
    if (li > BQC_LMAX || lj > BQC_LMAX || lk > BQC_LMAX || ll > BQC_LMAX) return 1;
    if ((li + lj + lk + ll) / 2 + 1 >= 12) return 2;
    // pair types keep l(bra) >= l(ket) (groups are sorted by l)
    if (lj > li || ll > lk) return 1;
    const int pij = li * (li + 1) / 2 + lj, pkl = lk * (lk + 1) / 2 + ll;
    const int swap = pkl > pij;
    FillKernel kern = swap ? g_fill[pkl][pij] : g_fill[pij][pkl];
    if (!kern) return 1;
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, (const void *)kern) != hipSuccess ||
        attr.localSizeBytes > BQC_FILL_MAX_SCRATCH)
        return 1;

    EriOut out;
    out.data = eri;
    for (int d = 0; d < 4; ++d) {
        out.stride[d] = strides[d];
        out.off[d] = ao_offsets[d];
    }
    PairData pd;
    pd.a12 = bp->a12;
    pd.e12 = bp->e12;
    pd.x12 = bp->x12;
    pd.y12 = bp->y12;
    pd.z12 = bp->z12;
    pd.bas_x = bp->bas_coords;
    pd.bas_y = bp->bas_coords + bp->nbas;
    pd.bas_z = bp->bas_coords + 2 * bp->nbas;
    pd.bra = bp->bas_pair2bra;
    pd.ket = bp->bas_pair2ket;
    pd.ao_loc = bp->ao_loc;
    TaskArgs ta;
    const int cp_p = swap ? cp_kl_id : cp_ij_id, cp_q = swap ? cp_ij_id : cp_kl_id;
    auto reset = [&]() {
        memset(&ta, 0, sizeof(ta));
        ta.bas_p0 = bp->bas_pairs_locs[cp_p];
        ta.bas_q0 = bp->bas_pairs_locs[cp_q];
        ta.prim_p0 = bp->primitive_pairs_locs[cp_p];
        ta.prim_q0 = bp->primitive_pairs_locs[cp_q];
        ta.nprim_p = bp->cptype[cp_p].nprim_12;
        ta.nprim_q = bp->cptype[cp_q].nprim_12;
        ta.swap = swap;
        ta.diag = 0;
        ta.pre[0] = 0;
    };
    auto launch = [&]() -> int {
        if (ta.nbins == 0) return 0;
        const long long nblk = (ta.pre[ta.nbins] + BQC_THREADS - 1) / BQC_THREADS;
        if (nblk > 0x7fffffffLL) return 1;
        hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(BQC_THREADS), 0, stream, out, pd,
                           ta);
        return hipGetLastError() == hipSuccess ? 0 : 1;
    };
    reset();
    // kl bin b meets the ij pairs of bins [0, nbins - b); at most
    // BQC_ARG_BINS rectangles per launch
    for (int b = 0; b < nbins; ++b) {
        const int k0 = bins_locs_kl[b], k1 = bins_locs_kl[b + 1];
        const int i0 = bins_locs_ij[0], i1 = bins_locs_ij[nbins - b];
        if (k1 <= k0 || i1 <= i0) continue;
        const int r = ta.nbins;
        ta.klb[r] = k0;
        ta.ij0[r] = i0;
        ta.nij[r] = i1 - i0;
        ta.pre[r + 1] = ta.pre[r] + (long long)(k1 - k0) * (i1 - i0);
        ta.nbins = r + 1;
        if (ta.nbins == BQC_ARG_BINS) {
            if (launch()) return 1;
            reset();
        }
    }
    return launch();
}
