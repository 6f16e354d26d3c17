// Four-index ERI fill over the screened shell quartets of two pair types:
// the GINTfill_int2e role of cuobc/lib/gint/nr_fill_ao_ints.cu:493-550 with
// the write-out of cuobc/lib/gint/fill_ints.cu:28-110.
//
// For kl bin b (Schwarz bins of the kl pair type) every kl pair of the bin
// meets the ij pairs of bins [0, nbins - b), as in the J/K build.  Each
// quartet (ij|kl) is computed once, without the diagonal restriction and
// without the factor 1/2 for ij == kl of the J/K path (the pair data keep
// their own diagonal factor), and added to
//   eri[(i - off_i) s_i + (j - off_j) s_j + (k - off_k) s_k + (l - off_l) s_l]
// for the Cartesian components i, j, k, l of the shells (bra, ket) of the
// two pairs.  Distinct quartets write distinct elements, so the update is a
// plain read-modify-write.  One lane evaluates one quartet; classes whose
// block does not fit registers loop over the components of the first shell.
#pragma once
#include "jk_kernel.h"

struct EriOut {
    double *data;
    long long stride[4];
    int off[4];
};

// Adds the block of the evaluated orientation (a b | c d) to the tensor;
// swap: the evaluated p pair is the call's kl pair.
template <int LA, int LB, int LC, int LD, int XN>
__device__ __forceinline__ void fill_write(const double *__restrict__ g, const EriOut &out,
                                           int a0, int b0, int c0, int d0, int swap, int x0) {
    using C = Cls<LA, LB, LC, LD>;
    const long long sp0 = swap ? out.stride[2] : out.stride[0];
    const long long sp1 = swap ? out.stride[3] : out.stride[1];
    const long long sq0 = swap ? out.stride[0] : out.stride[2];
    const long long sq1 = swap ? out.stride[1] : out.stride[3];
    const int op0 = swap ? out.off[2] : out.off[0], op1 = swap ? out.off[3] : out.off[1];
    const int oq0 = swap ? out.off[0] : out.off[2], oq1 = swap ? out.off[1] : out.off[3];
    gdouble *const E = glb_rw(out.data);
#pragma unroll C::UO
    for (int x = 0; x < XN; ++x)
#pragma unroll C::UO
        for (int y = 0; y < C::NFJ; ++y) {
            const long long pbase = (long long)(a0 + x0 + x - op0) * sp0 + (long long)(b0 + y - op1) * sp1;
#pragma unroll
            for (int z = 0; z < C::NFK; ++z)
#pragma unroll
                for (int w = 0; w < C::NFL; ++w) {
                    const long long idx =
                        pbase + (long long)(c0 + z - oq0) * sq0 + (long long)(d0 + w - oq1) * sq1;
                    E[idx] += g[((x * C::NFJ + y) * C::NFK + z) * C::NFL + w];
                }
        }
}

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(BQC_THREADS)
    __attribute__((amdgpu_waves_per_eu(1, Cls<LA, LB, LC, LD>::BIG ? 1 : 8)))
    fill_quartets(EriOut out, PairData pd, TaskArgs ta) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ta.pre[ta.nbins]) return;
    int b = 0;
    while (t >= ta.pre[b + 1]) ++b;
    const long long local = t - ta.pre[b];
    // p side (as evaluated): pairs of the ij list unless swapped
    const int kl = ta.klb[b] + (int)(local / ta.nij[b]);
    const int ij = ta.ij0[b] + (int)(local % ta.nij[b]);
    const int bp = ta.swap ? ta.bas_p0 + kl : ta.bas_p0 + ij;
    const int bq = ta.swap ? ta.bas_q0 + ij : ta.bas_q0 + kl;
    const int pp0 = ta.prim_p0 + (bp - ta.bas_p0) * ta.nprim_p;
    const int qq0 = ta.prim_q0 + (bq - ta.bas_q0) * ta.nprim_q;
    const gptr<int> bra = glb(pd.bra), ket = glb(pd.ket), ao = glb(pd.ao_loc);
    const gptr<double> bx = glb(pd.bas_x), by = glb(pd.bas_y), bz = glb(pd.bas_z);
    const int sa = bra[bp], sb = ket[bp], sc = bra[bq], sd = ket[bq];
    const double ra[3] = {bx[sa], by[sa], bz[sa]};
    const double rb[3] = {bx[sb], by[sb], bz[sb]};
    const double rc[3] = {bx[sc], by[sc], bz[sc]};
    const double rd[3] = {bx[sd], by[sd], bz[sd]};
    using C = Cls<LA, LB, LC, LD>;
    if constexpr (C::BIG) {
        for (int x0 = 0; x0 < C::NFI; ++x0) {
            double g[C::NF / C::NFI];
            eri_block<LA, LB, LC, LD, double, 1>(g, pd, pp0, ta.nprim_p, qq0, ta.nprim_q,
                                                 TWO_PI_52, ra, rb, rc, rd, x0);
            fill_write<LA, LB, LC, LD, 1>(g, out, ao[sa], ao[sb], ao[sc], ao[sd], ta.swap, x0);
        }
    } else {
        double g[C::NF];
        eri_block<LA, LB, LC, LD, double>(g, pd, pp0, ta.nprim_p, qq0, ta.nprim_q, TWO_PI_52, ra,
                                          rb, rc, rd);
        fill_write<LA, LB, LC, LD, C::NFI>(g, out, ao[sa], ao[sb], ao[sc], ao[sd], ta.swap, 0);
    }
}

typedef void (*FillKernel)(EriOut, PairData, TaskArgs);
typedef FillKernel FillKernelTable[BQC_NPAIRID][BQC_NPAIRID];

// all ket classes (LC LD) with pair-id(LC,LD) <= pair-id(LA,LB)
template <int LA, int LB>
static void reg_fill_bra(FillKernelTable &t) {
    constexpr int pab = LA * (LA + 1) / 2 + LB;
    t[pab][0] = fill_quartets<LA, LB, 0, 0>;
    if constexpr (pab >= 1) t[pab][1] = fill_quartets<LA, LB, 1, 0>;
    if constexpr (pab >= 2) t[pab][2] = fill_quartets<LA, LB, 1, 1>;
    if constexpr (pab >= 3) t[pab][3] = fill_quartets<LA, LB, 2, 0>;
    if constexpr (pab >= 4) t[pab][4] = fill_quartets<LA, LB, 2, 1>;
    if constexpr (pab >= 5) t[pab][5] = fill_quartets<LA, LB, 2, 2>;
    if constexpr (pab >= 6) t[pab][6] = fill_quartets<LA, LB, 3, 0>;
    if constexpr (pab >= 7) t[pab][7] = fill_quartets<LA, LB, 3, 1>;
    if constexpr (pab >= 8) t[pab][8] = fill_quartets<LA, LB, 3, 2>;
    if constexpr (pab >= 9) t[pab][9] = fill_quartets<LA, LB, 3, 3>;
}
