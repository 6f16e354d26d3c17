#pragma once
// Rys-quadrature ERI + J/K digestion kernels for gfx950 (kernel templates).
//
// Work decomposition.  One launch per task (pair type cp_ij with pair type
// cp_kl, cp_ij >= cp_kl), as in the reference's per-task call
// (cuobc/lib/gvhf/nr_jk_driver.cu:227-283), but with all Schwarz bins of the
// task folded into one grid: kl pairs of bin b meet the ij pairs of bins
// [0, nbins-b); the work items of bin b form a dense (kl x ij) rectangle and
// the rectangles are laid end to end (prefix array in the task descriptor).
// One lane evaluates one contracted shell quartet: it loops over the
// primitive pairs of both sides, evaluates the Rys roots once per primitive
// quartet and accumulates the Cartesian integral block in registers, then
// digests the block into J and K with FP64 atomics.  ij is the fastest work
// index, so a wavefront shares its kl pair (and its density block) for up to
// 64 lanes.
//
// Unique-quartet weights follow the reference: e12 of a pair with ish == jsh
// carries 1/2 (pair data), and a quartet with bas_ij == bas_kl carries 1/2
// (cuobc/lib/gvhf/g2e.cu:1022-1028).  Digestion (contract_jk.cu:28-316):
//   J: V[i,j] += g D[k,l], V[k,l] += g D[i,j]
//   K: V[i,k] += g D[j,l], V[i,l] += g D[j,k], V[j,k] += g D[i,l], V[j,l] += g D[i,k]
// followed by J = 2 (V + V^T), K = V + V^T (BQCsymmetrize_jk).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/byteqc_jk.h"
#include "bqc_device.h"

using namespace bqc;

// Schwarz bins per pair list of one task (a sanity limit: the bin offsets of
// the row and wave kernels live in the device work list)
#define BQC_MAX_BINS 4096
// Rectangles per launch of the generic kernels (their arguments travel by
// value; a task with more non-empty rectangles takes several launches)
#define BQC_ARG_BINS 32
#define BQC_THREADS 256
// Largest Cartesian AO count: J/K/D element offsets are 32-bit in the kernels
#define BQC_MAX_NAO 46340

struct TaskArgs {
    int bas_p0, bas_q0;    // first global pair index of the p / q side
    int prim_p0, prim_q0;  // first primitive pair of the p / q side
    int nprim_p, nprim_q;  // primitive pairs per shell pair
    int diag;              // cp_ij == cp_kl
    int swap;              // p side is the task's kl list
    int nbins;
    int klb[BQC_ARG_BINS + 1];
    int ij0[BQC_ARG_BINS];
    int nij[BQC_ARG_BINS];
    long long pre[BQC_ARG_BINS + 1];
};

struct JKArgs {
    double *vj;
    double *vk;
    const double *dm;
    int nao;
    int n_dm;
};

struct PairData {
    const double *a12, *e12, *x12, *y12, *z12;
    const double *bas_x, *bas_y, *bas_z;
    const int *bra, *ket, *ao_loc;
};

// Everything a work item needs about its task, in device memory: the pair
// data of the task's plan, the list orders of both sides, the density and
// J/K matrices and the Schwarz bins.  Items of many tasks -- of one molecule
// or of many -- can then share one launch.
struct TaskDesc {
    PairData pd;
    const int *perm_p, *perm_q;  // list position -> plan position
    // per Q list position {plan position, first shell, second shell,
    // AO offset of the first | AO offset of the second << 16}, filled on
    // the device by fill_qrec before the J/K launches of a build
    const int *qrec;  // 4 ints per position
    const double *dm;
    double *vj, *vk;
    int nao, n_dm;
    int np;                      // P pairs in bins
    int bas_p0, bas_q0, prim_p0, prim_q0, nprim_p, nprim_q;
    int diag, nbins;
    const int *pbins, *qbins;    // nbins + 1 bin offsets of each side
};

__device__ __forceinline__ void atomic_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Address-space views of device buffers.  Plan data, the density and the
// pair lists are not written while a J/K kernel runs: reads with a
// wave-uniform index go through the constant space (scalar loads), per-lane
// reads through the global space (global_load rather than flat loads, which
// would also hold up LDS and scalar waits).  J/K accumulators are updated
// with global atomics.
template <typename T> using cptr = const __attribute__((address_space(4))) T *;
template <typename T> using gptr = const __attribute__((address_space(1))) T *;
typedef __attribute__((address_space(1))) double gdouble;
template <typename T> __device__ __forceinline__ cptr<T> cst(const T *p) { return (cptr<T>)p; }
template <typename T> __device__ __forceinline__ gptr<T> glb(const T *p) { return (gptr<T>)p; }
__device__ __forceinline__ gdouble *glb_rw(double *p) { return (gdouble *)p; }
__device__ __forceinline__ void atomic_add(gdouble *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Integral block of (p | q): p = shells (a, b) of the "bra" pair, q = (c, d).
// T is the arithmetic type of the recurrence and of the contracted block;
// roots, weights, prefactors and recurrence coefficients are always FP64.
// XN consecutive Cartesian components of shell a starting at x0 are
// evaluated (XN = NFI: the whole block); gout holds XN*NFJ*NFK*NFL values.
// rys_tab: nullptr (table in global memory) or a pointer to the table of
// this class's root count in LDS.
template <int LA, int LB, int LC, int LD, typename T, int XN = Cls<LA, LB, LC, LD>::NFI,
          typename RysPtr = std::nullptr_t>
__device__ __forceinline__ void eri_block(T *__restrict__ gout, const PairData &pd,
                                          int pp0, int np_p, int qq0, int np_q, double norm,
                                          const double *ra, const double *rb,
                                          const double *rc, const double *rd, int x0 = 0,
                                          RysPtr rys_tab = nullptr,
                                          const double *p_first = nullptr) {
    using C = Cls<LA, LB, LC, LD>;
    constexpr int NR = C::NR;
    constexpr int NG = XN * C::NFJ * C::NFK * C::NFL;
#pragma unroll C::UN
    for (int n = 0; n < NG; ++n) gout[n] = (T)0;
    const double ab[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
    const double cd[3] = {rc[0] - rd[0], rc[1] - rd[1], rc[2] - rd[2]};
    // P primitives: per-lane reads; Q primitives: the same for all lanes of a
    // row-kernel wavefront (scalar loads there)
    const gptr<double> pa12 = glb(pd.a12), pe12 = glb(pd.e12), px12 = glb(pd.x12),
                       py12 = glb(pd.y12), pz12 = glb(pd.z12);
    const cptr<double> qa12 = cst(pd.a12), qe12 = cst(pd.e12), qx12 = cst(pd.x12),
                       qy12 = cst(pd.y12), qz12 = cst(pd.z12);
    for (int ip = 0; ip < np_p; ++ip) {
        const int p = pp0 + ip;
        // p_first: the first P primitive pair (a12, e12, x12, y12, z12) kept
        // in registers by the caller
        const bool pre = p_first != nullptr && ip == 0;
        const double aij = pre ? p_first[0] : pa12[p];
        const double eij = pre ? p_first[1] : pe12[p];
        const double P[3] = {pre ? p_first[2] : px12[p], pre ? p_first[3] : py12[p],
                             pre ? p_first[4] : pz12[p]};
        for (int iq = 0; iq < np_q; ++iq) {
            const int q = qq0 + iq;
            const double akl = qa12[q], ekl = qe12[q];
            const double Q[3] = {qx12[q], qy12[q], qz12[q]};
            const double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
            const double aijkl = aij + akl;
            const double a1 = aij * akl;
            const double a0 = a1 / aijkl;
            const double x = a0 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
            double u[NR], w[NR];
            if constexpr (std::is_same<RysPtr, std::nullptr_t>::value) rys_roots<NR>(x, u, w);
            else rys_roots<NR>(x, u, w, rys_tab);
            const double fac = eij * ekl / (sqrt(aijkl) * a1);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double u2 = a0 * u[r];
                const double tmp4 = 0.5 / (u2 * aijkl + a1);
                const double b00 = u2 * tmp4;
                const double b10 = b00 + tmp4 * akl;
                const double b01 = b00 + tmp4 * aij;
                const double tmp2 = 2.0 * b00 * akl;
                const double tmp3 = 2.0 * b00 * aij;
                T Ix[C::GSZ], Iy[C::GSZ], Iz[C::GSZ];
                rys2d<LA, LB, LC, LD, T>(Ix, (T)norm, (T)((P[0] - ra[0]) - tmp2 * pq[0]),
                                         (T)((Q[0] - rc[0]) + tmp3 * pq[0]), (T)b10, (T)b01,
                                         (T)b00, (T)ab[0], (T)cd[0]);
                rys2d<LA, LB, LC, LD, T>(Iy, (T)fac, (T)((P[1] - ra[1]) - tmp2 * pq[1]),
                                         (T)((Q[1] - rc[1]) + tmp3 * pq[1]), (T)b10, (T)b01,
                                         (T)b00, (T)ab[1], (T)cd[1]);
                rys2d<LA, LB, LC, LD, T>(Iz, (T)w[r], (T)((P[2] - ra[2]) - tmp2 * pq[2]),
                                         (T)((Q[2] - rc[2]) + tmp3 * pq[2]), (T)b10, (T)b01,
                                         (T)b00, (T)ab[2], (T)cd[2]);
#pragma unroll C::UO
                for (int x = 0; x < XN; ++x) {
                    const int ax = cart_x(LA, x0 + x), ay = cart_y(LA, x0 + x),
                              az = cart_z(LA, x0 + x);
#pragma unroll C::UO
                    for (int y = 0; y < C::NFJ; ++y) {
                        const int bx = cart_x(LB, y), by = cart_y(LB, y), bz = cart_z(LB, y);
#pragma unroll
                        for (int z = 0; z < C::NFK; ++z) {
                            const int cx = cart_x(LC, z), cy = cart_y(LC, z), cz = cart_z(LC, z);
#pragma unroll
                            for (int w = 0; w < C::NFL; ++w) {
                                gout[((x * C::NFJ + y) * C::NFK + z) * C::NFL + w] +=
                                    Ix[C::g(ax, bx, cx, cart_x(LD, w))] *
                                    Iy[C::g(ay, by, cy, cart_y(LD, w))] *
                                    Iz[C::g(az, bz, cz, cart_z(LD, w))];
                            }
                        }
                    }
                }
            }
            // the pair builder sorts primitive pairs by |e12| and zeroes the
            // negligible ones: stop before the next one if it is zero (tested
            // after the work, so single-primitive pairs never wait on it).
            // Not in the classes above 81 components, whose registers are
            // full: there the test added spills ((dp|ds) 240 -> 268 ms).
            if constexpr (C::NF <= 81)
                if (iq + 1 < np_q && qe12[q + 1] == 0.0) break;
        }
        if constexpr (C::NF <= 81)
            if (ip + 1 < np_p && pe12[p + 1] == 0.0) break;
    }
}

// J/K digestion of one integral block (a b | c d) with AO offsets a0..d0.
// Digests the XN components of shell a starting at x0 (see eri_block).
template <int LA, int LB, int LC, int LD, typename T, int XN = Cls<LA, LB, LC, LD>::NFI>
__device__ __forceinline__ void digest(const T *__restrict__ gout, const JKArgs &jk,
                                       int a0, int b0, int c0, int d0, int x0 = 0) {
    using C = Cls<LA, LB, LC, LD>;
    a0 += x0;
    const int nao = jk.nao;
    const size_t nao2 = (size_t)nao * nao;
    for (int idm = 0; idm < jk.n_dm; ++idm) {
        const gptr<double> D = glb(jk.dm + idm * nao2);
        if (jk.vj) {
            gdouble *V = glb_rw(jk.vj + idm * nao2);
#pragma unroll C::UO
            for (int ia = 0; ia < XN; ++ia)
#pragma unroll
                for (int ib = 0; ib < C::NFJ; ++ib) {
                    double s = 0.0;
#pragma unroll
                    for (int ic = 0; ic < C::NFK; ++ic)
#pragma unroll
                        for (int id = 0; id < C::NFL; ++id)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(c0 + ic) * nao + d0 + id];
                    atomic_add(V + (size_t)(a0 + ia) * nao + b0 + ib, s);
                }
#pragma unroll C::UO
            for (int ic = 0; ic < C::NFK; ++ic)
#pragma unroll
                for (int id = 0; id < C::NFL; ++id) {
                    double s = 0.0;
#pragma unroll C::UO
                    for (int ia = 0; ia < XN; ++ia)
#pragma unroll
                        for (int ib = 0; ib < C::NFJ; ++ib)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(a0 + ia) * nao + b0 + ib];
                    atomic_add(V + (size_t)(c0 + ic) * nao + d0 + id, s);
                }
        }
        if (jk.vk) {
            gdouble *V = glb_rw(jk.vk + idm * nao2);
            // V[a,c] += g D[b,d]; V[a,d] += g D[b,c]
#pragma unroll C::UO
            for (int ia = 0; ia < XN; ++ia) {
#pragma unroll
                for (int ic = 0; ic < C::NFK; ++ic) {
                    double s = 0.0;
#pragma unroll
                    for (int ib = 0; ib < C::NFJ; ++ib)
#pragma unroll
                        for (int id = 0; id < C::NFL; ++id)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(b0 + ib) * nao + d0 + id];
                    atomic_add(V + (size_t)(a0 + ia) * nao + c0 + ic, s);
                }
#pragma unroll
                for (int id = 0; id < C::NFL; ++id) {
                    double s = 0.0;
#pragma unroll
                    for (int ib = 0; ib < C::NFJ; ++ib)
#pragma unroll
                        for (int ic = 0; ic < C::NFK; ++ic)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(b0 + ib) * nao + c0 + ic];
                    atomic_add(V + (size_t)(a0 + ia) * nao + d0 + id, s);
                }
            }
            // V[b,c] += g D[a,d]; V[b,d] += g D[a,c]
#pragma unroll C::UO
            for (int ib = 0; ib < C::NFJ; ++ib) {
#pragma unroll
                for (int ic = 0; ic < C::NFK; ++ic) {
                    double s = 0.0;
#pragma unroll C::UO
                    for (int ia = 0; ia < XN; ++ia)
#pragma unroll
                        for (int id = 0; id < C::NFL; ++id)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(a0 + ia) * nao + d0 + id];
                    atomic_add(V + (size_t)(b0 + ib) * nao + c0 + ic, s);
                }
#pragma unroll
                for (int id = 0; id < C::NFL; ++id) {
                    double s = 0.0;
#pragma unroll C::UO
                    for (int ia = 0; ia < XN; ++ia)
#pragma unroll
                        for (int ic = 0; ic < C::NFK; ++ic)
                            s += (double)gout[((ia * C::NFJ + ib) * C::NFK + ic) * C::NFL + id] *
                                 D[(a0 + ia) * nao + c0 + ic];
                    atomic_add(V + (size_t)(b0 + ib) * nao + d0 + id, s);
                }
            }
        }
    }
}

template <int LA, int LB, int LC, int LD, typename T>
__global__ void __launch_bounds__(BQC_THREADS)
    __attribute__((amdgpu_waves_per_eu(1, Cls<LA, LB, LC, LD>::BIG ? 1 : 8)))
    jk_quartets(JKArgs jk, PairData pd, TaskArgs ta) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ta.pre[ta.nbins]) return;
    int b = 0;
    while (t >= ta.pre[b + 1]) ++b;
    const long long local = t - ta.pre[b];
    const int kl = ta.klb[b] + (int)(local / ta.nij[b]);
    const int ij = ta.ij0[b] + (int)(local % ta.nij[b]);
    // p side = bra pair as evaluated; the task's (ij, kl) roles swap when the
    // ket class is the larger one
    const int bas_ij = (ta.swap ? ta.bas_q0 : ta.bas_p0) + ij;
    const int bas_kl = (ta.swap ? ta.bas_p0 : ta.bas_q0) + kl;
    if (ta.diag && bas_ij < bas_kl) return;
    double norm = TWO_PI_52;
    if (bas_ij == bas_kl) norm *= 0.5;
    const int bp = ta.swap ? bas_kl : bas_ij;
    const int bq = ta.swap ? bas_ij : bas_kl;
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
        // one Cartesian component of shell a at a time keeps the block small
        for (int x0 = 0; x0 < C::NFI; ++x0) {
            T gout[C::NF / C::NFI];
            eri_block<LA, LB, LC, LD, T, 1>(gout, pd, pp0, ta.nprim_p, qq0, ta.nprim_q, norm, ra,
                                            rb, rc, rd, x0);
            digest<LA, LB, LC, LD, T, 1>(gout, jk, ao[sa], ao[sb], ao[sc], ao[sd], x0);
        }
    } else {
        T gout[C::NF];
        eri_block<LA, LB, LC, LD, T>(gout, pd, pp0, ta.nprim_p, qq0, ta.nprim_q, norm, ra, rb, rc,
                                     rd);
        digest<LA, LB, LC, LD, T>(gout, jk, ao[sa], ao[sb], ao[sc], ao[sd]);
    }
}


typedef void (*JKKernel)(JKArgs, PairData, TaskArgs);
#define BQC_LMAX 3      // classes with compiled-in shapes (row, wave, quartet kernels)
#define BQC_LMAX_GEN 5  // g and h shells: the run-time-class kernel (jk_general.hip)
#define BQC_NPAIRID ((BQC_LMAX + 1) * (BQC_LMAX + 2) / 2)
// [precision: 0 FP64, 1 FP32][bra class][ket class]
typedef JKKernel JKKernelTable[2][BQC_NPAIRID][BQC_NPAIRID];

template <int LA, int LB, int LC, int LD>
static void reg_kernel(JKKernelTable &tab) {
    tab[0][LA * (LA + 1) / 2 + LB][LC * (LC + 1) / 2 + LD] = jk_quartets<LA, LB, LC, LD, double>;
    tab[1][LA * (LA + 1) / 2 + LB][LC * (LC + 1) / 2 + LD] = jk_quartets<LA, LB, LC, LD, float>;
}

