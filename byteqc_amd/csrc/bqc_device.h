// Device building blocks of the Rys-quadrature J/K kernels (gfx950, wave64).
//
//  * Rys nodes/weights: piecewise Chebyshev table (rys_table.inc) below
//    XBIG(n) = 36 + 8n, Gauss-Hermite large-x limit above.  Evaluated in
//    FP64 for every precision mode.
//  * 2-D Rys integrals per Cartesian direction and root, for a compile-time
//    class (LI LJ | LK LL): vertical recurrence in (n = i+j, m = k+l) with the
//    coefficients c00, c0p, b00, b10, b01, then transfer of angular momentum
//    to j (A-B) and to l (C-D).  The recurrence follows the arithmetic of the
//    reference (cuobc/lib/gvhf/g2e.cu:32-490); here every loop has
//    compile-time bounds so the intermediates live in registers.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#define BQC_RYS_QUAL static __device__
#include "rys_table.inc"

namespace bqc {

constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
constexpr int nroots_of(int li, int lj, int lk, int ll) { return (li + lj + lk + ll) / 2 + 1; }
constexpr int rys_xbig(int n) { return 36 + 8 * n; }
constexpr int rys_offset(int n) {
    int off = 0;
    for (int m = 1; m < n; ++m) off += rys_xbig(m) * 2 * m * (BQC_RYS_DEG + 1);
    return off;
}

// Cartesian exponents of component c of a shell with angular momentum l
// (order: lx from l down to 0, then ly from l-lx down to 0).
__host__ __device__ constexpr int cart_x(int l, int c) {
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly, ++n)
            if (n == c) return lx;
    return 0;
}
__host__ __device__ constexpr int cart_y(int l, int c) {
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly, ++n)
            if (n == c) return ly;
    return 0;
}
__host__ __device__ constexpr int cart_z(int l, int c) { return l - cart_x(l, c) - cart_y(l, c); }

// 2*pi^(5/2): prefactor of the Rys integrals (cuobc/lib/gint/g2e.c:41)
constexpr double TWO_PI_52 = 34.98683665524972497;

// Number of doubles of the interpolation table of N roots.
constexpr int rys_table_size(int n) { return rys_xbig(n) * 2 * n * (BQC_RYS_DEG + 1); }

// Nodes u_i = t_i^2 / (1 - t_i^2) and weights of the N-point Rys rule for
// argument x.  tab: the N-root part of the table (a pointer into global
// memory, or into a copy in LDS made by the caller).
template <int N, typename TabPtr>
__device__ __forceinline__ void rys_roots(double x, double *__restrict__ u,
                                          double *__restrict__ w, TabPtr tab) {
    static_assert(N >= 1 && N <= BQC_RYS_NMAX, "unsupported number of roots");
    if (x >= (double)rys_xbig(N)) {
        const double sq = rsqrt(x);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double h = BQC_RYS_HERM_S[(N - 1) * BQC_RYS_NMAX + i];
            u[i] = h / (x - h);
            w[i] = BQC_RYS_HERM_W[(N - 1) * BQC_RYS_NMAX + i] * sq;
        }
        return;
    }
    const int it = (int)x;
    const double y = 2.0 * (x - (double)it) - 1.0;
    const double y2 = 2.0 * y;
    const auto c = tab + it * (2 * N * (BQC_RYS_DEG + 1));
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
        const auto cf = c + f * (BQC_RYS_DEG + 1);
        double b1 = 0.0, b2 = 0.0;
#pragma unroll
        for (int j = BQC_RYS_DEG; j >= 1; --j) {
            const double t = fma(y2, b1, cf[j] - b2);
            b2 = b1;
            b1 = t;
        }
        const double v = fma(y, b1, cf[0] - b2);
        if (f < N) u[f] = v;
        else w[f - N] = v;
    }
}

template <int N>
__device__ __forceinline__ void rys_roots(double x, double *__restrict__ u,
                                          double *__restrict__ w) {
    rys_roots<N>(x, u, w, (const double *)(BQC_RYS_COEF + rys_offset(N)));
}

// Root r (0 <= r < N) and its weight only: the same table and asymptotic
// form as rys_roots, two Chebyshev sums instead of 2N (lanes that need
// different roots of one x run the same instructions).
template <int N>
__device__ __forceinline__ void rys_root_one(double x, int r, double &u, double &w) {
    static_assert(N >= 1 && N <= BQC_RYS_NMAX, "unsupported number of roots");
    if (x >= (double)rys_xbig(N)) {
        const double h = BQC_RYS_HERM_S[(N - 1) * BQC_RYS_NMAX + r];
        u = h / (x - h);
        w = BQC_RYS_HERM_W[(N - 1) * BQC_RYS_NMAX + r] * rsqrt(x);
        return;
    }
    const int it = (int)x;
    const double y = 2.0 * (x - (double)it) - 1.0;
    const double y2 = 2.0 * y;
    const double *c = (const double *)(BQC_RYS_COEF + rys_offset(N)) + it * (2 * N * (BQC_RYS_DEG + 1));
    const double *cu = c + r * (BQC_RYS_DEG + 1), *cw = c + (N + r) * (BQC_RYS_DEG + 1);
    double bu1 = 0.0, bu2 = 0.0, bw1 = 0.0, bw2 = 0.0;
#pragma unroll
    for (int j = BQC_RYS_DEG; j >= 1; --j) {
        const double tu = fma(y2, bu1, cu[j] - bu2);
        const double tw = fma(y2, bw1, cw[j] - bw2);
        bu2 = bu1;
        bu1 = tu;
        bw2 = bw1;
        bw1 = tw;
    }
    u = fma(y, bu1, cu[0] - bu2);
    w = fma(y, bw1, cw[0] - bw2);
}

// Class traits of a quartet (LI LJ | LK LL).
template <int LI, int LJ, int LK, int LL>
struct Cls {
    static constexpr int NR = nroots_of(LI, LJ, LK, LL);
    static constexpr int NFI = ncart(LI), NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
    static constexpr int NF = NFI * NFJ * NFK * NFL;
    static constexpr int NMAX = LI + LJ, MMAX = LK + LL;
    // index of the 2-D integral I(i, j, k, l) of one direction
    static constexpr int S_L = 1, S_K = LL + 1, S_J = (LK + 1) * S_K, S_I = (LJ + 1) * S_J;
    static constexpr int GSZ = (LI + 1) * S_I;
    // classes whose integral block is too large for registers use rolled
    // outer loops (the block then lives in scratch memory)
    static constexpr bool BIG = NF > 160;
    static constexpr int UO = BIG ? 1 : 16;
    static constexpr int UN = BIG ? 8 : 2048;
    __host__ __device__ static constexpr int g(int i, int j, int k, int l) {
        return i * S_I + j * S_J + k * S_K + l;
    }
};

// 2-D Rys integrals of one direction for one root.  T = double or float.
template <int LI, int LJ, int LK, int LL, typename T>
__device__ __forceinline__ void rys2d(T *__restrict__ I, T seed, T c00, T c0p, T b10, T b01,
                                      T b00, T ab, T cd) {
    using C = Cls<LI, LJ, LK, LL>;
    constexpr int NMAX = C::NMAX, MMAX = C::MMAX;
    T G[NMAX + 1][MMAX + 1];
    G[0][0] = seed;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
        T v = c00 * G[n][0];
        if (n > 0) v += (T)n * b10 * G[n - 1][0];
        G[n + 1][0] = v;
    }
#pragma unroll
    for (int m = 0; m < MMAX; ++m) {
#pragma unroll
        for (int n = 0; n <= NMAX; ++n) {
            T v = c0p * G[n][m];
            if (m > 0) v += (T)m * b01 * G[n][m - 1];
            if (n > 0) v += (T)n * b00 * G[n - 1][m];
            G[n][m + 1] = v;
        }
    }
    // bra transfer, in place: after step j, W[n][m] = I(i=n, j, m)
    T H[LI + 1][LJ + 1][MMAX + 1];
#pragma unroll
    for (int i = 0; i <= LI; ++i)
#pragma unroll
        for (int m = 0; m <= MMAX; ++m) H[i][0][m] = G[i][m];
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
#pragma unroll
        for (int n = 0; n <= NMAX - j - 1; ++n)
#pragma unroll
            for (int m = 0; m <= MMAX; ++m) G[n][m] = G[n + 1][m] + ab * G[n][m];
#pragma unroll
        for (int i = 0; i <= LI; ++i)
#pragma unroll
            for (int m = 0; m <= MMAX; ++m) H[i][j + 1][m] = G[i][m];
    }
    // ket transfer for every (i, j)
#pragma unroll
    for (int i = 0; i <= LI; ++i) {
#pragma unroll
        for (int j = 0; j <= LJ; ++j) {
            T K[MMAX + 1];
#pragma unroll
            for (int m = 0; m <= MMAX; ++m) K[m] = H[i][j][m];
#pragma unroll
            for (int k = 0; k <= LK; ++k) I[C::g(i, j, k, 0)] = K[k];
#pragma unroll
            for (int l = 0; l < LL; ++l) {
#pragma unroll
                for (int k = 0; k <= MMAX - l - 1; ++k) K[k] = K[k + 1] + cd * K[k];
#pragma unroll
                for (int k = 0; k <= LK; ++k) I[C::g(i, j, k, l + 1)] = K[k];
            }
        }
    }
}

}  // namespace bqc
