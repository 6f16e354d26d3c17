// One-electron integrals on the GPU: overlap S, kinetic T and nuclear
// attraction V over the decontracted Cartesian shells of the J/K plan.
//
// (In the reference these come from PySCF/libcint on the CPU, called by the
// SCF driver around get_jk: cuobc/scf/hf.py:272-274 overrides only get_jk.)
//
// Primitive Cartesian Gaussians x^a y^b z^c exp(-alpha r^2), with the
// libcint angular factors 1/(2 sqrt(pi)) for s and sqrt(3/(4 pi)) for p.
//   overlap, 1-D (Obara-Saika):  S(i+1,j) = PA S(i,j) + (i S(i-1,j) + j S(i,j-1)) / 2p
//                                 S(i,j+1) = S(i+1,j) + AB S(i,j)
//   kinetic, 1-D:  T(i,j) = b(2j+1) S(i,j) - 2 b^2 S(i,j+2) - j(j-1)/2 S(i,j-2)
//   nuclear attraction (Rys, one centre C, roots of x = p |P-C|^2):
//     I(i+1) = (PA - t^2 PC) I(i) + i (1-t^2)/(2p) I(i-1),  then transfer to j;
//     V = -Z (2 pi / p) K_AB sum_r w_r Ix Iy Iz
// One lane evaluates one shell pair (i >= j by group order); blocks are
// written to both (i, j) and (j, i).
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/byteqc_jk.h"
#include "bqc_device.h"

using namespace bqc;

#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define PTR_EXP 5
#define PTR_COEFF 6
#define BAS_SLOTS 8
#define CHARGE_OF 0
#define PTR_COORD 1
#define ATM_SLOTS 6

__device__ __forceinline__ double fac_sp1(int l) {
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

// 1-D overlap table S[i][j], i <= LI, j <= LJ + 2 (for the kinetic operator)
template <int LI, int LJ2>
__device__ __forceinline__ void ovlp1d(double (&S)[LI + 1][LJ2 + 1], double pa, double ab,
                                       double inv2p, double s00) {
    constexpr int N = LI + LJ2;
    double G[N + 1];
    G[0] = s00;
    if (N > 0) G[1] = pa * s00;
#pragma unroll
    for (int n = 1; n < N; ++n) G[n + 1] = pa * G[n] + n * inv2p * G[n - 1];
    // transfer to j: (i, j+1) = (i+1, j) + AB (i, j)
    double W[N + 1];
#pragma unroll
    for (int n = 0; n <= N; ++n) W[n] = G[n];
#pragma unroll
    for (int i = 0; i <= LI; ++i) S[i][0] = W[i];
#pragma unroll
    for (int j = 0; j < LJ2; ++j) {
#pragma unroll
        for (int n = 0; n < N - j; ++n) W[n] = W[n + 1] + ab * W[n];
#pragma unroll
        for (int i = 0; i <= LI; ++i) S[i][j + 1] = W[i];
    }
}

template <int LI, int LJ>
__device__ __forceinline__ void rys1d(double (&I)[LI + 1][LJ + 1], double c, double b,
                                      double ab) {
    constexpr int N = LI + LJ;
    double W[N + 1];
    W[0] = 1.0;
    if (N > 0) W[1] = c;
#pragma unroll
    for (int n = 1; n < N; ++n) W[n + 1] = c * W[n] + n * b * W[n - 1];
#pragma unroll
    for (int i = 0; i <= LI; ++i) I[i][0] = W[i];
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
#pragma unroll
        for (int n = 0; n < N - j; ++n) W[n] = W[n + 1] + ab * W[n];
#pragma unroll
        for (int i = 0; i <= LI; ++i) I[i][j + 1] = W[i];
    }
}

template <int LI, int LJ>
__global__ void __launch_bounds__(128)
    int1e_kernel(double *S, double *T, double *V, const int *atm, int natm, const int *bas,
                 const double *env, const int *ao_loc, int nao, int p0, int p1, int q0, int q1,
                 int diag) {
    constexpr int NFI = ncart(LI), NFJ = ncart(LJ);
    constexpr int NR = (LI + LJ) / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nq = q1 - q0;
    if (t >= (p1 - p0) * nq) return;
    const int ish = p0 + t / nq, jsh = q0 + t % nq;
    if (diag && jsh > ish) return;
    const int *bi = bas + ish * BAS_SLOTS, *bj = bas + jsh * BAS_SLOTS;
    const double *ri = env + atm[bi[ATOM_OF] * ATM_SLOTS + PTR_COORD];
    const double *rj = env + atm[bj[ATOM_OF] * ATM_SLOTS + PTR_COORD];
    const double *ai = env + bi[PTR_EXP], *aj = env + bj[PTR_EXP];
    const double *ci = env + bi[PTR_COEFF], *cj = env + bj[PTR_COEFF];
    const int npi = bi[NPRIM_OF], npj = bj[NPRIM_OF];
    const double ab[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
    const double rr = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    const double fsp = fac_sp1(LI) * fac_sp1(LJ);
    double s[NFI * NFJ], k[NFI * NFJ], v[NFI * NFJ];
#pragma unroll
    for (int n = 0; n < NFI * NFJ; ++n) s[n] = k[n] = v[n] = 0.0;
    for (int ip = 0; ip < npi; ++ip) {
        for (int jp = 0; jp < npj; ++jp) {
            const double a = ai[ip], b = aj[jp];
            const double p = a + b;
            const double kab = exp(-a * b / p * rr);
            const double cc = fsp * ci[ip] * cj[jp] * kab;
            const double P[3] = {(a * ri[0] + b * rj[0]) / p, (a * ri[1] + b * rj[1]) / p,
                                 (a * ri[2] + b * rj[2]) / p};
            const double inv2p = 0.5 / p;
            const double s00 = sqrt(M_PI / p);
            double Sd[3][LI + 1][LJ + 3];
#pragma unroll
            for (int d = 0; d < 3; ++d) ovlp1d<LI, LJ + 2>(Sd[d], P[d] - ri[d], ab[d], inv2p, s00);
            double Td[3][LI + 1][LJ + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int i = 0; i <= LI; ++i)
#pragma unroll
                    for (int j = 0; j <= LJ; ++j) {
                        double tv = b * (2 * j + 1) * Sd[d][i][j] - 2 * b * b * Sd[d][i][j + 2];
                        if (j >= 2) tv -= 0.5 * j * (j - 1) * Sd[d][i][j - 2];
                        Td[d][i][j] = tv;
                    }
#pragma unroll
            for (int x = 0; x < NFI; ++x)
#pragma unroll
                for (int y = 0; y < NFJ; ++y) {
                    const int ix = cart_x(LI, x), iy = cart_y(LI, x), iz = cart_z(LI, x);
                    const int jx = cart_x(LJ, y), jy = cart_y(LJ, y), jz = cart_z(LJ, y);
                    const double sx = Sd[0][ix][jx], sy = Sd[1][iy][jy], sz = Sd[2][iz][jz];
                    s[x * NFJ + y] += cc * sx * sy * sz;
                    k[x * NFJ + y] += cc * (Td[0][ix][jx] * sy * sz + sx * Td[1][iy][jy] * sz +
                                            sx * sy * Td[2][iz][jz]);
                }
            // nuclear attraction
            const double pre = 2.0 * M_PI / p * cc;
            for (int ia = 0; ia < natm; ++ia) {
                const double z = (double)atm[ia * ATM_SLOTS + CHARGE_OF];
                if (z == 0.0) continue;
                const double *rc = env + atm[ia * ATM_SLOTS + PTR_COORD];
                const double pc[3] = {P[0] - rc[0], P[1] - rc[1], P[2] - rc[2]};
                const double x = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
                double u[NR], w[NR];
                rys_roots<NR>(x, u, w);
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const double t2 = u[r] / (1.0 + u[r]);
                    const double bb = (1.0 - t2) * inv2p;
                    double Ix[LI + 1][LJ + 1], Iy[LI + 1][LJ + 1], Iz[LI + 1][LJ + 1];
                    rys1d<LI, LJ>(Ix, P[0] - ri[0] - t2 * pc[0], bb, ab[0]);
                    rys1d<LI, LJ>(Iy, P[1] - ri[1] - t2 * pc[1], bb, ab[1]);
                    rys1d<LI, LJ>(Iz, P[2] - ri[2] - t2 * pc[2], bb, ab[2]);
                    const double f = -z * pre * w[r];
#pragma unroll
                    for (int xx = 0; xx < NFI; ++xx)
#pragma unroll
                        for (int y = 0; y < NFJ; ++y)
                            v[xx * NFJ + y] += f * Ix[cart_x(LI, xx)][cart_x(LJ, y)] *
                                               Iy[cart_y(LI, xx)][cart_y(LJ, y)] *
                                               Iz[cart_z(LI, xx)][cart_z(LJ, y)];
                }
            }
        }
    }
    const int i0 = ao_loc[ish], j0 = ao_loc[jsh];
#pragma unroll
    for (int x = 0; x < NFI; ++x)
#pragma unroll
        for (int y = 0; y < NFJ; ++y) {
            const size_t a = (size_t)(i0 + x) * nao + j0 + y, bt = (size_t)(j0 + y) * nao + i0 + x;
            S[a] = S[bt] = s[x * NFJ + y];
            T[a] = T[bt] = k[x * NFJ + y];
            V[a] = V[bt] = v[x * NFJ + y];
        }
}

typedef void (*Int1eKernel)(double *, double *, double *, const int *, int, const int *,
                            const double *, const int *, int, int, int, int, int, int);

static Int1eKernel int1e_kernel_for(int li, int lj) {
    switch (li * 4 + lj) {
        case 0: return int1e_kernel<0, 0>;
        case 4: return int1e_kernel<1, 0>;
        case 5: return int1e_kernel<1, 1>;
        case 8: return int1e_kernel<2, 0>;
        case 9: return int1e_kernel<2, 1>;
        case 10: return int1e_kernel<2, 2>;
        case 12: return int1e_kernel<3, 0>;
        case 13: return int1e_kernel<3, 1>;
        case 14: return int1e_kernel<3, 2>;
        case 15: return int1e_kernel<3, 3>;
    }
    return nullptr;
}

extern "C" int BQCint1e(double *S, double *T, double *V, const int *atm, int natm,
                        const int *bas, const double *env, const int *ao_loc, int nao,
                        const int *grp_offsets, int ngroups, const int *grp_l,
                        hipStream_t stream) {
    for (int g = 0; g < ngroups; ++g) {
        for (int h = 0; h <= g; ++h) {
            if (grp_l[g] < grp_l[h]) return 1;
            Int1eKernel k = int1e_kernel_for(grp_l[g], grp_l[h]);
            if (!k) return 1;
            int p0 = grp_offsets[g], p1 = grp_offsets[g + 1];
            int q0 = grp_offsets[h], q1 = grp_offsets[h + 1];
            long long n = (long long)(p1 - p0) * (q1 - q0);
            if (n == 0) continue;
            hipLaunchKernelGGL(k, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, S, T, V,
                               atm, natm, bas, env, ao_loc, nao, p0, p1, q0, q1, g == h);
            if (hipGetLastError() != hipSuccess) return 1;
        }
    }
    return 0;
}
