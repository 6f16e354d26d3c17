// Schwarz bounds q_ij = sqrt(max |(ij|ij)|) on the GPU.
//
// The reference obtains them on the CPU through PySCF's libcvhf
// (CVHFnr_int2e_q_cond, called from cuobc/scf/hf.py:323-324,341): for every
// shell pair the largest diagonal element (ij|ij) over the Cartesian
// components, square-rooted, 1e-100 when it vanishes.  Here one lane handles
// one shell pair of a (group, group) block, builds the primitive-pair data on
// the fly and evaluates only the diagonal components of the (ij|ij) block.
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
#define PTR_COORD 1
#define ATM_SLOTS 6

__device__ __forceinline__ double fac_sp(int l) {
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

template <int LI, int LJ>
__global__ void __launch_bounds__(256)
    q_cond_kernel(double *q, const int *atm, const int *bas, const double *env, int nbas,
                  int p0, int p1, int q0, int q1, int diag) {
    using C = Cls<LI, LJ, LI, LJ>;
    constexpr int NR = C::NR;
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
    const double fsp = fac_sp(LI) * fac_sp(LJ);
    constexpr int NDIAG = C::NFI * C::NFJ;
    double diagv[NDIAG];
#pragma unroll
    for (int n = 0; n < NDIAG; ++n) diagv[n] = 0.0;
    for (int p = 0; p < npi * npj; ++p) {
        const double a_i = ai[p / npj], a_j = aj[p % npj];
        const double aij = a_i + a_j;
        const double eij = fsp * ci[p / npj] * cj[p % npj] * exp(-rr * a_i * a_j / aij);
        const double P[3] = {(a_i * ri[0] + a_j * rj[0]) / aij, (a_i * ri[1] + a_j * rj[1]) / aij,
                             (a_i * ri[2] + a_j * rj[2]) / aij};
        for (int s = 0; s < npi * npj; ++s) {
            const double a_k = ai[s / npj], a_l = aj[s % npj];
            const double akl = a_k + a_l;
            const double ekl = fsp * ci[s / npj] * cj[s % npj] * exp(-rr * a_k * a_l / akl);
            const double Q[3] = {(a_k * ri[0] + a_l * rj[0]) / akl,
                                 (a_k * ri[1] + a_l * rj[1]) / akl,
                                 (a_k * ri[2] + a_l * rj[2]) / akl};
            const double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
            const double aijkl = aij + akl, a1 = aij * akl, a0 = a1 / aijkl;
            const double x = a0 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
            double u[NR], w[NR];
            rys_roots<NR>(x, u, w);
            const double fac = eij * ekl / (sqrt(aijkl) * a1);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double u2 = a0 * u[r];
                const double tmp4 = 0.5 / (u2 * aijkl + a1);
                const double b00 = u2 * tmp4;
                const double b10 = b00 + tmp4 * akl;
                const double b01 = b00 + tmp4 * aij;
                const double tmp2 = 2.0 * b00 * akl, tmp3 = 2.0 * b00 * aij;
                double Ix[C::GSZ], Iy[C::GSZ], Iz[C::GSZ];
                rys2d<LI, LJ, LI, LJ, double>(Ix, TWO_PI_52, (P[0] - ri[0]) - tmp2 * pq[0],
                                              (Q[0] - ri[0]) + tmp3 * pq[0], b10, b01, b00,
                                              ab[0], ab[0]);
                rys2d<LI, LJ, LI, LJ, double>(Iy, fac, (P[1] - ri[1]) - tmp2 * pq[1],
                                              (Q[1] - ri[1]) + tmp3 * pq[1], b10, b01, b00,
                                              ab[1], ab[1]);
                rys2d<LI, LJ, LI, LJ, double>(Iz, w[r], (P[2] - ri[2]) - tmp2 * pq[2],
                                              (Q[2] - ri[2]) + tmp3 * pq[2], b10, b01, b00,
                                              ab[2], ab[2]);
#pragma unroll
                for (int n = 0; n < NDIAG; ++n) {
                    const int c1 = n / C::NFJ, c2 = n % C::NFJ;
                    diagv[n] += Ix[C::g(cart_x(LI, c1), cart_x(LJ, c2), cart_x(LI, c1),
                                        cart_x(LJ, c2))] *
                                Iy[C::g(cart_y(LI, c1), cart_y(LJ, c2), cart_y(LI, c1),
                                        cart_y(LJ, c2))] *
                                Iz[C::g(cart_z(LI, c1), cart_z(LJ, c2), cart_z(LI, c1),
                                        cart_z(LJ, c2))];
                }
            }
        }
    }
    double m = 0.0;
#pragma unroll
    for (int n = 0; n < NDIAG; ++n) m = fmax(m, fabs(diagv[n]));
    const double qv = m > 0.0 ? sqrt(m) : 1e-100;
    q[(size_t)ish * nbas + jsh] = qv;
    q[(size_t)jsh * nbas + ish] = qv;
}

typedef void (*QKernel)(double *, const int *, const int *, const double *, int, int, int,
                        int, int, int);

static QKernel q_kernel(int li, int lj) {
    switch (li * 4 + lj) {
        case 0: return q_cond_kernel<0, 0>;
        case 4: return q_cond_kernel<1, 0>;
        case 5: return q_cond_kernel<1, 1>;
        case 8: return q_cond_kernel<2, 0>;
        case 9: return q_cond_kernel<2, 1>;
        case 10: return q_cond_kernel<2, 2>;
        case 12: return q_cond_kernel<3, 0>;
        case 13: return q_cond_kernel<3, 1>;
        case 14: return q_cond_kernel<3, 2>;
        case 15: return q_cond_kernel<3, 3>;
    }
    return nullptr;
}

// Pairs with a g or h shell (jk_general.hip).
int bqc_q_cond_general(double *q, const int *atm, const int *bas, const double *env, int nbas,
                       int p0, int p1, int q0, int q1, int diag, hipStream_t stream);

extern "C" int BQCq_cond(double *q, const int *atm, const int *bas, const double *env,
                         int nbas, const int *grp_offsets, int ngroups, const int *grp_l,
                         hipStream_t stream) {
    for (int g = 0; g < ngroups; ++g) {
        for (int h = 0; h <= g; ++h) {
            int p0 = grp_offsets[g], p1 = grp_offsets[g + 1];
            int q0 = grp_offsets[h], q1 = grp_offsets[h + 1];
            long long n = (long long)(p1 - p0) * (q1 - q0);
            if (n == 0) continue;
            if (grp_l[g] > 3 || grp_l[h] > 3) {
                if (grp_l[g] > 5 || grp_l[h] > 5 ||
                    bqc_q_cond_general(q, atm, bas, env, nbas, p0, p1, q0, q1, g == h, stream))
                    return 1;
                continue;
            }
            QKernel k = q_kernel(grp_l[g], grp_l[h]);
            if (!k) return 1;
            hipLaunchKernelGGL(k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, q,
                               atm, bas, env, nbas, p0, p1, q0, q1, g == h);
            if (hipGetLastError() != hipSuccess) return 1;
        }
    }
    return 0;
}
