// J/K and Schwarz bounds for classes with a g or h shell (l = 4, 5).
//
// The reference evaluates l <= 5 on the GPU (cuobc/scf/hf.py:47; kernels
// chosen by root count in cuobc/lib/gvhf/nr_jk_driver.cu:181-212) with
// per-thread 2-D tables.  Those tables do not fit a lane here ((hh|hh) has
// 194481 Cartesian components and 11 roots), so these classes take one
// 256-lane workgroup per contracted quartet with run-time angular momenta:
//   1. lanes r < NR evaluate Rys root r and its weight (FP64);
//   2. lanes t < 3 NR build the vertical-recurrence table
//      G[n][m] (n <= li+lj, m <= lk+ll) of direction t % 3 and root t / 3 in
//      LDS;
//   3. the lanes split the (a, b, c) component triples of shells i, j, k;
//      for each triple and root they apply the two horizontal transfers
//      (A-B to the bra, C-D to the ket) to the few G entries the triple
//      needs and sum the block elements (ab|cd) over roots for every
//      component d of shell l, in registers;
//   4. each element is digested into six LDS accumulators
//      J(a,b) += g D(c,d)   J(c,d) += g D(a,b)
//      K(a,c) += g D(b,d)   K(a,d) += g D(b,c)   K(b,c) += g D(a,d)
//      K(b,d) += g D(a,c)
//      (register sums over d for the targets that do not depend on d, LDS
//      atomics for the others), and added to J/K with global atomics once
//      per contracted quartet.
// Quartet weights and the digestion targets are those of the other J/K
// kernels (cuobc/lib/gvhf/g2e.cu:1022-1028, contract_jk.cu:28-316).
#include <hip/hip_runtime.h>
#include <math.h>

#include "jk_kernel.h"
#include "rys_table_hi.inc"

#define GEN_LMAX 5
#define GEN_NF (((GEN_LMAX) + 1) * ((GEN_LMAX) + 2) / 2)  // 21 components
#define GEN_NRMAX 11
#define GEN_G ((2 * GEN_LMAX + 1) * (2 * GEN_LMAX + 1))  // VRR table of one (dir, root)
#define GEN_THREADS 256

static_assert(BQC_RYS_HI_NMAX == GEN_NRMAX, "Rys table must cover 11 roots");

// Root r of the n-point Rys rule at x, 1 <= n <= 11 (run-time n): the
// table of rys_table.inc for n <= 7, rys_table_hi.inc above.
template <int N>
__device__ __forceinline__ void rys_root_hi(double x, int r, double &u, double &w) {
    if (x >= (double)rys_xbig(N)) {
        const double h = BQC_RYS_HI_HERM_S[(N - 1) * BQC_RYS_HI_NMAX + r];
        u = h / (x - h);
        w = BQC_RYS_HI_HERM_W[(N - 1) * BQC_RYS_HI_NMAX + r] * rsqrt(x);
        return;
    }
    const int it = (int)x;
    const double y = 2.0 * (x - (double)it) - 1.0;
    const double y2 = 2.0 * y;
    const double *c = BQC_RYS_HI_COEF + BQC_RYS_HI_OFFSET[N] + it * (2 * N * (BQC_RYS_DEG + 1));
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

__device__ __forceinline__ void rys_root_rt(int n, double x, int r, double &u, double &w) {
    switch (n) {
        case 1: rys_root_one<1>(x, r, u, w); break;
        case 2: rys_root_one<2>(x, r, u, w); break;
        case 3: rys_root_one<3>(x, r, u, w); break;
        case 4: rys_root_one<4>(x, r, u, w); break;
        case 5: rys_root_one<5>(x, r, u, w); break;
        case 6: rys_root_one<6>(x, r, u, w); break;
        case 7: rys_root_one<7>(x, r, u, w); break;
        case 8: rys_root_hi<8>(x, r, u, w); break;
        case 9: rys_root_hi<9>(x, r, u, w); break;
        case 10: rys_root_hi<10>(x, r, u, w); break;
        default: rys_root_hi<11>(x, r, u, w); break;
    }
}

// Cartesian exponent of component c of shell l in direction d (run time).
__device__ __forceinline__ int cart_e(int l, int c, int d) {
    // components: lx = l..0, then ly = l-lx..0; c = (l-lx)(l-lx+1)/2 + (l-lx-ly)
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= c) ++s;  // s = l - lx
    const int lz = c - s * (s + 1) / 2;
    const int lx = l - s, ly = s - lz;
    return d == 0 ? lx : (d == 1 ? ly : lz);
}

// Binomial coefficients C(n, k), n <= 10.
static __device__ const double kBinom[11][11] = {
    {1},
    {1, 1},
    {1, 2, 1},
    {1, 3, 3, 1},
    {1, 4, 6, 4, 1},
    {1, 5, 10, 10, 5, 1},
    {1, 6, 15, 20, 15, 6, 1},
    {1, 7, 21, 35, 35, 21, 7, 1},
    {1, 8, 28, 56, 70, 56, 28, 8, 1},
    {1, 9, 36, 84, 126, 126, 84, 36, 9, 1},
    {1, 10, 45, 120, 210, 252, 210, 120, 45, 10, 1}};

// Vertical recurrence of one direction and root into G (row n, column m,
// row stride MM = mmax + 1).
__device__ void gen_vrr(double *G, int nmax, int mmax, double seed, double c00, double c0p,
                        double b10, double b01, double b00) {
    const int MM = mmax + 1;
    G[0] = seed;
    for (int n = 0; n < nmax; ++n) {
        double v = c00 * G[n * MM];
        if (n > 0) v += n * b10 * G[(n - 1) * MM];
        G[(n + 1) * MM] = v;
    }
    for (int m = 0; m < mmax; ++m)
        for (int n = 0; n <= nmax; ++n) {
            double v = c0p * G[n * MM + m];
            if (m > 0) v += m * b01 * G[n * MM + m - 1];
            if (n > 0) v += n * b00 * G[(n - 1) * MM + m];
            G[n * MM + m + 1] = v;
        }
}

// Bra transfer of one vertical-table entry:
// I(i, j, k') = sum_a C(j, a) ab^(j-a) G[i+a][k'].
__device__ __forceinline__ double gen_bra(const double *G, int MM, int i, int j, int kp,
                                          double ab) {
    double s = 0.0, p = 1.0;
    for (int a = j; a >= 0; --a) {
        s += kBinom[j][a] * p * G[(i + a) * MM + kp];
        p *= ab;
    }
    return s;
}

// V[l] = I(i, j, k, l) for l = 0..LL of one direction and root: the bra
// transfer for k' = k..k+LL, then the ket transfer
// I(i, j, k, l) = sum_b C(l, b) cd^(l-b) I(i, j, k+b).
template <int LL>
__device__ __forceinline__ void gen_hrr(const double *G, int MM, int i, int j, int k, double ab,
                                        double cd, double (&V)[LL + 1]) {
    double W[LL + 1];
#pragma unroll
    for (int b = 0; b <= LL; ++b) W[b] = gen_bra(G, MM, i, j, k + b, ab);
#pragma unroll
    for (int l = 0; l <= LL; ++l) {
        double s = 0.0, p = 1.0;
#pragma unroll
        for (int b = l; b >= 0; --b) {
            s += kBinom[l][b] * p * W[b];
            p *= cd;
        }
        V[l] = s;
    }
}

// One element I(i, j, k, l) of one direction and root (both transfers).
__device__ __forceinline__ double gen_elem(const double *G, int MM, int i, int j, int k, int l,
                                           double ab, double cd) {
    double s = 0.0, p = 1.0;
    for (int b = l; b >= 0; --b) {
        s += kBinom[l][b] * p * gen_bra(G, MM, i, j, k + b, ab);
        p *= cd;
    }
    return s;
}

__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct GenClass {
    int li, lj, lk, ll, nr;
};

// One contracted quartet per workgroup; quartet t of the task's rectangles
// (as in jk_quartets, with the ij list on the bra side).  LL = l of the
// fourth shell (compile time, so a triple's elements stay in registers).
template <int LL>
__global__ void __launch_bounds__(GEN_THREADS)
    jk_general(JKArgs jk, PairData pd, TaskArgs ta, GenClass cl) {
    constexpr int ND = (LL + 1) * (LL + 2) / 2;
    __shared__ double sG[3 * GEN_NRMAX * GEN_G];
    __shared__ double sD[6 * GEN_NF * GEN_NF];
    __shared__ double sV[6 * GEN_NF * GEN_NF];
    __shared__ double sU[GEN_NRMAX], sW[GEN_NRMAX];
    const long long t = blockIdx.x;
    if (t >= ta.pre[ta.nbins]) return;
    int b = 0;
    while (t >= ta.pre[b + 1]) ++b;
    const long long local = t - ta.pre[b];
    const int kl = ta.klb[b] + (int)(local / ta.nij[b]);
    const int ij = ta.ij0[b] + (int)(local % ta.nij[b]);
    const int bas_ij = ta.bas_p0 + ij, bas_kl = ta.bas_q0 + kl;
    if (ta.diag && bas_ij < bas_kl) return;
    double norm = TWO_PI_52;
    if (bas_ij == bas_kl) norm *= 0.5;
    const int tid = threadIdx.x;
    const int li = cl.li, lj = cl.lj, lk = cl.lk, nr = cl.nr;
    const int NA = ncart(li), NB = ncart(lj), NC = ncart(lk);
    const int nmax = li + lj, mmax = lk + LL, MM = mmax + 1;
    const int sa = pd.bra[bas_ij], sb = pd.ket[bas_ij], sc = pd.bra[bas_kl], sd = pd.ket[bas_kl];
    const double ra[3] = {pd.bas_x[sa], pd.bas_y[sa], pd.bas_z[sa]};
    const double rb[3] = {pd.bas_x[sb], pd.bas_y[sb], pd.bas_z[sb]};
    const double rc[3] = {pd.bas_x[sc], pd.bas_y[sc], pd.bas_z[sc]};
    const double rd[3] = {pd.bas_x[sd], pd.bas_y[sd], pd.bas_z[sd]};
    const double ab[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
    const double cd[3] = {rc[0] - rd[0], rc[1] - rd[1], rc[2] - rd[2]};
    const int a0 = pd.ao_loc[sa], b0 = pd.ao_loc[sb], c0 = pd.ao_loc[sc], d0 = pd.ao_loc[sd];
    const int pp0 = ta.prim_p0 + ij * ta.nprim_p, qq0 = ta.prim_q0 + kl * ta.nprim_q;
    // LDS blocks, row-major by the first index: D_ab D_cd D_bd D_ad D_bc D_ac
    // and the accumulators V_ab V_cd V_ac V_ad V_bc V_bd
    const int NN = GEN_NF * GEN_NF;
    double *Dab = sD, *Dcd = sD + NN, *Dbd = sD + 2 * NN, *Dad = sD + 3 * NN,
           *Dbc = sD + 4 * NN, *Dac = sD + 5 * NN;
    double *Vab = sV, *Vcd = sV + NN, *Vac = sV + 2 * NN, *Vad = sV + 3 * NN, *Vbc = sV + 4 * NN,
           *Vbd = sV + 5 * NN;
    const int nao = jk.nao;
    const size_t nao2 = (size_t)nao * nao;
    const int ntrip = NA * NB * NC;
    for (int idm = 0; idm < jk.n_dm; ++idm) {
        const double *D = jk.dm + idm * nao2;
        for (int e = tid; e < NN; e += GEN_THREADS) {
            const int x = e / GEN_NF, y = e % GEN_NF;
            Dab[e] = x < NA && y < NB ? D[(size_t)(a0 + x) * nao + b0 + y] : 0.0;
            Dcd[e] = x < NC && y < ND ? D[(size_t)(c0 + x) * nao + d0 + y] : 0.0;
            Dbd[e] = x < NB && y < ND ? D[(size_t)(b0 + x) * nao + d0 + y] : 0.0;
            Dad[e] = x < NA && y < ND ? D[(size_t)(a0 + x) * nao + d0 + y] : 0.0;
            Dbc[e] = x < NB && y < NC ? D[(size_t)(b0 + x) * nao + c0 + y] : 0.0;
            Dac[e] = x < NA && y < NC ? D[(size_t)(a0 + x) * nao + c0 + y] : 0.0;
        }
        for (int e = tid; e < 6 * NN; e += GEN_THREADS) sV[e] = 0.0;
        __syncthreads();
        for (int ip = 0; ip < ta.nprim_p; ++ip) {
            const int p = pp0 + ip;
            const double aij = pd.a12[p], eij = pd.e12[p];
            const double P[3] = {pd.x12[p], pd.y12[p], pd.z12[p]};
            for (int iq = 0; iq < ta.nprim_q; ++iq) {
                const int q = qq0 + iq;
                const double akl = pd.a12[q], ekl = pd.e12[q];
                const double Q[3] = {pd.x12[q], pd.y12[q], pd.z12[q]};
                const double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                const double aijkl = aij + akl, a1 = aij * akl, a00 = a1 / aijkl;
                const double x = a00 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
                if (tid < nr) {
                    double u, w;
                    rys_root_rt(nr, x, tid, u, w);
                    sU[tid] = u;
                    sW[tid] = w;
                }
                __syncthreads();
                if (tid < 3 * nr) {
                    const int d = tid % 3, r = tid / 3;
                    const double u2 = a00 * sU[r];
                    const double tmp4 = 0.5 / (u2 * aijkl + a1);
                    const double b00 = u2 * tmp4;
                    const double b10 = b00 + tmp4 * akl;
                    const double b01 = b00 + tmp4 * aij;
                    const double tmp2 = 2.0 * b00 * akl, tmp3 = 2.0 * b00 * aij;
                    const double seed =
                        d == 0 ? norm : (d == 1 ? eij * ekl / (sqrt(aijkl) * a1) : sW[r]);
                    gen_vrr(sG + (r * 3 + d) * GEN_G, nmax, mmax, seed,
                            (P[d] - ra[d]) - tmp2 * pq[d], (Q[d] - rc[d]) + tmp3 * pq[d], b10,
                            b01, b00);
                }
                __syncthreads();
                for (int tr = tid; tr < ntrip; tr += GEN_THREADS) {
                    const int x = tr / (NB * NC), y = (tr / NC) % NB, z = tr % NC;
                    int ei[3], ej[3], ek[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        ei[d] = cart_e(li, x, d);
                        ej[d] = cart_e(lj, y, d);
                        ek[d] = cart_e(lk, z, d);
                    }
                    double g[ND];
#pragma unroll
                    for (int w = 0; w < ND; ++w) g[w] = 0.0;
                    for (int r = 0; r < nr; ++r) {
                        double V[3][LL + 1];
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            gen_hrr<LL>(sG + (r * 3 + d) * GEN_G, MM, ei[d], ej[d], ek[d], ab[d],
                                        cd[d], V[d]);
#pragma unroll
                        for (int w = 0; w < ND; ++w)
                            g[w] += V[0][cart_x(LL, w)] * V[1][cart_y(LL, w)] *
                                    V[2][cart_z(LL, w)];
                    }
                    double jab = 0.0, kac = 0.0, kbc = 0.0;
                    const double dab = Dab[x * GEN_NF + y], dbc = Dbc[y * GEN_NF + z],
                                 dac = Dac[x * GEN_NF + z];
#pragma unroll
                    for (int w = 0; w < ND; ++w) {
                        const double v = g[w];
                        jab += v * Dcd[z * GEN_NF + w];
                        kac += v * Dbd[y * GEN_NF + w];
                        kbc += v * Dad[x * GEN_NF + w];
                        lds_add(Vcd + z * GEN_NF + w, v * dab);
                        lds_add(Vad + x * GEN_NF + w, v * dbc);
                        lds_add(Vbd + y * GEN_NF + w, v * dac);
                    }
                    lds_add(Vab + x * GEN_NF + y, jab);
                    lds_add(Vac + x * GEN_NF + z, kac);
                    lds_add(Vbc + y * GEN_NF + z, kbc);
                }
                __syncthreads();
            }
        }
        // one global atomic per accumulator element
        for (int e = tid; e < NN; e += GEN_THREADS) {
            const int x = e / GEN_NF, y = e % GEN_NF;
            if (jk.vj) {
                double *VJ = jk.vj + idm * nao2;
                if (x < NA && y < NB) atomic_add(VJ + (size_t)(a0 + x) * nao + b0 + y, Vab[e]);
                if (x < NC && y < ND) atomic_add(VJ + (size_t)(c0 + x) * nao + d0 + y, Vcd[e]);
            }
            if (jk.vk) {
                double *VK = jk.vk + idm * nao2;
                if (x < NA && y < NC) atomic_add(VK + (size_t)(a0 + x) * nao + c0 + y, Vac[e]);
                if (x < NA && y < ND) atomic_add(VK + (size_t)(a0 + x) * nao + d0 + y, Vad[e]);
                if (x < NB && y < NC) atomic_add(VK + (size_t)(b0 + x) * nao + c0 + y, Vbc[e]);
                if (x < NB && y < ND) atomic_add(VK + (size_t)(b0 + x) * nao + d0 + y, Vbd[e]);
            }
        }
        __syncthreads();
    }
}

#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define PTR_EXP 5
#define PTR_COEFF 6
#define BAS_SLOTS 8
#define PTR_COORD 1
#define ATM_SLOTS 6

__device__ __forceinline__ double fac_sp_g(int l) {
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

// Schwarz bound of one shell pair (i, j) per workgroup for pairs with a g or
// h shell: q = sqrt(max |(ij|ij)|) over the diagonal Cartesian components.
// Workgroup t takes ish = p0 + t / (q1 - q0), jsh = q0 + t % (q1 - q0)
// (jsh <= ish only in a diagonal block).
__global__ void __launch_bounds__(GEN_THREADS)
    q_cond_general(double *qout, const int *atm, const int *bas, const double *env, int nbas,
                   int p0, int p1, int q0, int q1, int diag) {
    __shared__ double sG[3 * GEN_NRMAX * GEN_G];
    __shared__ double sU[GEN_NRMAX], sW[GEN_NRMAX];
    __shared__ double sMax[GEN_THREADS];
    const int t = blockIdx.x;
    const int nq = q1 - q0;
    if (t >= (p1 - p0) * nq) return;
    const int ish = p0 + t / nq, jsh = q0 + t % nq;
    if (diag && jsh > ish) return;
    const int tid = threadIdx.x;
    const int *bi = bas + ish * BAS_SLOTS, *bj = bas + jsh * BAS_SLOTS;
    const int li = bi[ANG_OF], lj = bj[ANG_OF];
    const int nr = li + lj + 1;
    const double *ri = env + atm[bi[ATOM_OF] * ATM_SLOTS + PTR_COORD];
    const double *rj = env + atm[bj[ATOM_OF] * ATM_SLOTS + PTR_COORD];
    const double *ai = env + bi[PTR_EXP], *aj = env + bj[PTR_EXP];
    const double *ci = env + bi[PTR_COEFF], *cj = env + bj[PTR_COEFF];
    const int npi = bi[NPRIM_OF], npj = bj[NPRIM_OF];
    const double ab[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
    const double rr = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    const double fsp = fac_sp_g(li) * fac_sp_g(lj);
    const int NA = ncart(li), NB = ncart(lj);
    const int nmax = li + lj, MM = nmax + 1;
    // each lane keeps the diagonal elements of its (a, b) components
    constexpr int PER = (GEN_NF * GEN_NF + GEN_THREADS - 1) / GEN_THREADS;
    double acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.0;
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
            const double aijkl = aij + akl, a1 = aij * akl, a00 = a1 / aijkl;
            const double x = a00 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
            if (tid < nr) {
                double u, w;
                rys_root_rt(nr, x, tid, u, w);
                sU[tid] = u;
                sW[tid] = w;
            }
            __syncthreads();
            if (tid < 3 * nr) {
                const int d = tid % 3, r = tid / 3;
                const double u2 = a00 * sU[r];
                const double tmp4 = 0.5 / (u2 * aijkl + a1);
                const double b00 = u2 * tmp4;
                const double b10 = b00 + tmp4 * akl;
                const double b01 = b00 + tmp4 * aij;
                const double tmp2 = 2.0 * b00 * akl, tmp3 = 2.0 * b00 * aij;
                const double seed =
                    d == 0 ? TWO_PI_52 : (d == 1 ? eij * ekl / (sqrt(aijkl) * a1) : sW[r]);
                gen_vrr(sG + (r * 3 + d) * GEN_G, nmax, nmax, seed, (P[d] - ri[d]) - tmp2 * pq[d],
                        (Q[d] - ri[d]) + tmp3 * pq[d], b10, b01, b00);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PER; ++m) {
                const int e = tid + m * GEN_THREADS;
                if (e < NA * NB) {
                    const int x = e / NB, y = e % NB;
                    double v = 0.0;
                    for (int r = 0; r < nr; ++r) {
                        double prod = 1.0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const int ex = cart_e(li, x, d), ey = cart_e(lj, y, d);
                            prod *= gen_elem(sG + (r * 3 + d) * GEN_G, MM, ex, ey, ex, ey, ab[d],
                                             ab[d]);
                        }
                        v += prod;
                    }
                    acc[m] += v;
                }
            }
            __syncthreads();
        }
    }
    double mx = 0.0;
#pragma unroll
    for (int m = 0; m < PER; ++m) mx = fmax(mx, fabs(acc[m]));
    sMax[tid] = mx;
    __syncthreads();
    for (int s = GEN_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) sMax[tid] = fmax(sMax[tid], sMax[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        const double qv = sMax[0] > 0.0 ? sqrt(sMax[0]) : 1e-100;
        qout[(size_t)ish * nbas + jsh] = qv;
        qout[(size_t)jsh * nbas + ish] = qv;
    }
}

// Largest LDS a workgroup of these kernels may use on the current device
// (checked before the first launch).
static bool general_fits() {
    static int ok = -1;
    if (ok < 0) {
        hipFuncAttributes a1, a2;
        int dev = 0, lds_max = 0;
        ok = hipFuncGetAttributes(&a1, (const void *)jk_general<GEN_LMAX>) == hipSuccess &&
             hipFuncGetAttributes(&a2, (const void *)q_cond_general) == hipSuccess &&
             hipGetDevice(&dev) == hipSuccess &&
             hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
                 hipSuccess &&
             a1.sharedSizeBytes <= (size_t)lds_max && a2.sharedSizeBytes <= (size_t)lds_max &&
             a1.localSizeBytes <= 32768 && a2.localSizeBytes <= 32768;
        if (!ok)
            fprintf(stderr, "bqcjk: g/h-shell kernels exceed this device's LDS or scratch "
                            "limits\n");
    }
    return ok == 1;
}

// Launches the quartets of one task's rectangles (ta, bra = ij list) of a
// class with a g or h shell.  Returns 0 ok, 1 error.
int bqc_launch_general(const JKArgs &jk, const PairData &pd, const TaskArgs &ta, int li, int lj,
                       int lk, int ll, hipStream_t stream) {
    if (li > GEN_LMAX || lj > GEN_LMAX || lk > GEN_LMAX || ll > GEN_LMAX) return 1;
    if (!general_fits()) return 1;
    const long long n = ta.pre[ta.nbins];
    if (n <= 0) return 0;
    if (n > 0x7fffffffLL) return 1;
    GenClass cl = {li, lj, lk, ll, (li + lj + lk + ll) / 2 + 1};
    void (*kern)(JKArgs, PairData, TaskArgs, GenClass) = nullptr;
    switch (ll) {
        case 0: kern = jk_general<0>; break;
        case 1: kern = jk_general<1>; break;
        case 2: kern = jk_general<2>; break;
        case 3: kern = jk_general<3>; break;
        case 4: kern = jk_general<4>; break;
        default: kern = jk_general<5>; break;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(GEN_THREADS), 0, stream, jk, pd, ta, cl);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Schwarz bounds of the shell pairs of the (group, group) block
// [p0, p1) x [q0, q1) (jsh <= ish when diag).
int bqc_q_cond_general(double *q, const int *atm, const int *bas, const double *env, int nbas,
                       int p0, int p1, int q0, int q1, int diag, hipStream_t stream) {
    const long long n = (long long)(p1 - p0) * (q1 - q0);
    if (n <= 0) return 0;
    if (n > 0x7fffffffLL || !general_fits()) return 1;
    hipLaunchKernelGGL(q_cond_general, dim3((unsigned)n), dim3(GEN_THREADS), 0, stream, q, atm,
                       bas, env, nbas, p0, p1, q0, q1, diag);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
