/*
 * CPU FP64 oracle for the direct-SCF J/K path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark load
 * this library, and only as a checker.  The shipped library never links it.
 *
 * It restates, on the CPU and with OpenMP, the arithmetic of the reference's
 * J/K path (paths relative to the reference tree):
 *   pair data  a12, e12, P      cuobc/lib/gint/pair_data.c:27-142
 *   class constants              cuobc/lib/gint/g2e.c:30-97
 *   Rys 2-D recurrence (VRR+HRR) cuobc/lib/gvhf/g2e.cu:32-490
 *   thread -> quartet, weights   cuobc/lib/gvhf/g2e.cu:998-1083
 *   J/K digestion                cuobc/lib/gvhf/contract_jk.cu:28-316
 *   bin loop over quartets       cuobc/lib/gvhf/nr_jk_driver.cu:227-283
 *   Schwarz q_cond               PySCF libcvhf CVHFnr_int2e_q_cond (pyscf 2.5.0,
 *                                not vendored; called from hf.py:323-324,341)
 * Rys roots and weights (reference: cuobc/lib/gint/rys_roots.cu:1314-1872,
 * fitted tables) are produced here by a different route: a discretised
 * Stieltjes procedure in long double (Gauss-Legendre discretisation of the
 * weight exp(-x t^2) on [-1,1]) evaluated at Chebyshev nodes of unit-width
 * intervals, turned into degree-20 Chebyshev interpolants at start-up; above
 * XBIG(n) the large-x limit (Gauss-Hermite nodes over sqrt(x)) is used.
 *
 * Pinning: rys_golden.npz (mpmath, 80 digits), closed-form (ss|ss), an
 * independent Obara-Saika ERI in mpmath, brute-force J/K from the full ERI
 * tensor, and invariants (symmetry, rotation).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

#define ORC_NMAX 11
#define ORC_DEG 20
#define ORC_NQUAD 160
#define ORC_LMAX 5
#define ORC_BLK (NCART(ORC_LMAX) * NCART(ORC_LMAX) * NCART(ORC_LMAX) * NCART(ORC_LMAX))

/* slots of _atm/_bas (cuobc/lib/gint/gint.h:28-46) */
#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define NCTR_OF 3
#define PTR_EXP 5
#define PTR_COEFF 6
#define BAS_SLOTS 8
#define PTR_COORD 1
#define ATM_SLOTS 6

/* ---------------------------------------------------------------------- */
/* Rys roots and weights                                                   */
/* ---------------------------------------------------------------------- */
static int orc_nint[ORC_NMAX + 1];
static double orc_xbig[ORC_NMAX + 1];
/* coefficient tables: [n][interval][2n functions][DEG+1] */
static double *orc_tab[ORC_NMAX + 1];
static double orc_herm_s[ORC_NMAX + 1][ORC_NMAX]; /* h_i^2 */
static double orc_herm_w[ORC_NMAX + 1][ORC_NMAX]; /* W_i */
static int orc_ready = 0;

/* Gauss-Legendre nodes/weights on [-1,1] in long double (Newton on P_N). */
static void gauss_legendre(int n, long double *x, long double *w) {
    for (int i = 0; i < (n + 1) / 2; i++) {
        long double z = cosl(M_PI * (i + 0.75L) / (n + 0.5L));
        long double pp = 0;
        for (int it = 0; it < 100; it++) {
            long double p1 = 1, p2 = 0;
            for (int j = 0; j < n; j++) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-30L) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2 / ((1 - z * z) * pp * pp);
    }
}

/* Symmetric tridiagonal eigenproblem by implicit QL with Wilkinson shifts.
 * d: diagonal (in: m, out: eigenvalues), e: sub-diagonal e[1..m-1],
 * z: first components of eigenvectors (in: e_0, out). */
static void tql_first(int m, long double *d, long double *e, long double *z) {
    for (int i = 1; i < m; i++) e[i - 1] = e[i];
    e[m - 1] = 0;
    for (int l = 0; l < m; l++) {
        int iter = 0, mm;
        do {
            for (mm = l; mm < m - 1; mm++) {
                long double dd = fabsl(d[mm]) + fabsl(d[mm + 1]);
                if (fabsl(e[mm]) + dd == dd) break;
            }
            if (mm != l) {
                if (iter++ == 200) break;
                long double g = (d[l + 1] - d[l]) / (2 * e[l]);
                long double r = hypotl(g, 1.0L);
                g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? fabsl(r) : -fabsl(r)));
                long double s = 1, c = 1, p = 0;
                int i;
                for (i = mm - 1; i >= l; i--) {
                    long double f = s * e[i], b = c * e[i];
                    r = hypotl(f, g);
                    e[i + 1] = r;
                    if (r == 0) {
                        d[i + 1] -= p;
                        e[mm] = 0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    f = z[i + 1];
                    z[i + 1] = s * z[i] + c * f;
                    z[i] = c * z[i] - s * f;
                }
                if (r == 0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[mm] = 0;
            }
        } while (mm != l);
    }
}

static long double gl_x[ORC_NQUAD], gl_w[ORC_NQUAD];

/* 2n-point Gauss rule for the weight exp(-x t^2) on [-1,1]; returns the n
 * positive nodes squared (ascending) and their weights. */
static void rys_stieltjes(int n, long double x, long double *s_out, long double *w_out) {
    int m = 2 * n;
    long double om[ORC_NQUAD], q[ORC_NQUAD], qp[ORC_NQUAD], r[ORC_NQUAD];
    long double alpha[2 * ORC_NMAX], sb[2 * ORC_NMAX + 1];
    long double b0 = 0;
    for (int k = 0; k < ORC_NQUAD; k++) {
        om[k] = gl_w[k] * expl(-x * gl_x[k] * gl_x[k]);
        b0 += om[k];
    }
    for (int k = 0; k < ORC_NQUAD; k++) {
        q[k] = 1 / sqrtl(b0);
        qp[k] = 0;
    }
    sb[0] = 0;
    for (int j = 0; j < m; j++) {
        long double a = 0;
        for (int k = 0; k < ORC_NQUAD; k++) a += om[k] * gl_x[k] * q[k] * q[k];
        alpha[j] = a;
        long double nrm = 0;
        for (int k = 0; k < ORC_NQUAD; k++) {
            r[k] = (gl_x[k] - a) * q[k] - sb[j] * qp[k];
            nrm += om[k] * r[k] * r[k];
        }
        sb[j + 1] = sqrtl(nrm);
        for (int k = 0; k < ORC_NQUAD; k++) {
            qp[k] = q[k];
            q[k] = r[k] / sb[j + 1];
        }
    }
    long double d[2 * ORC_NMAX], e[2 * ORC_NMAX], z[2 * ORC_NMAX];
    for (int j = 0; j < m; j++) {
        d[j] = alpha[j];
        e[j] = sb[j];
        z[j] = j == 0;
    }
    tql_first(m, d, e, z);
    /* pick the positive nodes, ascending */
    int idx[2 * ORC_NMAX];
    for (int j = 0; j < m; j++) idx[j] = j;
    for (int a = 0; a < m; a++)
        for (int b = a + 1; b < m; b++)
            if (d[idx[b]] < d[idx[a]]) {
                int t = idx[a];
                idx[a] = idx[b];
                idx[b] = t;
            }
    for (int i = 0; i < n; i++) {
        int j = idx[n + i];
        s_out[i] = d[j] * d[j];
        w_out[i] = b0 * z[j] * z[j];
    }
}

/* Gauss-Hermite 2n-point rule (weight exp(-h^2) on the real line): the n
 * positive nodes squared and their weights, via the Jacobi matrix with
 * off-diagonal sqrt(j/2). */
static void hermite_half(int n, double *s_out, double *w_out) {
    int m = 2 * n;
    long double d[2 * ORC_NMAX], e[2 * ORC_NMAX], z[2 * ORC_NMAX];
    for (int j = 0; j < m; j++) {
        d[j] = 0;
        e[j] = j == 0 ? 0 : sqrtl(j / 2.0L);
        z[j] = j == 0;
    }
    tql_first(m, d, e, z);
    int idx[2 * ORC_NMAX];
    for (int j = 0; j < m; j++) idx[j] = j;
    for (int a = 0; a < m; a++)
        for (int b = a + 1; b < m; b++)
            if (d[idx[b]] < d[idx[a]]) {
                int t = idx[a];
                idx[a] = idx[b];
                idx[b] = t;
            }
    long double mu0 = sqrtl((long double)M_PI);
    for (int i = 0; i < n; i++) {
        int j = idx[n + i];
        s_out[i] = (double)(d[j] * d[j]);
        w_out[i] = (double)(mu0 * z[j] * z[j]);
    }
}

void oracle_rys_init(void) {
    if (orc_ready) return;
    gauss_legendre(ORC_NQUAD, gl_x, gl_w);
    for (int n = 1; n <= ORC_NMAX; n++) {
        orc_xbig[n] = 40 + 8 * n;
        int nint = (int)orc_xbig[n];
        orc_nint[n] = nint;
        hermite_half(n, orc_herm_s[n], orc_herm_w[n]);
        double *tab = (double *)malloc(sizeof(double) * nint * 2 * n * (ORC_DEG + 1));
        orc_tab[n] = tab;
#pragma omp parallel for schedule(dynamic)
        for (int it = 0; it < nint; it++) {
            long double vals[ORC_DEG + 1][2 * ORC_NMAX];
            for (int k = 0; k <= ORC_DEG; k++) {
                long double y = cosl(M_PI * (k + 0.5L) / (ORC_DEG + 1));
                long double x = it + 0.5L + 0.5L * y;
                long double s[ORC_NMAX], w[ORC_NMAX];
                rys_stieltjes(n, x, s, w);
                for (int i = 0; i < n; i++) {
                    vals[k][i] = s[i];
                    vals[k][n + i] = w[i];
                }
            }
            for (int f = 0; f < 2 * n; f++) {
                for (int j = 0; j <= ORC_DEG; j++) {
                    long double c = 0;
                    for (int k = 0; k <= ORC_DEG; k++)
                        c += vals[k][f] * cosl(M_PI * j * (k + 0.5L) / (ORC_DEG + 1));
                    c *= 2.0L / (ORC_DEG + 1);
                    if (j == 0) c *= 0.5L;
                    tab[((size_t)it * 2 * n + f) * (ORC_DEG + 1) + j] = (double)c;
                }
            }
        }
    }
    orc_ready = 1;
}

/* Rys nodes in the form u = t^2/(1-t^2) and weights, n <= ORC_NMAX. */
void oracle_rys(int n, double x, double *u, double *w) {
    if (x >= orc_xbig[n]) {
        double rx = 1 / x, sq = sqrt(rx);
        for (int i = 0; i < n; i++) {
            double s = orc_herm_s[n][i] * rx;
            u[i] = s / (1 - s);
            w[i] = orc_herm_w[n][i] * sq;
        }
        return;
    }
    int it = (int)x;
    double y = 2 * (x - it) - 1;
    const double *tab = orc_tab[n] + (size_t)it * 2 * n * (ORC_DEG + 1);
    for (int f = 0; f < 2 * n; f++) {
        const double *c = tab + f * (ORC_DEG + 1);
        double b1 = 0, b2 = 0;
        for (int j = ORC_DEG; j >= 1; j--) {
            double t = 2 * y * b1 - b2 + c[j];
            b2 = b1;
            b1 = t;
        }
        double v = y * b1 - b2 + c[0];
        if (f < n) u[f] = v / (1 - v);
        else w[f - n] = v;
    }
}

/* Exact (long double Stieltjes) evaluation, for checking the tables. */
void oracle_rys_direct(int n, double x, double *s, double *w) {
    if (!orc_ready) oracle_rys_init();
    long double sl[ORC_NMAX], wl[ORC_NMAX];
    rys_stieltjes(n, x, sl, wl);
    for (int i = 0; i < n; i++) {
        s[i] = (double)sl[i];
        w[i] = (double)wl[i];
    }
}

/* Boys function F_m(x), m = 0..mmax, by downward recursion from a
 * convergent series (all terms positive). */
void oracle_boys(int mmax, double x, double *f) {
    int mt = mmax + 40;
    long double ex = expl(-(long double)x);
    if (x > 2 * mmax + 40) {
        /* upward recursion from F_0 = sqrt(pi/x) erf(sqrt x)/2, stable for x >> m */
        long double fm = 0.5L * sqrtl(M_PI / (long double)x) * erfl(sqrtl((long double)x));
        f[0] = (double)fm;
        for (int m = 0; m < mmax; m++) {
            fm = ((2 * m + 1) * fm - ex) / (2.0L * x);
            f[m + 1] = (double)fm;
        }
        return;
    }
    long double term = 1.0L / (2 * mt + 1), sum = term;
    for (int k = 1; k < 2000; k++) {
        term *= 2.0L * x / (2 * mt + 2 * k + 1);
        sum += term;
        if (term < 1e-30L * sum) break;
    }
    long double fm = ex * sum;
    for (int m = mt; m > mmax; m--) fm = (2 * x * fm + ex) / (2 * m - 1);
    f[mmax] = (double)fm;
    for (int m = mmax; m > 0; m--) {
        fm = (2 * x * fm + ex) / (2 * m - 1);
        f[m - 1] = (double)fm;
    }
}

/* ---------------------------------------------------------------------- */
/* Primitive-pair data (cuobc/lib/gint/pair_data.c:83-142)                */
/* ---------------------------------------------------------------------- */
static double common_fac_sp(int l) {
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

/* For every pair of the list: a12, e12, P (SoA, 5 x nprim_pairs).  e12 is
 * scaled by diag_fac when both shells are the same. */
void oracle_pair_data(double *aexyz, int nprim_pairs, const int *pair2bra,
                      const int *pair2ket, int npairs, double diag_fac,
                      const int *atm, const int *bas, const double *env) {
    double *a12 = aexyz, *e12 = a12 + nprim_pairs, *x12 = e12 + nprim_pairs;
    double *y12 = x12 + nprim_pairs, *z12 = y12 + nprim_pairs;
    size_t off = 0;
    for (int p = 0; p < npairs; p++) {
        int ish = pair2bra[p], jsh = pair2ket[p];
        const int *bi = bas + ish * BAS_SLOTS, *bj = bas + jsh * BAS_SLOTS;
        const double *ai = env + bi[PTR_EXP], *aj = env + bj[PTR_EXP];
        const double *ci = env + bi[PTR_COEFF], *cj = env + bj[PTR_COEFF];
        const double *ri = env + atm[bi[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        const double *rj = env + atm[bj[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        double fac = common_fac_sp(bi[ANG_OF]) * common_fac_sp(bj[ANG_OF]);
        if (ish == jsh) fac *= diag_fac;
        double dx = ri[0] - rj[0], dy = ri[1] - rj[1], dz = ri[2] - rj[2];
        double rr = dx * dx + dy * dy + dz * dz;
        for (int ip = 0; ip < bi[NPRIM_OF]; ip++)
            for (int jp = 0; jp < bj[NPRIM_OF]; jp++, off++) {
                double aij = ai[ip] + aj[jp];
                a12[off] = aij;
                e12[off] = fac * ci[ip] * cj[jp] * exp(-rr * ai[ip] * aj[jp] / aij);
                x12[off] = (ai[ip] * ri[0] + aj[jp] * rj[0]) / aij;
                y12[off] = (ai[ip] * ri[1] + aj[jp] * rj[1]) / aij;
                z12[off] = (ai[ip] * ri[2] + aj[jp] * rj[2]) / aij;
            }
    }
}

/* ---------------------------------------------------------------------- */
/* Contracted Cartesian ERI block of one shell quartet (Rys)              */
/* ---------------------------------------------------------------------- */
#define NCART(l) (((l) + 1) * ((l) + 2) / 2)
#define L1 (ORC_LMAX + 1)
#define NMAXG (2 * ORC_LMAX + 1)

typedef struct {
    int l[4];
    int ish[4];
    const double *r[4]; /* centres */
} Quartet;

/* Cartesian exponents in storage order (lx from l down, then ly). */
static void cart_exps(int l, int *ex, int *ey, int *ez) {
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--, n++) {
            ex[n] = lx;
            ey[n] = ly;
            ez[n] = l - lx - ly;
        }
}

/* One primitive quartet, all roots: accumulates the Cartesian block into
 * out[((i*nfj + j)*nfk + k)*nfl + l].  The per-direction 2-D integrals follow
 * g2e.cu:32-490: VRR in (n = i+j, m = k+l) with c00/c0p/b00/b10/b01, then
 * transfer of angular momentum to j (A-B) and to l (C-D); the three
 * directions are seeded with norm, the pair prefactor and the weight. */
static void prim_quartet(const Quartet *q, double aij, double eij, const double *P,
                         double akl, double ekl, const double *Q, double norm,
                         double *out) {
    int li = q->l[0], lj = q->l[1], lk = q->l[2], ll = q->l[3];
    int nmax = li + lj, mmax = lk + ll;
    int nroots = (li + lj + lk + ll) / 2 + 1;
    double aijkl = aij + akl, a1 = aij * akl, a0 = a1 / aijkl;
    double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
    double x = a0 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
    double u[ORC_NMAX], w[ORC_NMAX];
    oracle_rys(nroots, x, u, w);
    double fac = eij * ekl / (sqrt(aijkl) * a1);
    static const int dummy = 0;
    (void)dummy;
    int nfi = NCART(li), nfj = NCART(lj), nfk = NCART(lk), nfl = NCART(ll);
    int ix[4][NCART(ORC_LMAX)], iy[4][NCART(ORC_LMAX)], iz[4][NCART(ORC_LMAX)];
    for (int c = 0; c < 4; c++) cart_exps(q->l[c], ix[c], iy[c], iz[c]);

    /* I[d][i][j][k][l] for one root */
    double I[3][L1][L1][L1][L1];
    for (int r = 0; r < nroots; r++) {
        double u2 = a0 * u[r];
        double tmp4 = .5 / (u2 * aijkl + a1);
        double b00 = u2 * tmp4;
        double b10 = b00 + tmp4 * akl;
        double b01 = b00 + tmp4 * aij;
        double tmp2 = 2 * b00 * akl, tmp3 = 2 * b00 * aij;
        double seed[3] = {norm, fac, w[r]};
        for (int d = 0; d < 3; d++) {
            double c00 = (P[d] - q->r[0][d]) - tmp2 * pq[d];
            double c0p = (Q[d] - q->r[2][d]) + tmp3 * pq[d];
            double G[NMAXG][NMAXG];
            G[0][0] = seed[d];
            if (nmax > 0) G[1][0] = c00 * G[0][0];
            for (int n = 1; n < nmax; n++) G[n + 1][0] = c00 * G[n][0] + n * b10 * G[n - 1][0];
            for (int m = 0; m < mmax; m++) {
                for (int n = 0; n <= nmax; n++) {
                    double v = c0p * G[n][m];
                    if (m > 0) v += m * b01 * G[n][m - 1];
                    if (n > 0) v += n * b00 * G[n - 1][m];
                    G[n][m + 1] = v;
                }
            }
            /* transfer to j: (i, j+1) = (i+1, j) + (A-B)(i, j) */
            double ab = q->r[0][d] - q->r[1][d];
            double cd = q->r[2][d] - q->r[3][d];
            double H[NMAXG][L1][NMAXG]; /* [i][j][m] */
            for (int n = 0; n <= nmax; n++)
                for (int m = 0; m <= mmax; m++) H[n][0][m] = G[n][m];
            for (int j = 1; j <= lj; j++)
                for (int i = 0; i <= nmax - j; i++)
                    for (int m = 0; m <= mmax; m++)
                        H[i][j][m] = H[i + 1][j - 1][m] + ab * H[i][j - 1][m];
            /* transfer to l: (k, l+1) = (k+1, l) + (C-D)(k, l) */
            for (int i = 0; i <= li; i++)
                for (int j = 0; j <= lj; j++) {
                    double K[NMAXG][L1];
                    for (int m = 0; m <= mmax; m++) K[m][0] = H[i][j][m];
                    for (int l = 1; l <= ll; l++)
                        for (int k = 0; k <= mmax - l; k++)
                            K[k][l] = K[k + 1][l - 1] + cd * K[k][l - 1];
                    for (int k = 0; k <= lk; k++)
                        for (int l = 0; l <= ll; l++) I[d][i][j][k][l] = K[k][l];
                }
        }
        int n = 0;
        for (int i = 0; i < nfi; i++)
            for (int j = 0; j < nfj; j++)
                for (int k = 0; k < nfk; k++)
                    for (int l = 0; l < nfl; l++, n++)
                        out[n] += I[0][ix[0][i]][ix[1][j]][ix[2][k]][ix[3][l]] *
                                  I[1][iy[0][i]][iy[1][j]][iy[2][k]][iy[3][l]] *
                                  I[2][iz[0][i]][iz[1][j]][iz[2][k]][iz[3][l]];
    }
}

/* Contracted block of (ij|kl) for shells i,j,k,l using pair data. */
static void contracted_quartet(const Quartet *q, const double *aexyz_ij, int npp_ij,
                               int nprim_ij, int off_ij, const double *aexyz_kl,
                               int npp_kl, int nprim_kl, int off_kl, double norm,
                               double *out) {
    const double *a12 = aexyz_ij, *e12 = a12 + npp_ij, *x12 = e12 + npp_ij;
    const double *y12 = x12 + npp_ij, *z12 = y12 + npp_ij;
    const double *b12 = aexyz_kl, *f12 = b12 + npp_kl, *u12 = f12 + npp_kl;
    const double *v12 = u12 + npp_kl, *w12 = v12 + npp_kl;
    for (int p = off_ij; p < off_ij + nprim_ij; p++) {
        double P[3] = {x12[p], y12[p], z12[p]};
        for (int s = off_kl; s < off_kl + nprim_kl; s++) {
            double Q[3] = {u12[s], v12[s], w12[s]};
            prim_quartet(q, a12[p], e12[p], P, b12[s], f12[s], Q, norm, out);
        }
    }
}

/* Cartesian block of the contracted quartet (ish jsh | ksh lsh), no pair
 * weights, for brute-force checks.  out has nfi*nfj*nfk*nfl entries, index
 * ((i*nfj + j)*nfk + k)*nfl + l. */
void oracle_eri_shells(int ish, int jsh, int ksh, int lsh, const int *atm,
                       const int *bas, const double *env, double *out) {
    if (!orc_ready) oracle_rys_init();
    int sh[4] = {ish, jsh, ksh, lsh};
    Quartet q;
    int nf = 1;
    for (int c = 0; c < 4; c++) {
        const int *b = bas + sh[c] * BAS_SLOTS;
        q.l[c] = b[ANG_OF];
        q.ish[c] = sh[c];
        q.r[c] = env + atm[b[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        nf *= NCART(q.l[c]);
    }
    int np_ij = bas[ish * BAS_SLOTS + NPRIM_OF] * bas[jsh * BAS_SLOTS + NPRIM_OF];
    int np_kl = bas[ksh * BAS_SLOTS + NPRIM_OF] * bas[lsh * BAS_SLOTS + NPRIM_OF];
    double *pij = (double *)malloc(sizeof(double) * 5 * np_ij);
    double *pkl = (double *)malloc(sizeof(double) * 5 * np_kl);
    int b0[2] = {ish, ksh}, b1[2] = {jsh, lsh};
    oracle_pair_data(pij, np_ij, b0, b1, 1, 1.0, atm, bas, env);
    oracle_pair_data(pkl, np_kl, b0 + 1, b1 + 1, 1, 1.0, atm, bas, env);
    memset(out, 0, sizeof(double) * nf);
    double fac = 2 * M_PI * M_PI * M_PI / sqrt(M_PI); /* g2e.c:41 */
    contracted_quartet(&q, pij, np_ij, np_ij, 0, pkl, np_kl, np_kl, 0, fac, out);
    free(pij);
    free(pkl);
}

/* Schwarz bound per shell pair: sqrt(max_{i in I, j in J} |(ij|ij)|),
 * 1e-100 when that is zero (PySCF CVHFnr_int2e_q_cond).  q is nbas x nbas. */
void oracle_q_cond(double *qc, const int *atm, int natm, const int *bas, int nbas,
                   const double *env) {
    if (!orc_ready) oracle_rys_init();
    (void)natm;
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * ORC_BLK);
#pragma omp for schedule(dynamic, 4)
        for (int ish = 0; ish < nbas; ish++) {
            for (int jsh = 0; jsh <= ish; jsh++) {
                oracle_eri_shells(ish, jsh, ish, jsh, atm, bas, env, buf);
                int nfi = NCART(bas[ish * BAS_SLOTS + ANG_OF]);
                int nfj = NCART(bas[jsh * BAS_SLOTS + ANG_OF]);
                double m = 0;
                for (int i = 0; i < nfi; i++)
                    for (int j = 0; j < nfj; j++) {
                        double v = fabs(buf[((i * nfj + j) * nfi + i) * nfj + j]);
                        if (v > m) m = v;
                    }
                double qv = m > 0 ? sqrt(m) : 1e-100;
                qc[ish * nbas + jsh] = qv;
                qc[jsh * nbas + ish] = qv;
            }
        }
        free(buf);
    }
}

/* ---------------------------------------------------------------------- */
/* J/K over the quartet set of one or many tasks                          */
/* ---------------------------------------------------------------------- */
/*
 * Tasks follow hf.py:111-166 and nr_jk_driver.cu:227-283: task t covers
 * pair-type cp_ij[t] >= cp_kl[t]; for kl_bin in [0,nbins) the kl pairs of
 * that bin are paired with the ij pairs of bins [0, nbins-kl_bin); within
 * a task with cp_ij == cp_kl only bas_ij >= bas_kl is kept, with weight 1/2
 * when equal.  Pair data e12 already carries 1/2 for ish == jsh.
 *
 * Digestion (contract_jk.cu:28-316), V accumulated in the orientation
 * V[a][b] with a the row AO:
 *   J: V[i,j] += g D[k,l],  V[k,l] += g D[i,j]
 *   K: V[i,k] += g D[j,l],  V[i,l] += g D[j,k],
 *      V[j,k] += g D[i,l],  V[j,l] += g D[i,k]
 * and the caller forms J = 2 (V + V^T), K = V + V^T (hf.py:171-181).
 * kl_stride > 1 keeps only the kl pairs with kl % kl_stride == 0 (a uniform
 * sample of every task, used to time the CPU path on a bounded workload).
 */
int oracle_jk(int ntask, const int *task_cp, const int *task_bins, int nbins,
              const int *bas_pairs_locs, const int *prim_pairs_locs,
              const int *cp_nprim, const int *pair2bra, const int *pair2ket,
              int npairs_total, const double *aexyz, int nprim_pairs_total,
              const int *atm, const int *bas, const double *env, const int *ao_loc,
              int nao, int n_dm, const double *dm, double *vj, double *vk,
              long long *nquartets_out, int kl_stride) {
    if (!orc_ready) oracle_rys_init();
    (void)npairs_total;
    /* flatten (task, kl_bin, kl) work items */
    long long nitems = 0;
    for (int t = 0; t < ntask; t++) {
        const int *bk = task_bins + (size_t)t * 2 * (nbins + 1) + (nbins + 1);
        nitems += bk[nbins] - bk[0];
    }
    if (kl_stride < 1) kl_stride = 1;
    int *it_task = (int *)malloc(sizeof(int) * (nitems + 1));
    int *it_kl = (int *)malloc(sizeof(int) * (nitems + 1));
    int *it_bin = (int *)malloc(sizeof(int) * (nitems + 1));
    long long c = 0;
    for (int t = 0; t < ntask; t++) {
        const int *bk = task_bins + (size_t)t * 2 * (nbins + 1) + (nbins + 1);
        for (int b = 0; b < nbins; b++)
            for (int kl = bk[b]; kl < bk[b + 1]; kl++) {
                if (kl % kl_stride) continue;
                it_task[c] = t;
                it_kl[c] = kl;
                it_bin[c] = b;
                c++;
            }
    }
    size_t nao2 = (size_t)nao * nao;
    int nthreads = omp_get_max_threads();
    double *vpriv = (double *)calloc((size_t)nthreads * 2 * n_dm * nao2, sizeof(double));
    if (!vpriv) return 1;
    long long nq = 0;
    const double fac0 = 2 * M_PI * M_PI * M_PI / sqrt(M_PI); /* g2e.c:41 */
#pragma omp parallel reduction(+ : nq)
    {
        int tid = omp_get_thread_num();
        double *vjp = vpriv + (size_t)tid * 2 * n_dm * nao2;
        double *vkp = vjp + (size_t)n_dm * nao2;
        double *blk = (double *)malloc(sizeof(double) * ORC_BLK);
#pragma omp for schedule(dynamic, 16)
        for (long long item = 0; item < c; item++) {
            int t = it_task[item], kl = it_kl[item], kl_bin = it_bin[item];
            int cp_ij = task_cp[2 * t], cp_kl = task_cp[2 * t + 1];
            const int *bi = task_bins + (size_t)t * 2 * (nbins + 1);
            int ij1 = bi[nbins - kl_bin];
            int np_ij = cp_nprim[cp_ij], np_kl = cp_nprim[cp_kl];
            int bas_kl = bas_pairs_locs[cp_kl] + kl;
            int ksh = pair2bra[bas_kl], lsh = pair2ket[bas_kl];
            for (int ij = bi[0]; ij < ij1; ij++) {
                int bas_ij = bas_pairs_locs[cp_ij] + ij;
                if (bas_ij < bas_kl) continue;
                double norm = fac0;
                if (bas_ij == bas_kl) norm *= .5;
                int ish = pair2bra[bas_ij], jsh = pair2ket[bas_ij];
                int sh[4] = {ish, jsh, ksh, lsh};
                Quartet q;
                int nfs[4];
                for (int cc = 0; cc < 4; cc++) {
                    const int *b = bas + sh[cc] * BAS_SLOTS;
                    q.l[cc] = b[ANG_OF];
                    q.ish[cc] = sh[cc];
                    q.r[cc] = env + atm[b[ATOM_OF] * ATM_SLOTS + PTR_COORD];
                    nfs[cc] = NCART(q.l[cc]);
                }
                int nf = nfs[0] * nfs[1] * nfs[2] * nfs[3];
                memset(blk, 0, sizeof(double) * nf);
                int pij = prim_pairs_locs[cp_ij] + ij * np_ij;
                int pkl = prim_pairs_locs[cp_kl] + kl * np_kl;
                contracted_quartet(&q, aexyz, nprim_pairs_total, np_ij, pij, aexyz,
                                   nprim_pairs_total, np_kl, pkl, norm, blk);
                nq++;
                int i0 = ao_loc[ish], j0 = ao_loc[jsh], k0 = ao_loc[ksh], l0 = ao_loc[lsh];
                for (int idm = 0; idm < n_dm; idm++) {
                    const double *d = dm + idm * nao2;
                    double *vjd = vjp + idm * nao2, *vkd = vkp + idm * nao2;
                    int n = 0;
                    for (int i = i0; i < i0 + nfs[0]; i++)
                        for (int j = j0; j < j0 + nfs[1]; j++)
                            for (int k = k0; k < k0 + nfs[2]; k++)
                                for (int l = l0; l < l0 + nfs[3]; l++, n++) {
                                    double g = blk[n];
                                    if (vj) {
                                        vjd[i * nao + j] += g * d[k * nao + l];
                                        vjd[k * nao + l] += g * d[i * nao + j];
                                    }
                                    if (vk) {
                                        vkd[i * nao + k] += g * d[j * nao + l];
                                        vkd[i * nao + l] += g * d[j * nao + k];
                                        vkd[j * nao + k] += g * d[i * nao + l];
                                        vkd[j * nao + l] += g * d[i * nao + k];
                                    }
                                }
                }
            }
        }
        free(blk);
    }
    /* reduce thread-private accumulators */
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < (long long)(n_dm * nao2); e++) {
        double sj = 0, sk = 0;
        for (int tt = 0; tt < nthreads; tt++) {
            sj += vpriv[(size_t)tt * 2 * n_dm * nao2 + e];
            sk += vpriv[(size_t)tt * 2 * n_dm * nao2 + n_dm * nao2 + e];
        }
        if (vj) vj[e] += sj;
        if (vk) vk[e] += sk;
    }
    free(vpriv);
    free(it_task);
    free(it_kl);
    free(it_bin);
    if (nquartets_out) *nquartets_out = nq;
    return 0;
}

/* Threads of the oracle's OpenMP regions (for single-thread baselines). */
void oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

int oracle_get_threads(void) { return omp_get_max_threads(); }
