// Host-side shell-pair cache (C-ABI replacements of the reference's
// GINTinit_basis_prod_cpu / GINTinit_basis_prod_gpu / GINTdel_basis_prod,
// cuobc/lib/gint/nr_fill_ao_ints.cu:394-491 with pair_data.c:27-142).
//
// For every screened shell pair (ish, jsh) and every primitive pair (ip, jp):
//   a12 = ai + aj
//   e12 = fac_sp(li) fac_sp(lj) ci cj exp(-ai aj |Ri - Rj|^2 / a12)  (x diag_fac if ish == jsh)
//   P   = (ai Ri + aj Rj) / a12
// with fac_sp(0) = 1/(2 sqrt(pi)), fac_sp(1) = sqrt(3/(4 pi)), else 1.
//
// Primitive-pair screening (not in the reference, which keeps every
// primitive pair): inside each shell pair the primitive pairs are sorted by
// |e12|, largest first, and those below BQC_PRIM_CUT times the largest |e12|
// of the whole cache are set to zero.  The J/K kernels stop a primitive loop
// at the first zero.  A dropped pair's share of any integral is below
// ~1e-16 (|e12| <= 1e-20 max|e12|, times at most ~1e3 for the remaining
// factors of diffuse primitives), far under the 1e-13 Schwarz cutoff; in
// def2-SVP water clusters about 12 % of the primitive pairs of shell pairs
// that survive Schwarz screening are dropped.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#define BQC_PRIM_CUT 1e-20

#include "../../include/byteqc_jk.h"

#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define PTR_EXP 5
#define PTR_COEFF 6
#define BAS_SLOTS 8
#define PTR_COORD 1
#define ATM_SLOTS 6

static double fac_sp(int l) {
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

extern "C" {

void GINTinit_basis_prod_cpu(BasisProdCache **pbp, double diag_fac, double *bas_coords,
                             int *bas_pair2shls, int *bas_pairs_locs, int ncptype, int *atm,
                             int natm, int *bas, int nbas, double *env) {
    (void)natm;
    BasisProdCache *bp = (BasisProdCache *)calloc(1, sizeof(BasisProdCache));
    *pbp = bp;
    bp->nbas = nbas;
    bp->ncptype = ncptype;
    bp->bas_pair2shls = bas_pair2shls;
    bp->bas_pairs_locs = bas_pairs_locs;
    const int npairs = bas_pairs_locs[ncptype];
    const int *bra = bas_pair2shls, *ket = bas_pair2shls + npairs;
    bp->cptype = (ContractionProdType *)malloc(sizeof(ContractionProdType) * (ncptype > 0 ? ncptype : 1));
    bp->primitive_pairs_locs = (int *)malloc(sizeof(int) * (ncptype + 1));
    int nprim_pairs = 0;
    bp->primitive_pairs_locs[0] = 0;
    for (int n = 0; n < ncptype; ++n) {
        ContractionProdType *ct = bp->cptype + n;
        int p0 = bas_pairs_locs[n];
        int np = bas_pairs_locs[n + 1] - p0;
        ct->npairs = np;
        if (np > 0) {
            int ish = bra[p0], jsh = ket[p0];
            ct->l_bra = bas[ish * BAS_SLOTS + ANG_OF];
            ct->l_ket = bas[jsh * BAS_SLOTS + ANG_OF];
            ct->nprim_12 = bas[ish * BAS_SLOTS + NPRIM_OF] * bas[jsh * BAS_SLOTS + NPRIM_OF];
        } else {
            ct->l_bra = ct->l_ket = 0;
            ct->nprim_12 = 1;
        }
        nprim_pairs += np * ct->nprim_12;
        bp->primitive_pairs_locs[n + 1] = nprim_pairs;
    }
    for (int ib = 0; ib < nbas; ++ib) {
        const double *r = env + atm[bas[ib * BAS_SLOTS + ATOM_OF] * ATM_SLOTS + PTR_COORD];
        bas_coords[ib] = r[0];
        bas_coords[nbas + ib] = r[1];
        bas_coords[2 * nbas + ib] = r[2];
    }
    double *aexyz = (double *)malloc(sizeof(double) * 5 * (nprim_pairs > 0 ? nprim_pairs : 1));
    bp->aexyz = aexyz;
    double *a12 = aexyz, *e12 = a12 + nprim_pairs, *x12 = e12 + nprim_pairs;
    double *y12 = x12 + nprim_pairs, *z12 = y12 + nprim_pairs;
    size_t off = 0;
    for (int p = 0; p < npairs; ++p) {
        const int ish = bra[p], jsh = ket[p];
        const int *bi = bas + ish * BAS_SLOTS, *bj = bas + jsh * BAS_SLOTS;
        const double *ai = env + bi[PTR_EXP], *aj = env + bj[PTR_EXP];
        const double *ci = env + bi[PTR_COEFF], *cj = env + bj[PTR_COEFF];
        const double *ri = env + atm[bi[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        const double *rj = env + atm[bj[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        double fac = fac_sp(bi[ANG_OF]) * fac_sp(bj[ANG_OF]);
        if (ish == jsh) fac *= diag_fac;
        const double dx = ri[0] - rj[0], dy = ri[1] - rj[1], dz = ri[2] - rj[2];
        const double rr = dx * dx + dy * dy + dz * dz;
        for (int ip = 0; ip < bi[NPRIM_OF]; ++ip)
            for (int jp = 0; jp < bj[NPRIM_OF]; ++jp, ++off) {
                const double aij = ai[ip] + aj[jp];
                a12[off] = aij;
                e12[off] = fac * ci[ip] * cj[jp] * exp(-rr * ai[ip] * aj[jp] / aij);
                x12[off] = (ai[ip] * ri[0] + aj[jp] * rj[0]) / aij;
                y12[off] = (ai[ip] * ri[1] + aj[jp] * rj[1]) / aij;
                z12[off] = (ai[ip] * ri[2] + aj[jp] * rj[2]) / aij;
            }
    }
    double emax = 0.0;
    for (int k = 0; k < nprim_pairs; ++k) emax = std::max(emax, fabs(e12[k]));
    const double cut = BQC_PRIM_CUT * emax;
    std::vector<int> idx;
    std::vector<double> tmp;
    for (int n = 0; n < ncptype; ++n) {
        const int k = bp->cptype[n].nprim_12;
        if (k < 2) continue;
        idx.resize(k);
        tmp.resize(5 * (size_t)k);
        for (int pr = bas_pairs_locs[n]; pr < bas_pairs_locs[n + 1]; ++pr) {
            const size_t base = (size_t)bp->primitive_pairs_locs[n] +
                                (size_t)(pr - bas_pairs_locs[n]) * k;
            for (int i = 0; i < k; ++i) idx[i] = i;
            std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) {
                return fabs(e12[base + u]) > fabs(e12[base + v]);
            });
            double *arr[5] = {a12, e12, x12, y12, z12};
            for (int f = 0; f < 5; ++f) {
                for (int i = 0; i < k; ++i) tmp[(size_t)f * k + i] = arr[f][base + idx[i]];
                for (int i = 0; i < k; ++i) arr[f][base + i] = tmp[(size_t)f * k + i];
            }
            for (int i = 0; i < k; ++i)
                if (fabs(e12[base + i]) < cut) e12[base + i] = 0.0;
        }
    }
}

void GINTinit_basis_prod_gpu(BasisProdCache **in, BasisProdCache **out, int *ao_loc,
                             double *bas_coords, double *aexyz, int *bas_pair2shls) {
    BasisProdCache *src = *in;
    BasisProdCache *dst = (BasisProdCache *)malloc(sizeof(BasisProdCache));
    memcpy(dst, src, sizeof(BasisProdCache));
    const int npairs = src->bas_pairs_locs[src->ncptype];
    const int nprim_pairs = src->primitive_pairs_locs[src->ncptype];
    dst->ao_loc = ao_loc;
    dst->bas_coords = bas_coords;
    dst->bas_pair2bra = bas_pair2shls;
    dst->bas_pair2ket = bas_pair2shls + npairs;
    dst->a12 = aexyz;
    dst->e12 = aexyz + nprim_pairs;
    dst->x12 = aexyz + 2 * (size_t)nprim_pairs;
    dst->y12 = aexyz + 3 * (size_t)nprim_pairs;
    dst->z12 = aexyz + 4 * (size_t)nprim_pairs;
    *out = dst;
}

// Drops the first-shell order the J/K driver keeps per cache (jk.hip).
void bqc_forget_pair_order(const BasisProdCache *bp);

void GINTdel_basis_prod(BasisProdCache **pbp) {
    BasisProdCache *bp = *pbp;
    if (!bp) return;
    bqc_forget_pair_order(bp);
    free(bp->cptype);
    free(bp->primitive_pairs_locs);
    free(bp->aexyz);
    free(bp);
    *pbp = NULL;
}

void BQCdel_basis_prod_clone(BasisProdCache **pbp) {
    if (*pbp) bqc_forget_pair_order(*pbp);
    if (*pbp) free(*pbp);
    *pbp = NULL;
}

}  // extern "C"
