// Host-side task list of one J/K build (replaces the per-task Python loop of
// cuobc/scf/hf.py:111-145 and _make_s_index_offsets, hf.py:567-578).
//
// Pair types cp = 0..ncp-1 are (group gi, group gj) with gi >= gj in the
// order i = 0.., j = 0..i; task t = (a, b) with a >= b is decoded from the
// triangular index.  Density prescreen: sub = max of the six group-block
// bounds dm_cond[(gi,gj), (gk,gl), (gi,gk), (gj,gk), (gi,gl), (gj,gl)];
// a task is skipped if sub < cutoff (with prescreen) or a list is empty.
// Bins of a pair list sorted by descending log q, for cut = cutoff / sub:
//   s = floor(nbins / log(min(cut, 0.1)) * log q), counted for s < nbins.
#include <math.h>
#include <string.h>

extern "C" int BQCplan_tasks(int ncp, const int *cp_grp, const int *bas_pairs_locs,
                             const double *log_q, const double *dm_cond, int ngroups,
                             double cutoff, int nbins, int prescreen, int *task_cp,
                             int *task_bins, double *task_dmax, int *ntask_out) {
    int nt = 0;
    const int nb1 = nbins + 1;
    for (int a = 0; a < ncp; ++a) {
        for (int b = 0; b <= a; ++b) {
            const int na = bas_pairs_locs[a + 1] - bas_pairs_locs[a];
            const int nbb = bas_pairs_locs[b + 1] - bas_pairs_locs[b];
            if (na == 0 || nbb == 0) continue;
            const int i = cp_grp[2 * a], j = cp_grp[2 * a + 1];
            const int k = cp_grp[2 * b], l = cp_grp[2 * b + 1];
            const double *dc = dm_cond;
            const int G = ngroups;
            double sub = dc[i * G + j];
            const double c2[5] = {dc[k * G + l], dc[i * G + k], dc[j * G + k], dc[i * G + l],
                                  dc[j * G + l]};
            for (int m = 0; m < 5; ++m)
                if (c2[m] > sub) sub = c2[m];
            if (prescreen && sub < cutoff) continue;
            const double cut = sub > 0 ? cutoff / sub : 1e300;
            const double scale = nbins / log(cut < .1 ? cut : .1);
            int *out = task_bins + (size_t)nt * 2 * nb1;
            for (int side = 0; side < 2; ++side) {
                const int cp = side == 0 ? a : b;
                int *bins = out + side * nb1;
                memset(bins, 0, sizeof(int) * nb1);
                const double *lq = log_q + bas_pairs_locs[cp];
                const int n = bas_pairs_locs[cp + 1] - bas_pairs_locs[cp];
                for (int p = 0; p < n; ++p) {
                    const double sv = floor(scale * lq[p]);
                    if (sv < nbins) bins[(int)sv + 1]++;
                }
                for (int s = 0; s < nbins; ++s) bins[s + 1] += bins[s];
            }
            task_cp[2 * nt] = a;
            task_cp[2 * nt + 1] = b;
            task_dmax[nt] = sub;
            ++nt;
        }
    }
    *ntask_out = nt;
    return 0;
}
