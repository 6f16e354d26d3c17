// Class dispatch and C-ABI entry points of the J/K build.
//
// Canonical classes have LA >= LB, LC >= LD and pair-id(LA,LB) >=
// pair-id(LC,LD) with pair-id(a,b) = a(a+1)/2 + b; a task whose ket class is
// the larger one is evaluated with bra and ket exchanged.  Kernel
// instantiations live in jk_bra_XY.hip, one translation unit per bra class.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "jk_v2.h"

static int pair_id(int a, int b) { return a * (a + 1) / 2 + b; }

static JKKernelTable g_kernels;
static RowKernelTable g_rows;
static bool g_registered = false;

#define BQC_DECL_BRA(LA, LB) \
    void bqc_register_bra_##LA##LB(JKKernelTable &t1, RowKernelTable &t2);
BQC_DECL_BRA(0, 0)
BQC_DECL_BRA(1, 0)
BQC_DECL_BRA(1, 1)
BQC_DECL_BRA(2, 0)
BQC_DECL_BRA(2, 1)
BQC_DECL_BRA(2, 2)
BQC_DECL_BRA(3, 0)
BQC_DECL_BRA(3, 1)
BQC_DECL_BRA(3, 2)
BQC_DECL_BRA(3, 3)

static void register_kernels() {
    if (g_registered) return;
    bqc_register_bra_00(g_kernels, g_rows);
    bqc_register_bra_10(g_kernels, g_rows);
    bqc_register_bra_11(g_kernels, g_rows);
    bqc_register_bra_20(g_kernels, g_rows);
    bqc_register_bra_21(g_kernels, g_rows);
    bqc_register_bra_22(g_kernels, g_rows);
    bqc_register_bra_30(g_kernels, g_rows);
    bqc_register_bra_31(g_kernels, g_rows);
    bqc_register_bra_32(g_kernels, g_rows);
    bqc_register_bra_33(g_kernels, g_rows);
    g_registered = true;
}

static double g_last_flops[2] = {0, 0};

// Optional per-launch timing: HIP events around every kernel of a build,
// on the build's stream.  Records (class code, algorithmic FLOPs, events).
struct LaunchRec {
    int cls;       // li*1000 + lj*100 + lk*10 + ll of the evaluated class
    int kind;      // 0 quartet kernel, 1 row kernel
    double flops;  // algorithmic FLOPs of the launch
    double nquart;
    hipEvent_t e0, e1;
};
static bool g_prof = false;
static std::vector<LaunchRec> g_recs;
static std::vector<hipEvent_t> g_evpool;

static hipEvent_t prof_event() {
    if (g_evpool.empty()) {
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    hipEvent_t e = g_evpool.back();
    g_evpool.pop_back();
    return e;
}

// Algorithmic FLOP model of one contracted quartet (K primitive quartets).
static double quartet_flops(int li, int lj, int lk, int ll, int K) {
    int nr = (li + lj + lk + ll) / 2 + 1;
    int nf = ncart(li) * ncart(lj) * ncart(lk) * ncart(ll);
    int nmax = li + lj, mmax = lk + ll;
    int ijmin = li < lj ? li : lj, klmin = lk < ll ? lk : ll;
    double f2d = 4.0 * nmax + 4.0 * mmax + 6.0 * nmax * mmax;
    for (int j = 1; j <= ijmin; ++j) f2d += 2.0 * (mmax + 1) * (nmax - j + 1);
    for (int l = 1; l <= klmin; ++l) f2d += 2.0 * (li + 1) * (lj + 1) * (mmax - l + 1);
    return K * (20.0 + nr * (60.0 + 20.0 + 3.0 * f2d + 3.0 * nf)) + 12.0 * nf;
}

static int launch_task(BasisProdCache *bp, const JKArgs &jk, const PairData &pd,
                       const int *bins_ij, const int *bins_kl, int nbins, int cp_ij,
                       int cp_kl, hipStream_t stream, double *flops) {
    const ContractionProdType *tij = bp->cptype + cp_ij, *tkl = bp->cptype + cp_kl;
    int li = tij->l_bra, lj = tij->l_ket, lk = tkl->l_bra, ll = tkl->l_ket;
    if ((li + lj + lk + ll) / 2 + 1 >= 12) return 2;
    if (li > BQC_LMAX || lj > BQC_LMAX || lk > BQC_LMAX || ll > BQC_LMAX || lj > li ||
        ll > lk)
        return 1;
    if (nbins > BQC_MAX_BINS) return 3;
    register_kernels();
    TaskArgs ta;
    memset(&ta, 0, sizeof(ta));
    int pij = pair_id(li, lj), pkl = pair_id(lk, ll);
    ta.swap = pij < pkl;
    ta.diag = cp_ij == cp_kl;
    const int *locs = bp->bas_pairs_locs, *plocs = bp->primitive_pairs_locs;
    int cp_p = ta.swap ? cp_kl : cp_ij, cp_q = ta.swap ? cp_ij : cp_kl;
    ta.bas_p0 = locs[cp_p];
    ta.bas_q0 = locs[cp_q];
    ta.prim_p0 = plocs[cp_p];
    ta.prim_q0 = plocs[cp_q];
    ta.nprim_p = bp->cptype[cp_p].nprim_12;
    ta.nprim_q = bp->cptype[cp_q].nprim_12;
    // work rectangles: kl pairs of bin b x ij pairs of bins [0, nbins-b)
    int nb = 0;
    ta.pre[0] = 0;
    double nq = 0;
    for (int b = 0; b < nbins; ++b) {
        int nkl = bins_kl[b + 1] - bins_kl[b];
        int nij = bins_ij[nbins - b] - bins_ij[0];
        if (nkl <= 0 || nij <= 0) continue;
        ta.klb[nb] = bins_kl[b];
        ta.nij[nb] = nij;
        ta.pre[nb + 1] = ta.pre[nb] + (long long)nkl * nij;
        if (ta.diag) {
            for (int kl = bins_kl[b]; kl < bins_kl[b + 1]; ++kl) nq += nij > kl ? nij - kl : 0;
        } else {
            nq += (double)nkl * nij;
        }
        ++nb;
    }
    ta.nbins = nb;
    if (nb == 0) return 0;
    if (bins_ij[0] != 0) return 1;
    const double lflops = nq * quartet_flops(li, lj, lk, ll, tij->nprim_12 * tkl->nprim_12);
    if (flops) *flops += lflops;
    JKKernel kern = ta.swap ? g_kernels[pkl][pij] : g_kernels[pij][pkl];
    if (!kern) return 1;
    long long ntot = ta.pre[nb];
    long long nblk = (ntot + BQC_THREADS - 1) / BQC_THREADS;
    LaunchRec rec;
    if (g_prof) {
        int a = ta.swap ? 1 : 0;
        rec.cls = a ? lk * 1000 + ll * 100 + li * 10 + lj : li * 1000 + lj * 100 + lk * 10 + ll;
        rec.kind = 0;
        rec.flops = lflops;
        rec.nquart = nq;
        rec.e0 = prof_event();
        rec.e1 = prof_event();
        hipEventRecord(rec.e0, stream);
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(BQC_THREADS), 0, stream, jk, pd, ta);
    if (g_prof) {
        hipEventRecord(rec.e1, stream);
        g_recs.push_back(rec);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        fprintf(stderr, "bqcjk: launch failed for (%d%d|%d%d): %s\n", li, lj, lk, ll,
                hipGetErrorString(err));
        return 1;
    }
    return 0;
}

static PairData make_pair_data(const BasisProdCache *bp) {
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
    return pd;
}

// ---------------------------------------------------------------------------
// wavefront-stationary kernels: host-side list order and work items
// ---------------------------------------------------------------------------
struct RowLaunch {
    RowKernel kern;
    V2Task task;
    size_t off_perm_p, off_perm_q, off_items;
    int cls;
    double flops, nquart;
};

static int *g_dws = nullptr;         // device workspace (ints)
static size_t g_dws_cap = 0;
static int *g_hstage = nullptr;      // pinned staging
static size_t g_hstage_cap = 0;
static hipEvent_t g_stage_evt = nullptr;
static bool g_stage_pending = false;

static int stage_to_device(const std::vector<int> &stage, hipStream_t stream) {
    size_t n = stage.size();
    if (g_stage_pending) {
        hipEventSynchronize(g_stage_evt);
        g_stage_pending = false;
    }
    if (n > g_hstage_cap) {
        if (g_hstage) hipHostFree(g_hstage);
        size_t cap = n + n / 2 + 1024;
        if (hipHostMalloc((void **)&g_hstage, cap * sizeof(int), hipHostMallocDefault) != hipSuccess)
            return 1;
        g_hstage_cap = cap;
    }
    if (n > g_dws_cap) {
        // the previous workspace may still be read by queued kernels
        hipStreamSynchronize(stream);
        if (g_dws) hipFree(g_dws);
        size_t cap = n + n / 2 + 1024;
        if (hipMalloc((void **)&g_dws, cap * sizeof(int)) != hipSuccess) return 1;
        g_dws_cap = cap;
    }
    if (!g_stage_evt) hipEventCreateWithFlags(&g_stage_evt, hipEventDisableTiming);
    memcpy(g_hstage, stage.data(), n * sizeof(int));
    if (hipMemcpyAsync(g_dws, g_hstage, n * sizeof(int), hipMemcpyHostToDevice, stream) !=
        hipSuccess)
        return 1;
    hipEventRecord(g_stage_evt, stream);
    g_stage_pending = true;
    return 0;
}

// List order of one side: plan positions [0, bins[nbins]) grouped by bin,
// inside a bin by first shell (bra_order gives the plan positions of the
// pair type sorted by first shell).
static void side_order(const int *bins, int nbins, const int *order, int npairs_cp,
                       std::vector<int> &stage) {
    const int n = bins[nbins];
    std::vector<int> cursor(bins, bins + nbins);
    size_t base = stage.size();
    stage.resize(base + n);
    for (int k = 0; k < npairs_cp; ++k) {
        int pos = order[k];
        if (pos >= n) continue;
        int b = (int)(std::upper_bound(bins + 1, bins + nbins + 1, pos) - (bins + 1));
        stage[base + cursor[b]++] = pos;
    }
}

// returns -1 if the task's class has no row kernel (caller falls back)
static int plan_rows(BasisProdCache *bp, const int *bins_ij, const int *bins_kl, int nbins,
                     int cp_ij, int cp_kl, const int *bra_order, std::vector<int> &stage,
                     std::vector<RowLaunch> &rows, double *flops) {
    const ContractionProdType *tij = bp->cptype + cp_ij, *tkl = bp->cptype + cp_kl;
    int li = tij->l_bra, lj = tij->l_ket, lk = tkl->l_bra, ll = tkl->l_ket;
    if (li > BQC_LMAX || lj > BQC_LMAX || lk > BQC_LMAX || ll > BQC_LMAX || lj > li || ll > lk)
        return 1;
    int pij = pair_id(li, lj), pkl = pair_id(lk, ll);
    bool swap = pij < pkl;
    RowKernel kern = swap ? g_rows[pkl][pij] : g_rows[pij][pkl];
    if (!kern || !bra_order) return -1;
    const int *pbins = swap ? bins_kl : bins_ij;
    const int *qbins = swap ? bins_ij : bins_kl;
    int cp_p = swap ? cp_kl : cp_ij, cp_q = swap ? cp_ij : cp_kl;
    const int *locs = bp->bas_pairs_locs;
    int np = pbins[nbins], nq = qbins[nbins];
    if (pbins[0] != 0 || qbins[0] != 0) return 1;
    // quartet count for the FLOP model
    double nquart = 0;
    for (int b = 0; b < nbins; ++b) {
        int nkl = bins_kl[b + 1] - bins_kl[b];
        int nij = bins_ij[nbins - b];
        if (nkl <= 0 || nij <= 0) continue;
        if (cp_ij == cp_kl) {
            for (int kl = bins_kl[b]; kl < bins_kl[b + 1]; ++kl) nquart += nij > kl ? nij - kl : 0;
        } else {
            nquart += (double)nkl * nij;
        }
    }
    const double lflops = nquart * quartet_flops(li, lj, lk, ll, tij->nprim_12 * tkl->nprim_12);
    if (flops) *flops += lflops;
    if (np == 0 || nq == 0) return 0;
    RowLaunch r;
    memset(&r, 0, sizeof(r));
    r.kern = kern;
    r.cls = swap ? lk * 1000 + ll * 100 + li * 10 + lj : li * 1000 + lj * 100 + lk * 10 + ll;
    r.flops = lflops;
    r.nquart = nquart;
    V2Task &tk = r.task;
    tk.np = np;
    tk.bas_p0 = locs[cp_p];
    tk.bas_q0 = locs[cp_q];
    tk.prim_p0 = bp->primitive_pairs_locs[cp_p];
    tk.prim_q0 = bp->primitive_pairs_locs[cp_q];
    tk.nprim_p = bp->cptype[cp_p].nprim_12;
    tk.nprim_q = bp->cptype[cp_q].nprim_12;
    tk.diag = cp_ij == cp_kl;
    tk.nbins = nbins;
    for (int b = 0; b <= nbins; ++b) {
        tk.pbins[b] = pbins[b];
        tk.qbins[b] = qbins[b];
    }
    r.off_perm_p = stage.size();
    side_order(pbins, nbins, bra_order + locs[cp_p], locs[cp_p + 1] - locs[cp_p], stage);
    if (cp_p == cp_q) {
        r.off_perm_q = r.off_perm_p;
    } else {
        r.off_perm_q = stage.size();
        side_order(qbins, nbins, bra_order + locs[cp_q], locs[cp_q + 1] - locs[cp_q], stage);
    }
    if (stage.size() & 1) stage.push_back(0);  // int2 alignment
    r.off_items = stage.size();
    int nchunk = (np + 63) / 64;
    int nitems = 0;
    for (int c = 0; c < nchunk; ++c) {
        int first = c * 64;
        int s = (int)(std::upper_bound(pbins + 1, pbins + nbins + 1, first) - (pbins + 1));
        int qmax = qbins[nbins - s];
        for (int q0 = 0; q0 < qmax; q0 += BQC_V2_CK) {
            stage.push_back(c);
            stage.push_back(q0);
            ++nitems;
        }
    }
    tk.nitems = nitems;
    if (nitems) rows.push_back(r);
    return 0;
}

extern "C" {

int GINTbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                 int n_dm, int *bins_locs_ij, int *bins_locs_kl, int nbins, int cp_ij_id,
                 int cp_kl_id, hipStream_t stream) {
    JKArgs jk = {vj, vk, dm, nao, n_dm};
    PairData pd = make_pair_data(bpcache);
    return launch_task(bpcache, jk, pd, bins_locs_ij, bins_locs_kl, nbins, cp_ij_id,
                       cp_kl_id, stream, nullptr);
}

int BQCbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                int n_dm, int ntask, const int *task_cp, const int *task_bins, int nbins,
                const int *bra_order, int precision, const double *log_q,
                const double *task_dmax, double fp32_cut, hipStream_t stream) {
    (void)log_q;
    (void)task_dmax;
    (void)fp32_cut;
    if (precision != 0) return 1;
    if (nbins > BQC_MAX_BINS) return 3;
    register_kernels();
    JKArgs jk = {vj, vk, dm, nao, n_dm};
    PairData pd = make_pair_data(bpcache);
    double flops = 0;
    std::vector<RowLaunch> rows;
    std::vector<int> stage;
    for (int t = 0; t < ntask; ++t) {
        const int *bi = task_bins + (size_t)t * 2 * (nbins + 1);
        const int *bk = bi + nbins + 1;
        int cp_ij = task_cp[2 * t], cp_kl = task_cp[2 * t + 1];
        int err = plan_rows(bpcache, bi, bk, nbins, cp_ij, cp_kl, bra_order, stage, rows, &flops);
        if (err < 0) {
            err = launch_task(bpcache, jk, pd, bi, bk, nbins, cp_ij, cp_kl, stream, &flops);
        }
        if (err > 0) return err;
    }
    if (!rows.empty()) {
        int err = stage_to_device(stage, stream);
        if (err) return err;
        for (RowLaunch &r : rows) {
            r.task.perm_p = g_dws + r.off_perm_p;
            r.task.perm_q = g_dws + r.off_perm_q;
            r.task.items = (const int2 *)(g_dws + r.off_items);
            unsigned nblk = (unsigned)((r.task.nitems + 3) / 4);
            LaunchRec rec;
            if (g_prof) {
                rec.cls = r.cls;
                rec.kind = 1;
                rec.flops = r.flops;
                rec.nquart = r.nquart;
                rec.e0 = prof_event();
                rec.e1 = prof_event();
                hipEventRecord(rec.e0, stream);
            }
            hipLaunchKernelGGL(r.kern, dim3(nblk), dim3(256), 0, stream, jk, pd, r.task);
            if (g_prof) {
                hipEventRecord(rec.e1, stream);
                g_recs.push_back(rec);
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                fprintf(stderr, "bqcjk: row kernel launch failed: %s\n", hipGetErrorString(e));
                return 1;
            }
        }
    }
    g_last_flops[0] = flops;
    g_last_flops[1] = 0;
    return 0;
}

void BQCprofile_enable(int on) {
    g_prof = on != 0;
}

int BQCprofile_count(void) { return (int)g_recs.size(); }

// Drains the launch records: cls, kind, flops, quartets and duration (ms)
// per launch; waits for the recorded events.  Returns the number written.
int BQCprofile_read(int *cls, int *kind, double *flops, double *nquart, double *ms, int cap) {
    int n = 0;
    for (LaunchRec &r : g_recs) {
        if (n < cap) {
            hipEventSynchronize(r.e1);
            float t = 0;
            hipEventElapsedTime(&t, r.e0, r.e1);
            cls[n] = r.cls;
            kind[n] = r.kind;
            flops[n] = r.flops;
            nquart[n] = r.nquart;
            ms[n] = t;
            ++n;
        }
        g_evpool.push_back(r.e0);
        g_evpool.push_back(r.e1);
    }
    g_recs.clear();
    return n;
}

void BQClast_build_flops(double *flops) {
    flops[0] = g_last_flops[0];
    flops[1] = g_last_flops[1];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// J = 2 (V + V^T), K = V + V^T
// ---------------------------------------------------------------------------
__global__ void symmetrize_kernel(double *v, int nao, double fac) {
    const size_t nao2 = (size_t)nao * nao;
    double *m = v + blockIdx.z * nao2;
    const int i = blockIdx.y * 16 + threadIdx.y;
    const int j = blockIdx.x * 16 + threadIdx.x;
    if (i >= nao || j > i) return;
    const double s = fac * (m[(size_t)i * nao + j] + m[(size_t)j * nao + i]);
    m[(size_t)i * nao + j] = s;
    m[(size_t)j * nao + i] = s;
}

extern "C" int BQCsymmetrize_jk(double *vj, double *vk, int nao, int n_dm,
                                hipStream_t stream) {
    dim3 thr(16, 16);
    dim3 grd((nao + 15) / 16, (nao + 15) / 16, n_dm);
    if (vj) hipLaunchKernelGGL(symmetrize_kernel, grd, thr, 0, stream, vj, nao, 2.0);
    if (vk) hipLaunchKernelGGL(symmetrize_kernel, grd, thr, 0, stream, vk, nao, 1.0);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
