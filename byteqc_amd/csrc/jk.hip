// Class dispatch and C-ABI entry points of the J/K build.
//
// Canonical classes have LA >= LB, LC >= LD and pair-id(LA,LB) >=
// pair-id(LC,LD) with pair-id(a,b) = a(a+1)/2 + b; a task whose ket class is
// the larger one is evaluated with bra and ket exchanged.  Kernel
// instantiations live in jk_bra_XY.hip, one translation unit per bra class.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "jk_wave.h"

static int pair_id(int a, int b) { return a * (a + 1) / 2 + b; }

static JKKernelTable g_kernels;
static RowKernelTable g_rows;
static WaveKernelTable g_waves;

#define BQC_DECL_BRA(LA, LB) \
    void bqc_register_bra_##LA##LB(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3);
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

static void register_kernels_once() {
    bqc_register_bra_00(g_kernels, g_rows, g_waves);
    bqc_register_bra_10(g_kernels, g_rows, g_waves);
    bqc_register_bra_11(g_kernels, g_rows, g_waves);
    bqc_register_bra_20(g_kernels, g_rows, g_waves);
    bqc_register_bra_21(g_kernels, g_rows, g_waves);
    bqc_register_bra_22(g_kernels, g_rows, g_waves);
    bqc_register_bra_30(g_kernels, g_rows, g_waves);
    bqc_register_bra_31(g_kernels, g_rows, g_waves);
    bqc_register_bra_32(g_kernels, g_rows, g_waves);
    bqc_register_bra_33(g_kernels, g_rows, g_waves);
}
static std::once_flag g_register_once;
static void register_kernels() { std::call_once(g_register_once, register_kernels_once); }


// FLOPs of the last build made by the calling thread.
static thread_local double g_last_flops[2] = {0, 0};

// Optional per-launch timing: HIP events around every kernel of a build,
// on the build's stream.  Records (class code, algorithmic FLOPs, events).
struct LaunchRec {
    int cls;       // li*1000 + lj*100 + lk*10 + ll of the evaluated class
    int kind;      // 0 quartet kernel, 1 row kernel, 4 wave kernel; +2 for FP32
    double flops;  // algorithmic FLOPs of the launch
    double nquart;
    hipEvent_t e0, e1;
};
static bool g_prof = false;
static std::mutex g_prof_mu;  // guards g_recs and g_evpool
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

// Quartets of the rectangle ij in [i0, i1) x kl in [k0, k1); in a task with
// cp_ij == cp_kl only ij >= kl counts.
static double count_rect(int i0, int i1, int k0, int k1, bool diag) {
    if (i1 <= i0 || k1 <= k0) return 0;
    if (!diag) return (double)(i1 - i0) * (k1 - k0);
    // kl < i0 meets all i1 - i0 pairs; i0 <= kl < i1 meets i1 - kl; kl >= i1 none
    const long long below = (long long)(k1 < i0 ? k1 : i0) - k0;
    double n = below > 0 ? (double)below * (i1 - i0) : 0.0;
    const long long a = k0 > i0 ? k0 : i0, b = k1 < i1 ? k1 : i1;
    if (b > a) n += (double)(b - a) * i1 - (double)(a + b - 1) * (b - a) / 2;
    return n;
}

// One task of a build, with its precision rule: a quartet whose P bin and Q
// bin upper bounds satisfy ub_P + ub_Q < log_cut is evaluated in FP32.
struct TaskCtx {
    BasisProdCache *bp;
    int cp_ij, cp_kl, nbins;
    const int *bins_ij, *bins_kl;
    int li, lj, lk, ll, pij, pkl;
    bool swap, diag;
    bool gen;            // a g or h shell: the run-time-class kernel, FP64
    double log_cut;      // -inf: FP64 only
    const double *lq;    // host log q per pair (plan order) or NULL
    double qflops;       // FLOPs of one quartet of this class
};

// log q upper bound of bin b of pair type cp (first pair of the bin in the
// q-descending plan order); +inf for an empty bin or without log q
static double bin_ub(const TaskCtx &c, int cp, const int *bins, int b) {
    if (!c.lq || bins[b + 1] <= bins[b]) return 1e300;
    return c.lq[c.bp->bas_pairs_locs[cp] + bins[b]];
}

static int task_init(TaskCtx &c, BasisProdCache *bp, const int *bins_ij, const int *bins_kl,
                     int nbins, int cp_ij, int cp_kl, const double *lq, double dmax,
                     double fp32_cut, int precision) {
    c.bp = bp;
    c.cp_ij = cp_ij;
    c.cp_kl = cp_kl;
    c.nbins = nbins;
    c.bins_ij = bins_ij;
    c.bins_kl = bins_kl;
    const ContractionProdType *tij = bp->cptype + cp_ij, *tkl = bp->cptype + cp_kl;
    c.li = tij->l_bra;
    c.lj = tij->l_ket;
    c.lk = tkl->l_bra;
    c.ll = tkl->l_ket;
    if ((c.li + c.lj + c.lk + c.ll) / 2 + 1 >= 12) return 2;
    if (c.li > BQC_LMAX_GEN || c.lj > BQC_LMAX_GEN || c.lk > BQC_LMAX_GEN ||
        c.ll > BQC_LMAX_GEN || c.lj > c.li || c.ll > c.lk)
        return 1;
    c.gen = c.li > BQC_LMAX || c.lk > BQC_LMAX;
    if (nbins > BQC_MAX_BINS) return 3;
    if (bins_ij[0] != 0 || bins_kl[0] != 0) return 1;
    c.pij = pair_id(c.li, c.lj);
    c.pkl = pair_id(c.lk, c.ll);
    c.swap = !c.gen && c.pij < c.pkl;
    c.diag = cp_ij == cp_kl;
    c.lq = lq;
    c.log_cut = (precision == 1 && lq && dmax > 0 && !c.gen) ? log(fp32_cut / dmax) : -1e300;
    c.qflops = quartet_flops(c.li, c.lj, c.lk, c.ll, tij->nprim_12 * tkl->nprim_12);
    return 0;
}

static void record_launch(LaunchRec &rec, int cls, int kind, double flops, double nq,
                          hipStream_t stream) {
    rec.cls = cls;
    rec.kind = kind;
    rec.flops = flops;
    rec.nquart = nq;
    {
        std::lock_guard<std::mutex> lock(g_prof_mu);
        rec.e0 = prof_event();
        rec.e1 = prof_event();
    }
    hipEventRecord(rec.e0, stream);
}

static void keep_launch(const LaunchRec &rec, hipStream_t stream) {
    hipEventRecord(rec.e1, stream);
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_recs.push_back(rec);
}

static int eval_cls(const TaskCtx &c) {
    return c.swap ? c.lk * 1000 + c.ll * 100 + c.li * 10 + c.lj
                  : c.li * 1000 + c.lj * 100 + c.lk * 10 + c.ll;
}

// Classes with a g or h shell (jk_general.hip).
int bqc_launch_general(const JKArgs &jk, const PairData &pd, const TaskArgs &ta, int li, int lj,
                       int lk, int ll, hipStream_t stream);

// Per-lane scratch (private segment) limit for a launch; the runtime aborts
// the process when it cannot back a kernel's scratch, so oversized kernels
// are refused here instead.
// A kernel whose static LDS exceeds the device's per-workgroup limit is
// refused the same way.  Answers are cached per kernel.
#define BQC_MAX_SCRATCH_BYTES 32768
#define BQC_OK_CACHE 1024
static bool scratch_ok(const void *kern) {
    static std::mutex mu;
    static const void *keys[BQC_OK_CACHE];
    static bool vals[BQC_OK_CACHE];
    static int nkeys = 0;
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < nkeys; ++i)
        if (keys[i] == kern) return vals[i];
    hipFuncAttributes attr;
    bool ok = hipFuncGetAttributes(&attr, kern) == hipSuccess &&
              attr.localSizeBytes <= BQC_MAX_SCRATCH_BYTES;
    int dev = 0, lds_max = 0;
    if (ok && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
            hipSuccess)
        ok = attr.sharedSizeBytes <= (size_t)lds_max;
    if (nkeys < BQC_OK_CACHE) {
        keys[nkeys] = kern;
        vals[nkeys++] = ok;
    }
    return ok;
}

// Builds through the allocation-free entry point set this (per thread):
// their launches skip the optional per-launch timing, whose records live on
// the heap.
static thread_local bool t_noalloc = false;
static bool prof_on() { return g_prof && !t_noalloc; }

// Launch shape of the row kernels: persistent grids (workgroups that stay
// resident and pull work items from per-XCD counters) or four work items per
// workgroup: 0 never, 1 for launches that hold a molecule from
// BQC_TILE_BIG_NAO Cartesian AOs up, 2 always (the default).  Set by
// BQCset_launch_mode.
static int g_persistent = 2;

// Workgroups of kern (nthr lanes) that fit on the whole device at once, from
// the kernel's registers and LDS: waves per SIMD = 512 / VGPRs (granule 8, at
// most 8), four SIMDs per CU, LDS per CU as the device reports it (at least
// its per-workgroup limit), at most 32 workgroups per CU; never fewer than
// the runtime's occupancy answer.  A persistent grid of this size keeps the
// device full; extra workgroups would only wait.  Cached per kernel and
// device in a fixed table (no heap, so the allocation-free entry may use it).
#define BQC_OCC_CACHE 1024
static unsigned resident_blocks(const void *kern, int nthr) {
    static std::mutex mu;
    static const void *keys[BQC_OCC_CACHE];
    static int devs[BQC_OCC_CACHE];
    static unsigned vals[BQC_OCC_CACHE];
    static int nkeys = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < nkeys; ++i)
        if (keys[i] == kern && devs[i] == dev) return vals[i];
    unsigned v = 0;
    hipFuncAttributes a;
    int ncu = 0, lds_cu = 0, lds_blk = 0, api = 0;
    if (hipFuncGetAttributes(&a, kern) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        ncu > 0) {
        if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  dev) != hipSuccess)
            lds_cu = 0;
        if (hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess)
            lds_blk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, kern, nthr, 0) != hipSuccess)
            api = 0;
        lds_cu = lds_cu > lds_blk ? lds_cu : lds_blk;
        const int wpb = (nthr + 63) / 64;
        const int vgpr = a.numRegs > 0 ? (a.numRegs + 7) / 8 * 8 : 512;
        const int waves_simd = 512 / vgpr < 8 ? 512 / vgpr : 8;
        int per_cu = 4 * waves_simd / wpb;
        if (a.sharedSizeBytes > 0 && lds_cu > 0) {
            const int by_lds = lds_cu / (int)a.sharedSizeBytes;
            if (by_lds < per_cu) per_cu = by_lds;
        }
        if (per_cu > 32) per_cu = 32;
        if (per_cu < api) per_cu = api;
        if (per_cu < 1) per_cu = 1;
        v = (unsigned)per_cu * (unsigned)ncu;
    }
    if (nkeys < BQC_OCC_CACHE) {
        keys[nkeys] = kern;
        devs[nkeys] = dev;
        vals[nkeys++] = v;
    }
    return v;
}

// Generic kernels (one lane per quartet): rectangle kl-bin b x ij in
// [0, bins_ij[nbins-b]) split at the first FP32 ij bin; at most
// BQC_ARG_BINS rectangles per launch.
static int launch_quartets(const TaskCtx &c, const JKArgs &jk, const PairData &pd,
                           hipStream_t stream, double *flops) {
    BasisProdCache *bp = c.bp;
    const int nbins = c.nbins;
    const int *bins_ij = c.bins_ij, *bins_kl = c.bins_kl;
    const int *locs = bp->bas_pairs_locs, *plocs = bp->primitive_pairs_locs;
    const int cp_p = c.swap ? c.cp_kl : c.cp_ij, cp_q = c.swap ? c.cp_ij : c.cp_kl;
    TaskArgs ta[2];
    double nq[2] = {0, 0};
    auto reset = [&](int prec) {
        TaskArgs &t = ta[prec];
        memset(&t, 0, sizeof(t));
        t.swap = c.swap;
        t.diag = c.diag;
        t.bas_p0 = locs[cp_p];
        t.bas_q0 = locs[cp_q];
        t.prim_p0 = plocs[cp_p];
        t.prim_q0 = plocs[cp_q];
        t.nprim_p = bp->cptype[cp_p].nprim_12;
        t.nprim_q = bp->cptype[cp_q].nprim_12;
        t.pre[0] = 0;
        nq[prec] = 0;
    };
    auto launch = [&](int prec) -> int {
        TaskArgs &t = ta[prec];
        if (t.nbins == 0) return 0;
        const double lflops = nq[prec] * c.qflops;
        flops[prec] += lflops;
        if (c.gen) {
            LaunchRec rec;
            if (prof_on()) record_launch(rec, eval_cls(c), 0, lflops, nq[prec], stream);
            int e = bqc_launch_general(jk, pd, t, c.li, c.lj, c.lk, c.ll, stream);
            if (prof_on()) keep_launch(rec, stream);
            if (e)
                fprintf(stderr, "bqcjk: launch failed for (%d%d|%d%d)\n", c.li, c.lj, c.lk,
                        c.ll);
            return e;
        }
        JKKernel kern = c.swap ? g_kernels[prec][c.pkl][c.pij] : g_kernels[prec][c.pij][c.pkl];
        if (!kern) return 1;
        if (!scratch_ok((const void *)kern)) {
            fprintf(stderr, "bqcjk: (%d%d|%d%d) kernel needs more scratch than allowed\n", c.li,
                    c.lj, c.lk, c.ll);
            return 1;
        }
        long long nblk = (t.pre[t.nbins] + BQC_THREADS - 1) / BQC_THREADS;
        if (nblk > 0x7fffffffLL) return 1;
        LaunchRec rec;
        if (prof_on()) record_launch(rec, eval_cls(c), 2 * prec, lflops, nq[prec], stream);
        hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(BQC_THREADS), 0, stream, jk, pd, t);
        if (prof_on()) keep_launch(rec, stream);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) {
            fprintf(stderr, "bqcjk: launch failed for (%d%d|%d%d): %s\n", c.li, c.lj, c.lk, c.ll,
                    hipGetErrorString(err));
            return 1;
        }
        return 0;
    };
    reset(0);
    reset(1);
    for (int b = 0; b < nbins; ++b) {
        int k0 = bins_kl[b], k1 = bins_kl[b + 1];
        int iend_bin = nbins - b;
        if (k1 <= k0 || bins_ij[iend_bin] <= 0) continue;
        double ubk = bin_ub(c, c.cp_kl, bins_kl, b);
        int split = iend_bin;
        for (int bi = 0; bi < iend_bin; ++bi)
            if (bin_ub(c, c.cp_ij, bins_ij, bi) + ubk < c.log_cut) {
                split = bi;
                break;
            }
        int ranges[2][2] = {{0, bins_ij[split]}, {bins_ij[split], bins_ij[iend_bin]}};
        for (int prec = 0; prec < 2; ++prec) {
            int i0 = ranges[prec][0], i1 = ranges[prec][1];
            if (i1 <= i0) continue;
            TaskArgs &t = ta[prec];
            int r = t.nbins;
            t.klb[r] = k0;
            t.ij0[r] = i0;
            t.nij[r] = i1 - i0;
            t.pre[r + 1] = t.pre[r] + (long long)(k1 - k0) * (i1 - i0);
            t.nbins = r + 1;
            nq[prec] += count_rect(i0, i1, k0, k1, c.diag);
            if (t.nbins == BQC_ARG_BINS) {
                if (launch(prec)) return 1;
                reset(prec);
            }
        }
    }
    for (int prec = 0; prec < 2; ++prec)
        if (launch(prec)) return 1;
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
// Row and wave kernels: host-side list orders, task descriptors and items.
// A BuildPlan collects the tasks of one or many molecules; items are grouped
// by kernel so that every (class, precision, kernel kind) is one launch.
// ---------------------------------------------------------------------------
#define BQC_WAVE_CK 32

#define BQC_CK_ROWS_DEFAULT 64
#define BQC_CK_WAVE_DEFAULT 32
static int g_ck_rows = BQC_CK_ROWS_DEFAULT;  // Q pairs per row-kernel item
static int g_ck_wave = BQC_CK_WAVE_DEFAULT;  // Q pairs per wave-kernel item
static std::vector<std::pair<long long, int>> g_ck_override;  // (key, ck)

// Default tiles by molecule size.  Autotuner timings on MI355X: for large
// molecules (64 H2O cc-pVDZ, 2624 Cartesian AOs; 128 H2O cc-pVTZ) the row
// kernels of the s/p classes are fastest with 256 Q pairs per item (fewer
// flushes of the lane accumulators; 4-6 % over 64), the spilling 108-component
// classes with 32-64 and the wave kernels with 64; for batches of small
// molecules (4-40 H2O def2-SVP, at most ~1000 AOs) with 32, 16 and 16.
#define BQC_TILE_BIG_NAO 2048
#define BQC_TILE_HUGE_NAO 8192
static bool g_ck_all_set[2] = {false, false};  // a tile set for all classes of a kind

static int class_components(int cls) {
    auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
    return nc(cls / 1000) * nc(cls / 100 % 10) * nc(cls / 10 % 10) * nc(cls % 10);
}

static int tile_for(int kind, int cls, int prec, int nao) {
    long long key = ((long long)kind * 10000 + cls) * 2 + prec;
    for (auto &kv : g_ck_override)
        if (kv.first == key) return kv.second;
    if (g_ck_all_set[kind == 1 ? 0 : 1]) return kind == 1 ? g_ck_rows : g_ck_wave;
    // three size bands, read off the autotuner on MI355X (batches of 4-40
    // waters, 64 waters cc-pVDZ, 128 waters cc-pVTZ): longer Q ranges per
    // item pay as the molecule grows
    // (batches of small molecules with persistent grids: longer ranges than
    // mid-size molecules, except (ss|ss) and (pp|pp): 256 molecules of 4-40
    // waters def2-SVP 4.65 -> 4.52 s per batch)
    const int band = nao >= BQC_TILE_HUGE_NAO ? 2 : (nao >= BQC_TILE_BIG_NAO ? 1 : 0);
    static const int wave[3] = {32, 32, 64}, rows_big[3] = {128, 64, 128},
                     rows[3] = {128, 64, 256};
    if (kind == 4) return wave[band];
    if (class_components(cls) > 81) return rows_big[band];
    if (band == 0 && cls == 0) return 32;
    if (band == 0 && cls == 1111) return 16;
    return rows[band];
}

// Orders of the pair lists of a plan by first shell (which = 0; for the
// reference-ABI entry GINTbuild_jk, whose caller has no such array) and by
// second shell (which = 1; the b-side lists of split row classes): per pair
// type the plan positions sorted by that shell, stable (a counting sort of a
// row of the host bas_pair2shls, which the cache aliases).  Built per pair
// type on first use and kept per cache until BQCdel_basis_prod_clone or
// GINTdel_basis_prod forgets it.
struct PairOrder {
    const int *pair2shls = nullptr;
    int npairs = 0, ncptype = 0;
    std::vector<int> order[2];
    std::vector<char> done[2];
};
static std::mutex g_order_mu;
static std::map<const BasisProdCache *, PairOrder> g_orders;

static const int *shell_order(const BasisProdCache *bp, int which, int cp_a, int cp_b) {
    const int *locs = bp->bas_pairs_locs;
    const int npairs = locs[bp->ncptype];
    std::lock_guard<std::mutex> lock(g_order_mu);
    PairOrder &po = g_orders[bp];
    if (po.pair2shls != bp->bas_pair2shls || po.npairs != npairs ||
        po.ncptype != bp->ncptype) {
        po.pair2shls = bp->bas_pair2shls;
        po.npairs = npairs;
        po.ncptype = bp->ncptype;
        for (int w = 0; w < 2; ++w) {
            po.order[w].assign(npairs > 0 ? npairs : 1, 0);
            po.done[w].assign(bp->ncptype, 0);
        }
    }
    const int cps[2] = {cp_a, cp_b};
    for (int cp : cps) {
        if (po.done[which][cp]) continue;
        const int p0 = locs[cp], n = locs[cp + 1] - locs[cp];
        const int *key = bp->bas_pair2shls + (which ? npairs : 0) + p0;
        int lo = 0, hi = -1;
        for (int k = 0; k < n; ++k) {
            if (k == 0 || key[k] < lo) lo = key[k];
            if (key[k] > hi) hi = key[k];
        }
        std::vector<int> cnt(n > 0 ? hi - lo + 2 : 1, 0);
        for (int k = 0; k < n; ++k) ++cnt[key[k] - lo + 1];
        for (size_t v = 1; v < cnt.size(); ++v) cnt[v] += cnt[v - 1];
        for (int k = 0; k < n; ++k) po.order[which][p0 + cnt[key[k] - lo]++] = k;
        po.done[which][cp] = 1;
    }
    return po.order[which].data();
}

extern "C" void bqc_forget_pair_order(const BasisProdCache *bp) {
    std::lock_guard<std::mutex> lock(g_order_mu);
    g_orders.erase(bp);
}

// Split row classes (a-side and b-side passes) by class code; -1 entries are
// unset (default on).  Set by BQCset_row_split.
#define BQC_SPLIT_MAX 64
static int g_split_cls[BQC_SPLIT_MAX], g_split_val[BQC_SPLIT_MAX], g_split_n = 0;
static int g_split_all = -1;  // -1: the built-in per-class rule
static bool split_on(int cls) {
    for (int i = 0; i < g_split_n; ++i)
        if (g_split_cls[i] == cls) return g_split_val[i] != 0;
    if (g_split_all >= 0) return g_split_all != 0;
    // measured per class on MI355X (64 waters cc-pVDZ, 128 waters cc-pVTZ):
    // (ps|ps) ran 20 ms faster in one pass on the larger molecule and equal
    // on the smaller, (dp|ss) 0.5 ms faster on the smaller; the other split
    // candidates gain 1-150 ms from the two passes
    return cls != 1010 && cls != 2100;
}

template <class T> using StdVec = std::vector<T>;

template <template <class> class V>
struct GroupT {
    RowKernel kern;
    int wave, cls, prec;
    double flops, nquart;
    V<int> items;     // int4 per item
    int role_b = 0;   // b-side pass of a split row class
    int nao_max = 0;  // largest molecule (Cartesian AOs) with items here
};

template <template <class> class V>
struct BuildPlanT {
    V<TaskDesc> tasks;
    V<size_t> perm_p_off, perm_q_off, bins_off;  // in ints
    V<size_t> qrec_off;                          // in int4, per task
    size_t nqrec = 0;                            // Q list records in all
    int qmax = 0;                                // longest Q list
    V<int> ints;                                 // list orders and bins
    V<GroupT<V>> groups;
    double flops[2] = {0, 0};

    GroupT<V> &group(RowKernel k, int wave, int cls, int prec) {
        for (GroupT<V> &g : groups)
            if (g.kern == k) return g;
        groups.push_back(GroupT<V>{k, wave, cls, prec, 0, 0, {}, 0, 0});
        return groups.back();
    }
};
using BuildPlan = BuildPlanT<StdVec>;

// List order of one side: plan positions [0, bins[nbins]) grouped by bin,
// inside a bin by first shell (bra_order gives the plan positions of the
// pair type sorted by first shell, then second shell).
template <class IntVec>
static void side_order(const int *bins, int nbins, const int *order, int npairs_cp,
                       IntVec &out) {
    const int n = bins[nbins];
    IntVec cursor(bins, bins + nbins);
    size_t base = out.size();
    out.resize(base + n);
    for (int k = 0; k < npairs_cp; ++k) {
        int pos = order[k];
        if (pos >= n) continue;
        int b = (int)(std::upper_bound(bins + 1, bins + nbins + 1, pos) - (bins + 1));
        out[base + cursor[b]++] = pos;
    }
}

// Adds one task to the plan.  Each precision takes the row kernel of the
// class when there is one within the scratch and LDS limits, else the wave
// kernel.  Returns -1 if a precision has neither (the generic kernels take
// the task).
template <template <class> class V>
static int plan_task(BuildPlanT<V> &bp_, const TaskCtx &c, const int *bra_order,
                     const JKArgs &jk, const PairData &pd) {
    if (!bra_order || c.gen) return -1;
    RowKernel k[2];
    int wv[2];
    bool fp32 = true;
    for (int prec = 0; prec < 2; ++prec) {
        RowKernel r = c.swap ? g_rows[0][prec][c.pkl][c.pij] : g_rows[0][prec][c.pij][c.pkl];
        RowKernel w = c.swap ? g_waves[prec][c.pkl][c.pij] : g_waves[prec][c.pij][c.pkl];
        if (r && scratch_ok((const void *)r)) {
            k[prec] = r;
            wv[prec] = 0;
        } else if (w && scratch_ok((const void *)w)) {
            k[prec] = w;
            wv[prec] = 1;
        } else if (prec == 1 && !wv[0]) {
            k[1] = k[0];
            wv[1] = 0;
            fp32 = false;
        } else {
            return -1;
        }
    }
    // a row class without an FP32 instance runs all its quartets in FP64
    const double log_cut = fp32 ? c.log_cut : -1e300;
    // split row class: a-side and b-side passes (FP64 builds of the class only)
    RowKernel ks[2] = {nullptr, nullptr};
    const int *order2 = nullptr;
    if (!wv[0] && !t_noalloc && log_cut <= -1e299 && jk.nao >= BQC_TILE_BIG_NAO &&
        split_on(eval_cls(c)) && c.bp->bas_pair2shls) {
        for (int role = 1; role <= 2; ++role)
            ks[role - 1] = c.swap ? g_rows[role][0][c.pkl][c.pij] : g_rows[role][0][c.pij][c.pkl];
        if (ks[0] && ks[1] && scratch_ok((const void *)ks[0]) && scratch_ok((const void *)ks[1])) {
            const int cp_p = c.swap ? c.cp_kl : c.cp_ij;
            order2 = shell_order(c.bp, 1, cp_p, cp_p);
            k[0] = ks[0];
        }
    }
    BasisProdCache *bp = c.bp;
    const int nbins = c.nbins;
    const int *pbins = c.swap ? c.bins_kl : c.bins_ij;
    const int *qbins = c.swap ? c.bins_ij : c.bins_kl;
    int cp_p = c.swap ? c.cp_kl : c.cp_ij, cp_q = c.swap ? c.cp_ij : c.cp_kl;
    const int *locs = bp->bas_pairs_locs;
    int np = pbins[nbins], nq = qbins[nbins];
    // FLOPs, split between precisions at bin granularity (total is exact)
    double nquart[2] = {0, 0};
    for (int bq = 0; bq < nbins; ++bq) {
        int q0 = qbins[bq], q1 = qbins[bq + 1];
        if (q1 <= q0) continue;
        double ubq = bin_ub(c, cp_q, qbins, bq);
        for (int bpn = 0; bpn < nbins - bq; ++bpn) {
            int p0 = pbins[bpn], p1 = pbins[bpn + 1];
            if (p1 <= p0) continue;
            int prec = bin_ub(c, cp_p, pbins, bpn) + ubq < log_cut ? 1 : 0;
            nquart[prec] += count_rect(p0, p1, q0, q1, c.diag);
        }
    }
    bp_.flops[0] += nquart[0] * c.qflops;
    bp_.flops[1] += nquart[1] * c.qflops;
    if (np == 0 || nq == 0) return 0;
    const int ti = (int)bp_.tasks.size();
    TaskDesc td;
    memset(&td, 0, sizeof(td));
    td.pd = pd;
    td.dm = jk.dm;
    td.vj = jk.vj;
    td.vk = jk.vk;
    td.nao = jk.nao;
    td.n_dm = jk.n_dm;
    td.np = np;
    td.bas_p0 = locs[cp_p];
    td.bas_q0 = locs[cp_q];
    td.prim_p0 = bp->primitive_pairs_locs[cp_p];
    td.prim_q0 = bp->primitive_pairs_locs[cp_q];
    td.nprim_p = bp->cptype[cp_p].nprim_12;
    td.nprim_q = bp->cptype[cp_q].nprim_12;
    td.diag = c.diag;
    td.nbins = nbins;
    bp_.tasks.push_back(td);
    bp_.bins_off.push_back(bp_.ints.size());
    bp_.ints.insert(bp_.ints.end(), pbins, pbins + nbins + 1);
    bp_.ints.insert(bp_.ints.end(), qbins, qbins + nbins + 1);
    size_t off_p = bp_.ints.size();
    side_order(pbins, nbins, bra_order + locs[cp_p], locs[cp_p + 1] - locs[cp_p], bp_.ints);
    size_t off_q = off_p;
    if (cp_p != cp_q) {
        off_q = bp_.ints.size();
        side_order(qbins, nbins, bra_order + locs[cp_q], locs[cp_q + 1] - locs[cp_q], bp_.ints);
    }
    bp_.perm_p_off.push_back(off_p);
    bp_.perm_q_off.push_back(off_q);
    bp_.qrec_off.push_back(bp_.nqrec);
    bp_.nqrec += nq;
    if (nq > bp_.qmax) bp_.qmax = nq;
    const int cls = eval_cls(c);
    GroupT<V> *g[2] = {&bp_.group(k[0], wv[0], cls, 0), nullptr};
    g[1] = &bp_.group(k[1], wv[1], cls, 1);
    g[0] = &bp_.group(k[0], wv[0], cls, 0);  // re-fetch: group() may reallocate
    for (int prec = 0; prec < 2; ++prec) {
        g[prec]->flops += nquart[prec] * c.qflops;
        g[prec]->nquart += nquart[prec];
        if (jk.nao > g[prec]->nao_max) g[prec]->nao_max = jk.nao;
    }
    // Q end of P position p's bin and the first Q position computed in FP32
    auto q_limits = [&](int p, int &qmax, int &qsplit) {
        int s = (int)(std::upper_bound(pbins + 1, pbins + nbins + 1, p) - (pbins + 1));
        qmax = qbins[nbins - s];
        double ubp = bin_ub(c, cp_p, pbins, s);
        int split = nbins - s;
        for (int bq = 0; bq < nbins - s; ++bq)
            if (ubp + bin_ub(c, cp_q, qbins, bq) < log_cut) {
                split = bq;
                break;
            }
        qsplit = qbins[split];
    };
    // When one precision runs rows of 64 P pairs, the split of a row's first
    // pair holds for every pair of the row in both precisions, so the two
    // precisions cover each pair's Q range exactly once.
    const bool by_row = !wv[0] || !wv[1];
    for (int prec = 0; prec < 2; ++prec) {
        const int ck = tile_for(wv[prec] ? 4 : 1, cls, prec, jk.nao);
        const int step = wv[prec] ? 1 : 64;
        V<int> &it = g[prec]->items;
        for (int p = 0; p < np; p += step) {
            int qmax, qsplit;
            q_limits(p, qmax, qsplit);
            if (by_row && (p & 63)) {
                int rmax;
                q_limits(p & ~63, rmax, qsplit);
            }
            const int lo = prec ? qsplit : 0;
            const int hi = prec ? qmax : (qsplit < qmax ? qsplit : qmax);
            const int pitem = wv[prec] ? p : p / 64;
            for (int q0 = lo; q0 < hi; q0 += ck) {
                int q1 = q0 + ck < hi ? q0 + ck : hi;
                it.push_back(ti);
                it.push_back(pitem);
                it.push_back(q0);
                it.push_back(q1);
            }
        }
    }
    if (order2) {
        // b-side pass: the same pairs sorted by second shell inside each bin,
        // the same Q lists, bins and work items
        const int ti2 = (int)bp_.tasks.size();
        const TaskDesc td2 = bp_.tasks[ti];
        const size_t bins2 = bp_.bins_off[ti];
        bp_.tasks.push_back(td2);
        bp_.bins_off.push_back(bins2);
        const size_t permq2 = bp_.perm_q_off[ti];
        const size_t qrec2 = bp_.qrec_off[ti];
        bp_.perm_p_off.push_back(bp_.ints.size());
        bp_.perm_q_off.push_back(permq2);
        bp_.qrec_off.push_back(qrec2);  // the same Q list and records
        side_order(pbins, nbins, order2 + locs[cp_p], locs[cp_p + 1] - locs[cp_p], bp_.ints);
        GroupT<V> &g1 = bp_.group(k[0], 0, cls, 0);
        const size_t n1 = g1.items.size();
        size_t first = n1;
        while (first >= 4 && g1.items[first - 4] == ti) first -= 4;
        GroupT<V> &g2 = bp_.group(ks[1], 0, cls, 0);
        g2.role_b = 1;
        if (jk.nao > g2.nao_max) g2.nao_max = jk.nao;
        GroupT<V> &g1b = bp_.group(k[0], 0, cls, 0);  // group() may reallocate
        for (size_t e = first; e < n1; e += 4) {
            g2.items.push_back(ti2);
            g2.items.push_back(g1b.items[e + 1]);
            g2.items.push_back(g1b.items[e + 2]);
            g2.items.push_back(g1b.items[e + 3]);
        }
    }
    return 0;
}

// Device image of a BuildPlan: [TaskDesc array][list orders][items].
template <template <class> class V>
struct StagedT {
    size_t bytes = 0, off_perm = 0, off_tickets = 0, off_qrec = 0;
    V<size_t> off_items;  // per group
};
using Staged = StagedT<StdVec>;

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

template <template <class> class V>
static StagedT<V> stage_layout(const BuildPlanT<V> &p) {
    StagedT<V> st;
    size_t off = align16(p.tasks.size() * sizeof(TaskDesc));
    st.off_perm = off;
    off = align16(off + p.ints.size() * sizeof(int));
    st.off_tickets = off;  // work counters, BQC_TICKET_SLICES per group, zero in the image
    off = align16(off + p.groups.size() * BQC_TICKET_SLICES * sizeof(int));
    st.off_qrec = off;     // Q list records, written on the device
    off = align16(off + p.nqrec * sizeof(int4));
    for (const GroupT<V> &g : p.groups) {
        st.off_items.push_back(off);
        off = align16(off + g.items.size() * sizeof(int));
    }
    st.bytes = off;
    return st;
}

// Writes the image for device address dev into host buffer host.
template <template <class> class V>
static void stage_fill(BuildPlanT<V> &p, const StagedT<V> &st, char *dev, char *host) {
    int *perm_dev = (int *)(dev + st.off_perm);
    for (size_t t = 0; t < p.tasks.size(); ++t) {
        p.tasks[t].perm_p = perm_dev + p.perm_p_off[t];
        p.tasks[t].perm_q = perm_dev + p.perm_q_off[t];
        p.tasks[t].pbins = perm_dev + p.bins_off[t];
        p.tasks[t].qbins = perm_dev + p.bins_off[t] + p.tasks[t].nbins + 1;
        p.tasks[t].qrec = (const int *)(dev + st.off_qrec) + 4 * p.qrec_off[t];
    }
    memcpy(host, p.tasks.data(), p.tasks.size() * sizeof(TaskDesc));
    memcpy(host + st.off_perm, p.ints.data(), p.ints.size() * sizeof(int));
    memset(host + st.off_tickets, 0, p.groups.size() * BQC_TICKET_SLICES * sizeof(int));
    for (size_t gi = 0; gi < p.groups.size(); ++gi)
        memcpy(host + st.off_items[gi], p.groups[gi].items.data(),
               p.groups[gi].items.size() * sizeof(int));
}

// Q list records of every task (task t = blockIdx.y): plan position, both
// shells and their AO offsets per list position, so that the J/K kernels
// read one record per step instead of a chain of three dependent loads.
__global__ void fill_qrec(const TaskDesc *tasks, int ntask) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    for (int t = blockIdx.y; t < ntask; t += gridDim.y) {
        const TaskDesc &tk = tasks[t];
        if (q >= tk.qbins[tk.nbins]) continue;
        const int qpos = tk.perm_q[q];
        const int gq = tk.bas_q0 + qpos;
        const int sc = tk.pd.bra[gq], sd = tk.pd.ket[gq];
        const unsigned c0 = (unsigned)tk.pd.ao_loc[sc], d0 = (unsigned)tk.pd.ao_loc[sd];
        ((int4 *)tk.qrec)[q] = make_int4(qpos, sc, sd, (int)(c0 | d0 << 16));
    }
}

template <template <class> class V>
static int launch_groups(const BuildPlanT<V> &p, const StagedT<V> &st, char *dev,
                         hipStream_t stream) {
    const TaskDesc *tasks = (const TaskDesc *)dev;
    if (!p.tasks.empty() && p.qmax > 0) {
        const int nt = (int)p.tasks.size();
        hipLaunchKernelGGL(fill_qrec, dim3((p.qmax + 255) / 256, nt < 65535 ? nt : 65535),
                           dim3(256), 0, stream, tasks, nt);
        if (hipGetLastError() != hipSuccess) return 1;
    }
    for (size_t gi = 0; gi < p.groups.size(); ++gi) {
        const GroupT<V> &g = p.groups[gi];
        int nitems = (int)g.items.size() / 4;
        if (!nitems) continue;
        const int4 *items = (const int4 *)(dev + st.off_items[gi]);
        unsigned nblk = g.wave ? (unsigned)nitems : (unsigned)((nitems + 3) / 4);
        unsigned nthr = g.wave ? 64 : 256;
        int *ticket = nullptr;
        // persistent grids: 64 waters cc-pVDZ 143.0 -> 136.5 ms, 256 small
        // def2-SVP molecules 5.06 -> 4.69 s per batch (with per-XCD counters)
        if (!g.wave &&
            (g_persistent == 2 || (g_persistent == 1 && g.nao_max >= BQC_TILE_BIG_NAO))) {
            // as many workgroups as stay resident at once; they drain the list
            const unsigned cap = resident_blocks((const void *)g.kern, (int)nthr);
            if (cap > 0 && cap < nblk) {
                nblk = cap;
                ticket = (int *)(dev + st.off_tickets) + gi * BQC_TICKET_SLICES;
            }
        }
        LaunchRec rec;
        if (prof_on())
            record_launch(rec, g.cls, (g.wave ? 4 : 1) + 2 * g.prec + 8 * g.role_b, g.flops,
                          g.nquart, stream);
        hipLaunchKernelGGL(g.kern, dim3(nblk), dim3(nthr), 0, stream, tasks, items, nitems,
                           ticket);
        if (prof_on()) keep_launch(rec, stream);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            fprintf(stderr, "bqcjk: kernel launch failed: %s\n", hipGetErrorString(e));
            return 1;
        }
    }
    return 0;
}

// Device images of the work lists of direct (non-graph) builds: a ring of
// BQC_STAGE_SLOTS (pinned host buffer, device buffer) pairs per stream.
// Builds on one stream run in order, so a slot's device buffer may be
// refilled by a copy queued behind the kernels of earlier builds; its host
// buffer only once the previous copy out of it has completed, which lets the
// host plan a build while the device still runs the one before (the
// reference-ABI entry GINTbuild_jk is called once per task).  Builds on
// different streams have separate rings and may overlap on the device.
// One host thread drives a given stream at a time.
#define BQC_STAGE_SLOTS 2
struct StageSlot {
    char *dws = nullptr;
    size_t dws_cap = 0;
    char *hstage = nullptr;
    size_t hstage_cap = 0;
    hipEvent_t evt = nullptr;
    bool pending = false;
};
struct StreamStage {
    StageSlot slot[BQC_STAGE_SLOTS];
    int next = 0;
};
static std::mutex g_stage_mu;
static std::map<hipStream_t, StreamStage> g_stages;

static StreamStage &stage_for(hipStream_t stream) {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    return g_stages[stream];
}

static int stage_and_launch(BuildPlan &p, hipStream_t stream) {
    if (p.groups.empty()) return 0;
    Staged st = stage_layout(p);
    StreamStage &sst = stage_for(stream);
    StageSlot &ss = sst.slot[sst.next];
    sst.next = (sst.next + 1) % BQC_STAGE_SLOTS;
    if (ss.pending) {
        // the previous copy from this slot's pinned buffer must be done
        hipEventSynchronize(ss.evt);
        ss.pending = false;
    }
    if (st.bytes > ss.hstage_cap) {
        if (ss.hstage) hipHostFree(ss.hstage);
        size_t cap = st.bytes + st.bytes / 2 + 4096;
        if (hipHostMalloc((void **)&ss.hstage, cap, hipHostMallocDefault) != hipSuccess) {
            ss.hstage = nullptr;
            ss.hstage_cap = 0;
            return 1;
        }
        ss.hstage_cap = cap;
    }
    if (st.bytes > ss.dws_cap) {
        // kernels queued earlier on this stream may still read the old image
        hipStreamSynchronize(stream);
        if (ss.dws) hipFree(ss.dws);
        size_t cap = st.bytes + st.bytes / 2 + 4096;
        if (hipMalloc((void **)&ss.dws, cap) != hipSuccess) {
            ss.dws = nullptr;
            ss.dws_cap = 0;
            return 1;
        }
        ss.dws_cap = cap;
    }
    if (!ss.evt) hipEventCreateWithFlags(&ss.evt, hipEventDisableTiming);
    stage_fill(p, st, ss.dws, ss.hstage);
    if (hipMemcpyAsync(ss.dws, ss.hstage, st.bytes, hipMemcpyHostToDevice, stream) != hipSuccess)
        return 1;
    hipEventRecord(ss.evt, stream);
    ss.pending = true;
    return launch_groups(p, st, ss.dws, stream);
}

struct MolArgs {
    BasisProdCache *bp;
    double *vj, *vk, *dm;
    int nao, n_dm, ntask;
    const int *task_cp, *task_bins, *bra_order;
    const double *log_q, *task_dmax;
};

struct Deferred {
    TaskCtx c;
    JKArgs jk;
    PairData pd;
};

// Appends plan src (tasks planned after those of dst) to dst: task indices
// in the work items, list-order, bin and record offsets move by what dst
// already holds.  Groups keep the order of their first appearance, so
// planning blocks of tasks separately and merging them in order gives the
// plan of planning them all in one pass.
static void merge_plan(BuildPlan &dst, const BuildPlan &src) {
    const int tbase = (int)dst.tasks.size();
    const size_t ibase = dst.ints.size(), qbase = dst.nqrec;
    dst.tasks.insert(dst.tasks.end(), src.tasks.begin(), src.tasks.end());
    for (size_t x : src.perm_p_off) dst.perm_p_off.push_back(x + ibase);
    for (size_t x : src.perm_q_off) dst.perm_q_off.push_back(x + ibase);
    for (size_t x : src.bins_off) dst.bins_off.push_back(x + ibase);
    for (size_t x : src.qrec_off) dst.qrec_off.push_back(x + qbase);
    dst.nqrec += src.nqrec;
    dst.qmax = std::max(dst.qmax, src.qmax);
    dst.ints.insert(dst.ints.end(), src.ints.begin(), src.ints.end());
    dst.flops[0] += src.flops[0];
    dst.flops[1] += src.flops[1];
    for (const GroupT<StdVec> &g : src.groups) {
        GroupT<StdVec> &d = dst.group(g.kern, g.wave, g.cls, g.prec);
        d.flops += g.flops;
        d.nquart += g.nquart;
        d.role_b |= g.role_b;
        d.nao_max = std::max(d.nao_max, g.nao_max);
        const size_t n0 = d.items.size();
        d.items.insert(d.items.end(), g.items.begin(), g.items.end());
        for (size_t e = n0; e < d.items.size(); e += 4) d.items[e] += tbase;
    }
}

// Host threads planning one list of tasks: at most BQCset_plan_threads'
// value, else BQC_PLAN_THREADS, else the hardware threads up to 16; one per
// 16 tasks.
static std::atomic<int> g_plan_threads(0);
static int plan_threads(size_t ntask) {
    static const int dflt = [] {
        const char *e = getenv("BQC_PLAN_THREADS");
        int n = e ? atoi(e) : (int)std::thread::hardware_concurrency();
        if (!e && n > 16) n = 16;
        if (n > 256) n = 256;
        return n < 1 ? 1 : n;
    }();
    const int set = g_plan_threads.load();
    const int cap = set > 0 ? set : dflt;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)cap, ntask / 16));
}

// One task of a build: molecule and task index.
struct TaskRef {
    int im, t;
};

static int task_ctx(TaskCtx &c, const MolArgs &a, int t, int nbins, int precision,
                    double fp32_cut) {
    const int *bi = a.task_bins + (size_t)t * 2 * (nbins + 1);
    const int *bk = bi + nbins + 1;
    return task_init(c, a.bp, bi, bk, nbins, a.task_cp[2 * t], a.task_cp[2 * t + 1], a.log_q,
                     a.task_dmax ? a.task_dmax[t] : 0, fp32_cut, precision);
}

// Plans tasks list[0, n) into plan; generic-kernel tasks go to *deferred.
static int plan_refs(BuildPlan &plan, const MolArgs *m, const TaskRef *list, size_t n,
                     int nbins, int precision, double fp32_cut, std::vector<Deferred> *deferred) {
    for (size_t i = 0; i < n; ++i) {
        const MolArgs &a = m[list[i].im];
        TaskCtx c;
        int err = task_ctx(c, a, list[i].t, nbins, precision, fp32_cut);
        if (err) return err;
        JKArgs jk = {a.vj, a.vk, a.dm, a.nao, a.n_dm};
        PairData pd = make_pair_data(a.bp);
        err = plan_task(plan, c, a.bra_order, jk, pd);
        if (err < 0) {
            deferred->push_back(Deferred{c, jk, pd});
            err = 0;
        }
        if (err > 0) return err;
    }
    return 0;
}

// Plans tasks list[0, n), with several host threads when there are enough:
// blocks of tasks handed out in order are planned into plans of their own
// and merged in block order (the same plan as one pass over the list).
// Generic-kernel tasks are appended to deferred in list order.
static int plan_list(BuildPlan &plan, const MolArgs *m, const TaskRef *list, size_t n,
                     int nbins, int precision, double fp32_cut, std::vector<Deferred> &deferred) {
    const int nth = plan_threads(n);
    if (nth < 2) return plan_refs(plan, m, list, n, nbins, precision, fp32_cut, &deferred);
    const size_t bsz = std::max<size_t>(4, n / ((size_t)nth * 8));
    const size_t nblk = (n + bsz - 1) / bsz;
    std::vector<BuildPlan> part(nblk);
    std::vector<std::vector<Deferred>> part_def(nblk);
    std::vector<int> part_err(nblk, 0);
    std::atomic<size_t> next(0);
    // planning may query kernel attributes (scratch and LDS limits): the
    // workers use the calling thread's device, not a new thread's device 0
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    auto work = [&]() {
        (void)hipSetDevice(dev);
        for (size_t b; (b = next.fetch_add(1)) < nblk;) {
            const size_t t0 = b * bsz, t1 = std::min(n, t0 + bsz);
            part_err[b] = plan_refs(part[b], m, list + t0, t1 - t0, nbins, precision, fp32_cut,
                                    &part_def[b]);
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nth; ++i) th.emplace_back(work);
    work();
    for (std::thread &t : th) t.join();
    for (size_t b = 0; b < nblk; ++b) {
        if (part_err[b]) return part_err[b];
        merge_plan(plan, part[b]);
        deferred.insert(deferred.end(), part_def[b].begin(), part_def[b].end());
    }
    return 0;
}

static int check_build_args(const MolArgs *m, int nmol, int nbins, int precision) {
    if (precision != 0 && precision != 1) return 1;
    if (nbins > BQC_MAX_BINS) return 3;
    for (int im = 0; im < nmol; ++im)
        if (m[im].nao > BQC_MAX_NAO) return 4;
    return 0;
}

// Direct builds: tasks grouped by integral class (classes in order of first
// appearance, tasks in order inside a class), and each class staged and
// launched as soon as it is planned, so the device runs the first classes
// while the host plans the rest instead of waiting for the whole plan
// (64 waters cc-pVDZ: 5.4 ms of planning ahead of the first launch).
static int build_by_class(const MolArgs *m, int nmol, int nbins, int precision, double fp32_cut,
                          hipStream_t stream, double flops[2]) {
    flops[0] = flops[1] = 0;
    int err = check_build_args(m, nmol, nbins, precision);
    if (err) return err;
    register_kernels();
    std::vector<int> keys;
    std::vector<std::vector<TaskRef>> by_cls;
    for (int im = 0; im < nmol; ++im)
        for (int t = 0; t < m[im].ntask; ++t) {
            TaskCtx c;
            err = task_ctx(c, m[im], t, nbins, precision, fp32_cut);
            if (err) return err;
            const int key = eval_cls(c);
            size_t k = 0;
            while (k < keys.size() && keys[k] != key) ++k;
            if (k == keys.size()) {
                keys.push_back(key);
                by_cls.emplace_back();
            }
            by_cls[k].push_back(TaskRef{im, t});
        }
    for (const std::vector<TaskRef> &list : by_cls) {
        BuildPlan plan;
        std::vector<Deferred> def;
        err = plan_list(plan, m, list.data(), list.size(), nbins, precision, fp32_cut, def);
        for (Deferred &d : def)
            if (!err) err = launch_quartets(d.c, d.jk, d.pd, stream, plan.flops);
        if (!err) err = stage_and_launch(plan, stream);
        flops[0] += plan.flops[0];
        flops[1] += plan.flops[1];
        if (err) return err;
    }
    return 0;
}

// Plans all tasks of all molecules.  Classes without row/wave kernels go to
// the generic kernels: launched right away, or queued in *deferred.
template <template <class> class V>
static int plan_molecules(BuildPlanT<V> &plan, const MolArgs *m, int nmol, int nbins,
                          int precision, double fp32_cut, hipStream_t stream,
                          std::vector<Deferred> *deferred = nullptr) {
    const int cerr = check_build_args(m, nmol, nbins, precision);
    if (cerr) return cerr;
    register_kernels();
    if constexpr (std::is_same<BuildPlanT<V>, BuildPlan>::value) {
        std::vector<TaskRef> list;
        for (int im = 0; im < nmol; ++im)
            for (int t = 0; t < m[im].ntask; ++t) list.push_back(TaskRef{im, t});
        std::vector<Deferred> def;
        const int err = plan_list(plan, m, list.data(), list.size(), nbins, precision, fp32_cut,
                                  def);
        if (err) return err;
        for (Deferred &d : def) {
            if (deferred) {
                deferred->push_back(d);
            } else {
                const int e = launch_quartets(d.c, d.jk, d.pd, stream, plan.flops);
                if (e) return e;
            }
        }
        return 0;
    }
    for (int im = 0; im < nmol; ++im) {
        const MolArgs &a = m[im];
        JKArgs jk = {a.vj, a.vk, a.dm, a.nao, a.n_dm};
        PairData pd = make_pair_data(a.bp);
        for (int t = 0; t < a.ntask; ++t) {
            const int *bi = a.task_bins + (size_t)t * 2 * (nbins + 1);
            const int *bk = bi + nbins + 1;
            TaskCtx c;
            int err = task_init(c, a.bp, bi, bk, nbins, a.task_cp[2 * t], a.task_cp[2 * t + 1],
                                a.log_q, a.task_dmax ? a.task_dmax[t] : 0, fp32_cut, precision);
            if (err) return err;
            err = plan_task(plan, c, a.bra_order, jk, pd);
            if (err < 0 && deferred) {
                deferred->push_back(Deferred{c, jk, pd});
                err = 0;
            }
            if (err < 0) err = launch_quartets(c, jk, pd, stream, plan.flops);
            if (err > 0) return err;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free build (BQCbuild_jk_noalloc).  The device work list lives in
// the file-scope array bqc_noalloc_scratch (each GPU the library runs on
// holds its own copy until the process ends), the host side of the plan in a
// static arena; neither the device nor the host heap is touched.  The
// array's host-side name is not a device address: the device address of
// each GPU's copy comes from hipGetSymbolAddress, once per device, and only
// that address is handed to copies and kernels.
// ---------------------------------------------------------------------------
#define BQC_NOALLOC_BYTES (64ull << 20)        // device work list, per GPU
#define BQC_NOALLOC_ARENA (256ull << 20)       // host plan arena (growth slack)
#define BQC_NOALLOC_DEVICES 64
static __device__ __align__(256) char bqc_noalloc_scratch[BQC_NOALLOC_BYTES];

struct HostArena {
    alignas(256) char mem[BQC_NOALLOC_ARENA];
    size_t off = 0;
};
static HostArena g_arena;  // guarded by g_noalloc_mu
// Host copy of the work list handed to the device copy (pageable); refilled
// only after the previous copy out of it has completed.
alignas(256) static char g_noalloc_host[BQC_NOALLOC_BYTES];
static std::mutex g_noalloc_mu;
static char *g_noalloc_dev[BQC_NOALLOC_DEVICES];
static hipEvent_t g_noalloc_copied = nullptr;
static bool g_noalloc_pending = false;

struct ArenaFull {};
template <class T>
struct ArenaAlloc {
    using value_type = T;
    ArenaAlloc() = default;
    template <class U>
    ArenaAlloc(const ArenaAlloc<U> &) {}
    T *allocate(size_t n) {
        const size_t off = (g_arena.off + 15) & ~(size_t)15;
        if (n > (BQC_NOALLOC_ARENA - off) / sizeof(T)) throw ArenaFull();
        g_arena.off = off + n * sizeof(T);
        return (T *)(g_arena.mem + off);
    }
    void deallocate(T *, size_t) {}
    template <class U>
    bool operator==(const ArenaAlloc<U> &) const { return true; }
    template <class U>
    bool operator!=(const ArenaAlloc<U> &) const { return false; }
};
template <class T> using ArenaVec = std::vector<T, ArenaAlloc<T>>;

// Device address of this GPU's copy of the scratch array (set-up on the
// first call on each device, with that device current).
static char *noalloc_device_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= BQC_NOALLOC_DEVICES)
        return nullptr;
    if (!g_noalloc_dev[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(bqc_noalloc_scratch)) != hipSuccess || !ptr) {
            fprintf(stderr, "bqcjk: hipGetSymbolAddress of the no-alloc scratch failed\n");
            return nullptr;
        }
        g_noalloc_dev[dev] = (char *)ptr;
    }
    return g_noalloc_dev[dev];
}


extern "C" {

int GINTbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                 int n_dm, int *bins_locs_ij, int *bins_locs_kl, int nbins, int cp_ij_id,
                 int cp_kl_id, hipStream_t stream) {
    register_kernels();
    if (nao > BQC_MAX_NAO) return 4;
    if (nbins < 1 || cp_ij_id < 0 || cp_kl_id < 0 || cp_ij_id >= bpcache->ncptype ||
        cp_kl_id >= bpcache->ncptype)
        return 1;
    JKArgs jk = {vj, vk, dm, nao, n_dm};
    PairData pd = make_pair_data(bpcache);
    TaskCtx c;
    int err = task_init(c, bpcache, bins_locs_ij, bins_locs_kl, nbins, cp_ij_id, cp_kl_id,
                        nullptr, 0, 0, 0);
    if (err) return err;
    const int *order = bpcache->bas_pair2shls
                           ? shell_order(bpcache, 0, cp_ij_id, cp_kl_id)
                           : nullptr;
    BuildPlan plan;
    err = plan_task(plan, c, order, jk, pd);
    if (err < 0) err = launch_quartets(c, jk, pd, stream, plan.flops);
    else if (err == 0) err = stage_and_launch(plan, stream);
    g_last_flops[0] = plan.flops[0];
    g_last_flops[1] = plan.flops[1];
    return err;
}

int BQCbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                int n_dm, int ntask, const int *task_cp, const int *task_bins, int nbins,
                const int *bra_order, int precision, const double *log_q,
                const double *task_dmax, double fp32_cut, hipStream_t stream) {
    MolArgs m = {bpcache, vj, vk, dm, nao, n_dm, ntask, task_cp, task_bins, bra_order,
                 log_q, task_dmax};
    double flops[2];
    const int err = build_by_class(&m, 1, nbins, precision, fp32_cut, stream, flops);
    g_last_flops[0] = flops[0];
    g_last_flops[1] = flops[1];
    return err;
}

int BQCbuild_jk_batch(int nmol, BasisProdCache **bps, double **vj, double **vk, double **dm,
                      const int *nao, int n_dm, const int *ntask, const int **task_cp,
                      const int **task_bins, int nbins, const int **bra_order, int precision,
                      const double **log_q, const double **task_dmax, double fp32_cut,
                      hipStream_t stream) {
    std::vector<MolArgs> m(nmol);
    for (int i = 0; i < nmol; ++i)
        m[i] = MolArgs{bps[i], vj ? vj[i] : nullptr, vk ? vk[i] : nullptr, dm[i], nao[i], n_dm,
                       ntask[i], task_cp[i], task_bins[i], bra_order[i],
                       log_q ? log_q[i] : nullptr, task_dmax ? task_dmax[i] : nullptr};
    double flops[2];
    const int err = build_by_class(m.data(), nmol, nbins, precision, fp32_cut, stream, flops);
    g_last_flops[0] = flops[0];
    g_last_flops[1] = flops[1];
    return err;
}

int BQCbuild_jk_noalloc(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                        int n_dm, int ntask, const int *task_cp, const int *task_bins,
                        int nbins, const int *bra_order, int precision, const double *log_q,
                        const double *task_dmax, double fp32_cut, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    char *dev = noalloc_device_ptr();
    if (!dev) return 1;
    if (!g_noalloc_copied &&
        hipEventCreateWithFlags(&g_noalloc_copied, hipEventDisableTiming) != hipSuccess)
        return 1;
    MolArgs m = {bpcache, vj, vk, dm, nao, n_dm, ntask, task_cp, task_bins, bra_order,
                 log_q, task_dmax};
    g_arena.off = 0;
    t_noalloc = true;
    int err = 0;
    try {
        BuildPlanT<ArenaVec> plan;
        err = plan_molecules(plan, &m, 1, nbins, precision, fp32_cut, stream);
        g_last_flops[0] = plan.flops[0];
        g_last_flops[1] = plan.flops[1];
        if (!err && !plan.groups.empty()) {
            StagedT<ArenaVec> st = stage_layout(plan);
            if (st.bytes > BQC_NOALLOC_BYTES) {
                fprintf(stderr,
                        "bqcjk: work list of %zu bytes exceeds the %llu-byte scratch of "
                        "BQCbuild_jk_noalloc\n",
                        st.bytes, (unsigned long long)BQC_NOALLOC_BYTES);
                err = 5;
            } else {
                if (g_noalloc_pending) {
                    hipEventSynchronize(g_noalloc_copied);
                    g_noalloc_pending = false;
                }
                stage_fill(plan, st, dev, g_noalloc_host);
                if (hipMemcpyAsync(dev, g_noalloc_host, st.bytes, hipMemcpyHostToDevice,
                                   stream) != hipSuccess)
                    err = 1;
                if (!err) {
                    hipEventRecord(g_noalloc_copied, stream);
                    g_noalloc_pending = true;
                    err = launch_groups(plan, st, dev, stream);
                }
            }
        }
    } catch (const ArenaFull &) {
        fprintf(stderr, "bqcjk: host plan exceeds the %llu-byte arena of BQCbuild_jk_noalloc\n",
                (unsigned long long)BQC_NOALLOC_ARENA);
        err = 5;
    }
    t_noalloc = false;
    return err;
}

__global__ void zero_ints(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// A captured batch: its own device image and the instantiated graph.
struct BQCGraph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    char *dev;
    double flops[2];
};

void *BQCgraph_capture(int nmol, BasisProdCache **bps, double **vj, double **vk, double **dm,
                       const int *nao, int n_dm, const int *ntask, const int **task_cp,
                       const int **task_bins, int nbins, const int **bra_order, int precision,
                       const double **log_q, const double **task_dmax, double fp32_cut,
                       hipStream_t stream) {
    std::vector<MolArgs> m(nmol);
    for (int i = 0; i < nmol; ++i)
        m[i] = MolArgs{bps[i], vj ? vj[i] : nullptr, vk ? vk[i] : nullptr, dm[i], nao[i], n_dm,
                       ntask[i], task_cp[i], task_bins[i], bra_order[i],
                       log_q ? log_q[i] : nullptr, task_dmax ? task_dmax[i] : nullptr};
    BuildPlan plan;
    std::vector<Deferred> deferred;
    bool was_prof = g_prof;
    g_prof = false;
    int err = plan_molecules(plan, m.data(), nmol, nbins, precision, fp32_cut, stream, &deferred);
    BQCGraph *g = new BQCGraph();
    g->dev = nullptr;
    g->graph = nullptr;
    Staged st;
    if (!err && !plan.groups.empty()) {
        st = stage_layout(plan);
        std::vector<char> host(st.bytes);
        if (hipMalloc((void **)&g->dev, st.bytes) != hipSuccess) err = 1;
        if (!err) {
            stage_fill(plan, st, g->dev, host.data());
            if (hipMemcpy(g->dev, host.data(), st.bytes, hipMemcpyHostToDevice) != hipSuccess)
                err = 1;
        }
    }
    hipStream_t cs = nullptr;
    if (!err && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) err = 1;
    if (!err && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) err = 1;
    if (!err) {
        double fl[2] = {0, 0};
        for (Deferred &d : deferred)
            if (!err) err = launch_quartets(d.c, d.jk, d.pd, cs, fl);
        plan.flops[0] += fl[0];
        plan.flops[1] += fl[1];
        if (!err && !plan.groups.empty()) {
            // replays start from zero work counters (a kernel, not a memset)
            const int ng = (int)plan.groups.size() * BQC_TICKET_SLICES;
            hipLaunchKernelGGL(zero_ints, dim3((ng + 255) / 256), dim3(256), 0, cs,
                               (int *)(g->dev + st.off_tickets), ng);
            err = launch_groups(plan, st, g->dev, cs);
        }
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(cs, &graph);
        if (ce != hipSuccess || !graph) err = 1;
        g->graph = graph;
    }
    if (!err && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) err = 1;
    g_prof = was_prof;
    if (cs) hipStreamDestroy(cs);
    if (err) {
        if (g->graph) hipGraphDestroy(g->graph);
        if (g->dev) hipFree(g->dev);
        delete g;
        return nullptr;
    }
    g->flops[0] = plan.flops[0];
    g->flops[1] = plan.flops[1];
    return g;
}

int BQCgraph_launch(void *handle, hipStream_t stream) {
    BQCGraph *g = (BQCGraph *)handle;
    if (!g) return 1;
    g_last_flops[0] = g->flops[0];
    g_last_flops[1] = g->flops[1];
    return hipGraphLaunch(g->exec, stream) == hipSuccess ? 0 : 1;
}

void BQCgraph_destroy(void *handle) {
    BQCGraph *g = (BQCGraph *)handle;
    if (!g) return;
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    if (g->dev) hipFree(g->dev);
    delete g;
}

// Q pairs per work item: kind 1 row kernels, 4 wave kernels; cls = class code
// li*1000+lj*100+lk*10+ll as evaluated, or -1 for the kind's default; kind 0
// drops every override and restores the defaults.
int BQCset_tile(int kind, int cls, int prec, int ck) {
    if (kind == 0) {
        g_ck_override.clear();
        g_ck_rows = BQC_CK_ROWS_DEFAULT;
        g_ck_wave = BQC_CK_WAVE_DEFAULT;
        g_ck_all_set[0] = g_ck_all_set[1] = false;
        return 0;
    }
    if (ck < 1 || ck > 4096 || (kind != 1 && kind != 4)) return 1;
    if (cls < 0) {
        if (kind == 1) g_ck_rows = ck;
        else g_ck_wave = ck;
        g_ck_all_set[kind == 1 ? 0 : 1] = true;
        return 0;
    }
    long long key = ((long long)kind * 10000 + cls) * 2 + prec;
    for (auto &kv : g_ck_override)
        if (kv.first == key) {
            kv.second = ck;
            return 0;
        }
    g_ck_override.push_back({key, ck});
    return 0;
}

int BQCset_row_split(int cls, int on) {
    if (on != 0 && on != 1 && !(cls < 0 && on == -1)) return 1;
    if (cls < 0) {
        g_split_all = on;
        g_split_n = 0;
        return 0;
    }
    for (int i = 0; i < g_split_n; ++i)
        if (g_split_cls[i] == cls) {
            g_split_val[i] = on;
            return 0;
        }
    if (g_split_n >= BQC_SPLIT_MAX) return 1;
    g_split_cls[g_split_n] = cls;
    g_split_val[g_split_n++] = on;
    return 0;
}

int BQCset_launch_mode(int persistent) {
    if (persistent < 0 || persistent > 2) return 1;
    g_persistent = persistent;
    return 0;
}

int BQCset_plan_threads(int n) {
    if (n < 0 || n > 256) return 1;
    g_plan_threads.store(n);
    return 0;
}

int BQCget_tile(int kind, int cls, int prec, int nao) {
    if (kind != 1 && kind != 4) return -1;
    return tile_for(kind, cls, prec, nao);
}

void BQCprofile_enable(int on) {
    g_prof = on != 0;
}

int BQCprofile_count(void) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    return (int)g_recs.size();
}

// Drains the launch records: cls, kind, flops, quartets and duration (ms)
// per launch; waits for the recorded events.  Returns the number written.
int BQCprofile_read(int *cls, int *kind, double *flops, double *nquart, double *ms, int cap) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
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
