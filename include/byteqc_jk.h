/*
 * C-ABI of libbqcjk.so -- MI355X (gfx950) direct-SCF J/K engine.
 *
 * The first block of entry points keeps the names, argument meaning and
 * return codes of the reference's native interface for this path, so that
 * the reference's host code (cuobc/scf/hf.py, which binds them through
 * ctypes) can load this library in place of libgint.so / libgvhf.so.
 * The second block holds entry points of this library that have no
 * counterpart there.
 *
 * Conventions shared by all entry points:
 *   - plain pointers and sizes; "device" pointers are HIP device memory,
 *     "host" pointers are ordinary host memory;
 *   - all J/K output is accumulated (+=); the caller zeroes it;
 *   - work is queued on the given HIP stream and returns before it ends;
 *   - return 0 on success, non-zero on error (see each entry).
 */
#ifndef BYTEQC_JK_H
#define BYTEQC_JK_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* (l_bra, l_ket, nprim_bra*nprim_ket, number of pairs) of one pair type.
 * Same layout as the reference (cuobc/lib/gint/gint.h:110-116). */
typedef struct {
    int l_bra;
    int l_ket;
    int nprim_12;
    int npairs;
} ContractionProdType;

/* Shell-pair cache.  Same field order and meaning as the reference
 * (cuobc/lib/gint/gint.h:127-147; ctypes mirror cuobc/scf/hf.py:460-478):
 * the first seven fields are host data, the rest device data once the
 * struct has been cloned by GINTinit_basis_prod_gpu. */
typedef struct {
    int nbas;
    int ncptype;
    ContractionProdType *cptype;
    int *bas_pairs_locs;
    int *primitive_pairs_locs;
    int *bas_pair2shls;
    double *aexyz;
    double *bas_coords;
    int *bas_pair2bra;
    int *bas_pair2ket;
    int *ao_loc;
    double *a12;
    double *e12;
    double *x12;
    double *y12;
    double *z12;
} BasisProdCache;

/* ---- entry points with a reference counterpart ------------------------- */

/* Replaces GINTinit_basis_prod_cpu (cuobc/lib/gint/nr_fill_ao_ints.cu:452-470).
 * Allocates *pbp (host) with the pair types, primitive-pair offsets and the
 * primitive-pair table aexyz = [a12 | e12 | x12 | y12 | z12]; e12 of pairs
 * with ish == jsh is scaled by diag_fac.  Keeps (aliases) the caller's
 * bas_pair2shls (2 x npairs, bra row then ket row) and bas_pairs_locs
 * (ncptype+1), which must outlive the cache.  Fills bas_coords (3 x nbas,
 * x row, y row, z row) from atm/env. */
void GINTinit_basis_prod_cpu(BasisProdCache **pbp, double diag_fac, double *bas_coords,
                             int *bas_pair2shls, int *bas_pairs_locs, int ncptype,
                             int *atm, int natm, int *bas, int nbas, double *env);

/* Replaces GINTinit_basis_prod_gpu (cuobc/lib/gint/nr_fill_ao_ints.cu:472-491).
 * *out becomes a host-side clone of *in whose device fields point into the
 * caller-owned device arrays ao_loc, bas_coords, aexyz, bas_pair2shls. */
void GINTinit_basis_prod_gpu(BasisProdCache **in, BasisProdCache **out, int *ao_loc,
                             double *bas_coords, double *aexyz, int *bas_pair2shls);

/* Replaces GINTdel_basis_prod (cuobc/lib/gint/nr_fill_ao_ints.cu:394-410).
 * Frees what GINTinit_basis_prod_cpu allocated (the aliased caller arrays
 * are not freed).  Clones made by GINTinit_basis_prod_gpu are freed with
 * BQCdel_basis_prod_clone. */
void GINTdel_basis_prod(BasisProdCache **pbp);

/* Replaces GINTbuild_jk (cuobc/lib/gvhf/nr_jk_driver.cu:227-283).
 * One task: pair type cp_ij_id with pair type cp_kl_id (cp_ij_id >= cp_kl_id).
 * bins_locs_ij / bins_locs_kl (host, nbins+1) split each pair list into
 * Schwarz bins; kl pairs of bin b meet ij pairs of bins [0, nbins-b).
 * vj / vk / dm: device, n_dm x nao x nao (vj or vk may be NULL).
 * The quartets run in the same row and wavefront kernels as BQCbuild_jk;
 * the pair lists are walked grouped by first shell, an order built from the
 * host bra row of bas_pair2shls (aliased by the cache) on first use of each
 * pair type and kept until the cache is deleted.  Calls on one stream are
 * ordered; the host may plan the next task while the device runs this one.
 * Returns 0 ok, 1 launch error or unsupported class, 2 nroots >= 12; codes
 * this library adds: 3 more than 4096 bins, 4 nao above 46340 (element
 * offsets are 32-bit). */
int GINTbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                 int n_dm, int *bins_locs_ij, int *bins_locs_kl, int nbins, int cp_ij_id,
                 int cp_kl_id, hipStream_t stream);

/* ---- entry points of this library only ---------------------------------- */

/* Four-index ERIs of the screened quartets of pair types cp_ij_id (bra)
 * and cp_kl_id (ket), added to eri (device) at
 *   (i - ao_offsets[0]) strides[0] + (j - ao_offsets[1]) strides[1]
 *   + (k - ao_offsets[2]) strides[2] + (l - ao_offsets[3]) strides[3]
 * for the Cartesian AOs i, j (pair bra, ket) and k, l of every quartet; kl
 * bin b meets the ij pairs of bins [0, nbins - b) (host bin offsets as for
 * GINTbuild_jk).  Same name, arguments and return codes (0 ok, 1 launch
 * error / unsupported class, 2 nroots >= 12) as the reference
 * (cuobc/lib/gint/nr_fill_ao_ints.cu:493-550). */
int GINTfill_int2e(BasisProdCache *bpcache, double *eri, int nao, const int *strides,
                   const int *ao_offsets, const int *bins_locs_ij, const int *bins_locs_kl,
                   int nbins, int cp_ij_id, int cp_kl_id, hipStream_t stream);

/* Frees a clone made by GINTinit_basis_prod_gpu (the reference leaks it). */
void BQCdel_basis_prod_clone(BasisProdCache **pbp);

/* All tasks of one J/K build in one call.  task_cp (host, 2 x ntask):
 * (cp_ij, cp_kl) per task; task_bins (host, ntask x 2 x (nbins+1)): ij bins
 * then kl bins per task, as for GINTbuild_jk.  bra_order (host, one entry
 * per pair, or NULL): for every pair type, the positions of its pairs in
 * the pair list ordered by first shell (stable); it lets the kernels walk
 * each Schwarz bin grouped by first shell.  The library keeps its device
 * work lists per HIP stream: builds on one stream are ordered, builds on
 * different streams may overlap on the device, one host thread drives a
 * stream at a time, and BQCset_tile must not run concurrently with a build.
 * precision: 0 = FP64
 * throughout; 1 = mixed: a quartet is evaluated in FP32 (Rys roots and
 * weights, prefactors and all J/K accumulation stay FP64) when the upper
 * bound q_P q_Q dmax of its Schwarz bins, taken over the 64 pairs sharing a
 * wavefront, falls below fp32_cut.  log_q (host, one per pair, plan order)
 * gives log q; task_dmax (host, one per task) the task's density bound.
 * log_q is used for this split only (the bins come from task_bins), so it
 * may hold the unclipped log q of every pair.
 * Returns 0 ok, 1 launch error / unsupported class, 2 nroots >= 12,
 * 3 more than 4096 bins, 4 nao above 46340. */
int BQCbuild_jk(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                int n_dm, int ntask, const int *task_cp, const int *task_bins, int nbins,
                const int *bra_order, int precision, const double *log_q,
                const double *task_dmax, double fp32_cut, hipStream_t stream);

/* BQCbuild_jk without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, work and return
 * codes as BQCbuild_jk, plus 5 when the build's device work list exceeds the
 * 64 MiB scratch array this library keeps per GPU (or its host plan the
 * 256 MiB host arena).  The work list goes to a file-scope device array
 * (one per GPU, held until the process ends) instead of memory from
 * hipMalloc/hipHostMalloc; no device or host heap allocation is made.
 * Every call on a GPU shares that GPU's array: calls on one stream are
 * ordered, calls that overlap on one GPU (other streams, other threads)
 * race on it.  Calls are serialised on the host. */
int BQCbuild_jk_noalloc(BasisProdCache *bpcache, double *vj, double *vk, double *dm, int nao,
                        int n_dm, int ntask, const int *task_cp, const int *task_bins,
                        int nbins, const int *bra_order, int precision, const double *log_q,
                        const double *task_dmax, double fp32_cut, hipStream_t stream);

/* The same build for many independent molecules in one set of launches
 * (one launch per integral class and kernel kind across all molecules).
 * Arrays of length nmol hold each molecule's plan, device J/K/density
 * pointers (vj or vk may be NULL as a whole), nao, task list and first-shell
 * order exactly as for BQCbuild_jk; log_q / task_dmax may be NULL for FP64.
 * Returns as BQCbuild_jk. */
int BQCbuild_jk_batch(int nmol, BasisProdCache **bps, double **vj, double **vk, double **dm,
                      const int *nao, int n_dm, const int *ntask, const int **task_cp,
                      const int **task_bins, int nbins, const int **bra_order, int precision,
                      const double **log_q, const double **task_dmax, double fp32_cut,
                      hipStream_t stream);

/* Captures the launches of a BQCbuild_jk_batch call into a HIP graph with
 * its own device work list (the density and J/K buffers passed here are the
 * ones the graph reads and accumulates into on every replay; the caller
 * zeroes J/K before a replay with a kernel, not a memset).  Returns an
 * opaque handle, NULL on error. */
void *BQCgraph_capture(int nmol, BasisProdCache **bps, double **vj, double **vk, double **dm,
                       const int *nao, int n_dm, const int *ntask, const int **task_cp,
                       const int **task_bins, int nbins, const int **bra_order, int precision,
                       const double **log_q, const double **task_dmax, double fp32_cut,
                       hipStream_t stream);
int BQCgraph_launch(void *graph, hipStream_t stream);
void BQCgraph_destroy(void *graph);

/* Q pairs per work item of the row kernels (kind 1) or wave kernels
 * (kind 4), for class code cls (li*1000+lj*100+lk*10+ll as evaluated) and
 * precision prec (0 FP64, 1 FP32), or for all classes of the kind (cls < 0);
 * kind 0 clears these settings and restores the default tiles, which depend
 * on the molecule's Cartesian AO count (read off the autotuner on MI355X:
 * rows / rows of classes above 81 components / waves 128 / 128 / 32 below
 * 2048 AOs (32 for (ss|ss), 16 for (pp|pp)), 64 / 64 / 32 from 2048,
 * 256 / 128 / 64 from 8192; 64 waters cc-pVDZ, 2624 AOs, ran 6 ms faster on
 * 64 than on 256 Q pairs per row item).  Precedence when a build is
 * planned: per-class setting,
 * all-class setting, default.  Process-wide; used by the tile autotuner;
 * not to be called while another thread plans a build.
 * No counterpart in the reference (its launch shapes are fixed per class,
 * nr_jk_driver.cu:46-225).  Returns 0 ok, 1 bad argument. */
int BQCset_tile(int kind, int cls, int prec, int ck);

/* Whether the row-kernel class cls (code li*1000+lj*100+lk*10+ll as
 * evaluated; -1: all classes, clearing per-class settings) runs as two
 * passes in FP64 builds: an a-side pass (J(a,b), J(c,d), K(a,c), K(a,d) over
 * pair lists sorted by first shell) and a b-side pass (K(b,c), K(b,d) over the
 * same pairs sorted by second shell, summed over runs of equal b before the
 * atomics), twice the integral work for several times fewer atomic
 * requests.  Only the small s/p classes have the two passes compiled, and
 * builds use them for molecules from 2048 Cartesian AOs up (below that the
 * doubled integral work cost more than the atomics saved: 256 molecules of
 * 4-40 waters took 7.2 s against 5.5 s per batch).  on: 0 or 1; for
 * cls = -1 also -1, the built-in rule (the default): two passes for every
 * candidate but (ps|ps) and (dp|ss), which measured as fast or faster in one
 * pass on MI355X.  Process-wide; not to be called while another thread plans
 * a build.  No counterpart in the reference.  Returns 0 ok, 1 bad argument or
 * too many per-class settings. */
int BQCset_row_split(int cls, int on);

/* Launch shape of the row kernels.  Persistent grids have as many
 * workgroups as fit on the device at once, each wavefront pulling work items
 * from a counter in the build's work list until it is exhausted; the other
 * shape has one workgroup per four work items.  (The wavefront kernels always
 * take one workgroup per item: the item loop cost them registers.)
 * persistent: 0 never; 1 for launches holding a molecule of 2048 or more
 * Cartesian AOs; 2 (default) always (measured on MI355X: 64 waters cc-pVDZ
 * 143.0 -> 136.5 ms, 128 waters cc-pVTZ 7.90 -> 7.72 s, batches of 256
 * molecules of 4-40 waters 5.06 -> 4.69 s with the work counters split per
 * XCD; before that split those batches ran 5 % slower persistent).  Same
 * quartets and sums in every mode.
 * Process-wide; not to be called while another thread plans a build.  No
 * counterpart in the reference (it launches one grid per task and bin,
 * nr_jk_driver.cu:252-280).  Returns 0 ok, 1 bad argument. */
int BQCset_launch_mode(int persistent);

/* Host threads that plan a list of tasks (one per 16 tasks, up to n; 0
 * restores the default: the BQC_PLAN_THREADS environment variable, else the
 * hardware threads up to 16).  BQCbuild_jk and BQCbuild_jk_batch plan and
 * launch class by class (the device runs the first classes while the host
 * plans the rest), each class's tasks as one list; BQCgraph_capture plans
 * all tasks as one list.  Blocks of tasks are planned separately and merged
 * in list order, so the work list is the one a single thread builds.
 * Process-wide.  No counterpart in the reference (its host loop plans one
 * task per call, hf.py:111-166).  Returns 0 ok, 1 if n is outside 0..256. */
int BQCset_plan_threads(int n);

/* The tile a build of a molecule with nao Cartesian AOs planned now would
 * use for (kind, cls, prec); -1 for a bad kind.  No counterpart in the
 * reference. */
int BQCget_tile(int kind, int cls, int prec, int nao);

/* Task list of one build on the host (the per-task loop of
 * cuobc/scf/hf.py:111-145 with _make_s_index_offsets, hf.py:567-578).
 * cp_grp (2 x ncp): group pair (gi, gj) of each pair type; log_q (plan
 * order) and bas_pairs_locs as in the plan; dm_cond (ngroups^2): block
 * bounds of |D|.  Writes task_cp (2 per task), task_bins (2 x (nbins+1) per
 * task) and task_dmax; the output arrays hold ncp(ncp+1)/2 tasks. */
int BQCplan_tasks(int ncp, const int *cp_grp, const int *bas_pairs_locs, const double *log_q,
                  const double *dm_cond, int ngroups, double cutoff, int nbins, int prescreen,
                  int *task_cp, int *task_bins, double *task_dmax, int *ntask_out);

/* Block-wise max |D| of nmat device matrices over their (group, group) AO
 * blocks.  mats (device array of device pointers), nao, off, ng, outoff:
 * device int arrays per matrix; loc (device): AO offsets of the groups of
 * matrix m at loc[off[m] .. off[m]+ng[m]]; out (device) receives ng[m]^2
 * values per matrix at outoff[m] and must be zeroed by the caller;
 * gmax = max ng (<= 256), naomax = max nao.  Replaces the host-side
 * lib.condense('absmax', ...) of cuobc/scf/hf.py:106-107. */
int BQCdm_cond(const double *const *mats, const int *nao, const int *loc, const int *off,
               const int *ng, const int *outoff, double *out, int nmat, int gmax, int naomax,
               hipStream_t stream);

/* Schwarz bounds on the GPU: q[i*nbas + j] = sqrt(max |(ij|ij)|) over the
 * Cartesian components of shells i, j (1e-100 when zero), for the shell
 * groups [grp_offsets[g], grp_offsets[g+1]) (ngroups groups, host array);
 * atm/bas/env are device copies of the integral tables.  Returns 0 ok. */
int BQCq_cond(double *q, const int *atm, const int *bas, const double *env, int nbas,
              const int *grp_offsets, int ngroups, const int *grp_l, hipStream_t stream);

/* One-electron integrals S, T, V (nuclear attraction of all atoms in atm)
 * on the GPU, written as full symmetric nao x nao Cartesian matrices over the
 * shells of bas (device tables, shells grouped by non-decreasing l as in the
 * plan: grp_offsets/grp_l host arrays, ngroups groups).  No reference
 * counterpart on this path: the reference's SCF driver takes them from
 * PySCF/libcint on the CPU.  Returns 0 ok, 1 unsupported l / launch error. */
int BQCint1e(double *S, double *T, double *V, const int *atm, int natm, const int *bas,
             const double *env, const int *ao_loc, int nao, const int *grp_offsets, int ngroups,
             const int *grp_l, hipStream_t stream);

/* J = 2 (V + V^T), K = V + V^T in place on device n_dm x nao x nao arrays
 * (either may be NULL). */
int BQCsymmetrize_jk(double *vj, double *vk, int nao, int n_dm, hipStream_t stream);

/* Per-launch timing of BQCbuild_jk / GINTbuild_jk: when enabled, every
 * kernel launch is bracketed by HIP events on the build's stream.
 * BQCprofile_read drains the records (waiting for their events): per launch
 * the evaluated class li*1000+lj*100+lk*10+ll, the kernel kind (0 quartet
 * kernel, 1 row kernel), the algorithmic FLOPs and quartet count of the
 * launch and its duration in ms; returns the number written (<= cap). */
void BQCprofile_enable(int on);
int BQCprofile_count(void);
int BQCprofile_read(int *cls, int *kind, double *flops, double *nquart, double *ms, int cap);

/* Number of FP64 (and FP32) floating-point operations of the last
 * BQCbuild_jk call's quartet list by the algorithmic model of the path
 * (host counters, written into flops[0..1]). */
void BQClast_build_flops(double *flops);

#ifdef __cplusplus
}
#endif
#endif
