// Wavefront-stationary J/K kernel ("pair-row" digestion).
//
// One wavefront owns 64 consecutive pairs (a b) of the P list -- the pair
// type with the larger class -- and walks a range of the Q list (c d) in
// lock-step: at every step all 64 lanes share the same (c d) pair.  Both
// lists are in "bin-major, bra-shell-minor" order: the Schwarz bins of the
// task stay contiguous (so the quartet set is exactly the reference's: a P
// pair of bin s meets the Q pairs of bins [0, nbins - s)), and inside a bin
// pairs with the same first shell are adjacent.
//
// Digestion of the integral block g(ab|cd) at each step (V symmetrised at the
// end, so a target may be written as V[x][y] or V[y][x]):
//   J(a,b) += g D(c,d)    lane register accumulator, flushed once per item
//   J(c,d) += g D(a,b)    summed over the wavefront, one atomic per component
//   K(a,c) += g D(b,d)    lane register accumulator over a run of equal c,
//                         flushed by a segmented sum over runs of equal a
//   K(a,d) += g D(b,c)    segmented sum over runs of equal a at every step
//   K(b,c) += g D(a,d)    lane register accumulator over a run of equal c
//   K(b,d) += g D(a,c)    one atomic per lane and component (smallest block)
//
// The per-lane K(b,c), K(b,d) atomics dominate the atomic requests of the
// small s/p classes.  For those classes every quartet can instead be
// evaluated twice: by the a-side instance (ROLE 1: J(a,b), J(c,d), K(a,c),
// K(a,d)) over P lists sorted by a, and by the b-side instance (ROLE 2:
// K(b,c), K(b,d), as register sums over runs of equal c flushed by segmented
// sums over runs of equal b, and per-step segmented sums over runs of equal
// b) over the same pairs sorted by b.  The two instances together add the
// same six targets as ROLE 0 with twice the integral work and several times
// fewer atomic requests.
#pragma once
#include "jk_kernel.h"

#define BQC_V2_CK 64
// Work counters per persistent launch: the item list is cut into this many
// contiguous slices, one per XCD (workgroups are dispatched to the XCDs in
// turn, so workgroup b runs on XCD b % 8)
#define BQC_TICKET_SLICES 8

// Cross-lane moves of a double through DPP (two 32-bit halves).  Lanes
// whose source is outside the pattern, or whose row is masked off, read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// DPP controls (GFX9 encoding)
#define BQC_DPP_ROW_SHR(k) (0x110 + (k))
#define BQC_DPP_ROW_BCAST15 0x142
#define BQC_DPP_ROW_BCAST31 0x143

// Inclusive prefix sum of v over the lanes of the run containing this lane
// (runs are maximal blocks of consecutive lanes; run_start is the first lane
// of this lane's run): Hillis-Steele within rows of 16 lanes, then the row
// broadcasts of lane 15 and lane 31.  The last lane of a run holds the run
// total.  With run_start = 0 everywhere, lane 63 holds the wave total.
__device__ __forceinline__ double run_scan(double v, int lane, int run_start) {
    double t = dpp_d<BQC_DPP_ROW_SHR(1)>(v);
    if (lane - 1 >= run_start) v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(2)>(v);
    if (lane - 2 >= run_start) v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(4)>(v);
    if (lane - 4 >= run_start) v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(8)>(v);
    if (lane - 8 >= run_start) v += t;
    // rows 1 and 3 receive lane 15 and lane 47
    t = dpp_d<BQC_DPP_ROW_BCAST15, 0xa>(v);
    if ((lane & 16) && run_start <= (lane | 15) - 16) v += t;
    // rows 2 and 3 receive lane 31
    t = dpp_d<BQC_DPP_ROW_BCAST31, 0xc>(v);
    if (lane >= 32 && run_start <= 31) v += t;
    return v;
}

// Sum of v over the wavefront, valid in lane 63.
__device__ __forceinline__ double wave_sum63(double v) {
    double t = dpp_d<BQC_DPP_ROW_SHR(1)>(v);
    v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(2)>(v);
    v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(4)>(v);
    v += t;
    t = dpp_d<BQC_DPP_ROW_SHR(8)>(v);
    v += t;
    t = dpp_d<BQC_DPP_ROW_BCAST15, 0xa>(v);
    v += t;
    t = dpp_d<BQC_DPP_ROW_BCAST31, 0xc>(v);
    v += t;
    return v;
}

// Ordering point for LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-packed atomics.  Every lane holds NV values for elements of V at
// offsets off[v] (-1: nothing to add), the components of one lane adjacent
// in memory.  Staged through LDS, atomic instruction j then covers entries
// 64j .. 64j+63 in lane-major order, so the components of a lane fall into
// one instruction and mostly into one 64-byte segment: the atomic units take
// one request per segment and instruction, so this needs about NV times
// fewer requests than one instruction per component.
template <int NV>
__device__ __forceinline__ void packed_atomic_add(gdouble *V, const double (&val)[NV],
                                                  const int (&off)[NV], double *sv, int *so,
                                                  int lane) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        sv[lane * NV + v] = val[v];
        so[lane * NV + v] = off[v];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const int e = j * 64 + lane;
        const int o = so[e];
        const double x = sv[e];
        if (o >= 0) atomic_add(V + o, x);
    }
    wave_sync();
}

// Largest packed group of a row-kernel class (groups above 9 components,
// and single components, are issued directly).
constexpr int pack_nv(int n) { return n >= 2 && n <= 9 ? n : 0; }
constexpr int max2(int a, int b) { return a > b ? a : b; }
template <int NA, int NB, int NC, int ND>
constexpr int row_pack_max() {
    return max2(max2(pack_nv(NB * ND), pack_nv(NA * ND)),
                max2(pack_nv(NA * NC), pack_nv(NB * NC)));
}

// Work item: (task, chunk of 64 P pairs, first and end Q position).  With
// ticket == NULL wavefront w of block b takes item 4b + w; otherwise the grid
// is persistent and the wavefronts pull items from the counter *ticket
// (zero before the launch).
template <int LA, int LB, int LC, int LD, typename T, int ROLE = 0>
__global__ void __launch_bounds__(256)
    jk_rows(const TaskDesc *__restrict__ tasks, const int4 *__restrict__ items, int nitems,
            int *ticket) {
    using C = Cls<LA, LB, LC, LD>;
    constexpr int NA = C::NFI, NB = C::NFJ, NC = C::NFK, ND = C::NFL;
    // Rys interpolation table of this class's root count staged in LDS
    // (8 KB for 1 root, 20 KB for 2) before any wave leaves
    constexpr bool RYS_LDS = C::NR <= 2;
    constexpr int RYS_N = RYS_LDS ? rys_table_size(C::NR) : 1;
    __shared__ double rys_lds[RYS_N];
    if (RYS_LDS) {
        const double *src = BQC_RYS_COEF + rys_offset(C::NR);
        for (int i = threadIdx.x; i < RYS_N; i += blockDim.x) rys_lds[i] = src[i];
        __syncthreads();
    }
    typedef const __attribute__((address_space(3))) double *lds_ptr;
    const lds_ptr rys_tab = (lds_ptr)rys_lds;
    // per-wavefront staging of packed atomics
    constexpr int PK = row_pack_max<NA, NB, NC, ND>();
    __shared__ double pk_val[4 * 64 * (PK > 0 ? PK : 1)];
    __shared__ int pk_off[4 * 64 * (PK > 0 ? PK : 1)];
    double *const sv = pk_val + (threadIdx.x >> 6) * 64 * (PK > 0 ? PK : 1);
    int *const so = pk_off + (threadIdx.x >> 6) * 64 * (PK > 0 ? PK : 1);
    const int lane = threadIdx.x & 63;
    auto run_item = [&](const int item) {
        const int4 it = items[item];
        const TaskDesc &tk = tasks[__builtin_amdgcn_readfirstlane(it.x)];
        const PairData &pd = tk.pd;
        const JKArgs jk = {tk.vj, tk.vk, tk.dm, tk.nao, tk.n_dm};
        const int chunk = __builtin_amdgcn_readfirstlane(it.y);
        const int q0 = __builtin_amdgcn_readfirstlane(it.z);
        const int qe = __builtin_amdgcn_readfirstlane(it.w);
        const int pidx = chunk * 64 + lane;
        const bool pvalid = pidx < tk.np;
        // bin of this P pair -> end of its Q prefix
        int qlim = 0, ppos = 0;
        if (pvalid) {
            int s = 0;
            while (s < tk.nbins && pidx >= tk.pbins[s + 1]) ++s;
            qlim = tk.qbins[tk.nbins - s];
            ppos = glb(tk.perm_p)[pidx];
        }
        const int qend_item = min(qe, tk.qbins[tk.nbins]);
        int qlim_max = qlim;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) qlim_max = max(qlim_max, __shfl_xor(qlim_max, d, 64));
        const int q1 = min(qend_item, __builtin_amdgcn_readfirstlane(qlim_max));
        if (q1 <= q0) return;

        // per-lane (P side) and wave-uniform (Q side) views of the plan data
        const gptr<int> lbra = glb(pd.bra), lket = glb(pd.ket), lao = glb(pd.ao_loc);
        const gptr<double> lbx = glb(pd.bas_x), lby = glb(pd.bas_y), lbz = glb(pd.bas_z);
        // Q list records {plan position, first shell, second shell, AO offsets}
        const cptr<int> uqrec = cst(tk.qrec);
        const cptr<double> ubx = cst(pd.bas_x), uby = cst(pd.bas_y), ubz = cst(pd.bas_z);
        const int gp = tk.bas_p0 + ppos;
        const int sa = pvalid ? lbra[gp] : -1;
        const int sb = pvalid ? lket[gp] : 0;
        const int pp0 = tk.prim_p0 + ppos * tk.nprim_p;
        const int a0 = pvalid ? lao[sa] : 0;
        const int b0 = pvalid ? lao[sb] : 0;
        double ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
        double p_first[5] = {1, 0, 0, 0, 0};
        if (pvalid) {
            ra[0] = lbx[sa]; ra[1] = lby[sa]; ra[2] = lbz[sa];
            rb[0] = lbx[sb]; rb[1] = lby[sb]; rb[2] = lbz[sb];
            p_first[0] = glb(pd.a12)[pp0]; p_first[1] = glb(pd.e12)[pp0];
            p_first[2] = glb(pd.x12)[pp0]; p_first[3] = glb(pd.y12)[pp0];
            p_first[4] = glb(pd.z12)[pp0];
        }
        // runs of equal first shell a among the 64 lanes (of equal second
        // shell b for the b-side role, whose lists are sorted by b)
        const int rkey = ROLE == 2 ? (pvalid ? sb : -1) : sa;
        const int key_next = __shfl_down(rkey, 1, 64);
        const bool tail = lane == 63 || key_next != rkey;
        const int key_prev = __shfl_up(rkey, 1, 64);
        const unsigned long long heads = __ballot(lane == 0 || key_prev != rkey);
        const int run_start = 63 - __builtin_clzll(heads & (~0ull >> (63 - lane)));

        const int nao = jk.nao;
        const size_t nao2 = (size_t)nao * nao;
        for (int idm = 0; idm < jk.n_dm; ++idm) {
            const gptr<double> D = glb(jk.dm + idm * nao2);      // per-lane reads
            const cptr<double> DU = cst(jk.dm + idm * nao2);     // wave-uniform reads
            gdouble *VJ = jk.vj ? glb_rw(jk.vj + idm * nao2) : nullptr;
            gdouble *VK = jk.vk ? glb_rw(jk.vk + idm * nao2) : nullptr;
            double dab[NA * NB];
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) dab[x * NB + y] = pvalid ? D[(size_t)(a0 + x) * nao + b0 + y] : 0.0;
            double jab[NA * NB], kac[NA * NC], kbc[NB * NC];
#pragma unroll
            for (int n = 0; n < NA * NB; ++n) jab[n] = 0.0;
#pragma unroll
            for (int n = 0; n < NA * NC; ++n) kac[n] = 0.0;
#pragma unroll
            for (int n = 0; n < NB * NC; ++n) kbc[n] = 0.0;

            // the records of steps [qb, qb + 64), one per lane, from one
            // coalesced load; a step reads its own and the next step's first
            // shell with v_readlane, so the Q walk issues no scalar loads
            int rqx = __builtin_amdgcn_readfirstlane(uqrec[4 * q0]), rqy = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 1]);
            int rqz = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 2]), rqw = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 3]);
            int sc = rqy;
            // D(b,c), D(a,c): loaded at the start of every run of equal c
            double dbc[NB * NC], dac[NA * NC];
            bool new_c = true;
            for (int q = q0; q < q1; ++q) {
                const int qpos = rqx, sd = rqz;
                const int c0 = rqw & 0xffff, d0 = (int)((unsigned)rqw >> 16);
                int rnx = rqx, rny = rqy, rnz = rqz, rnw = rqw;
                if (q + 1 < q1) {
                    rnx = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 4]);
                    rny = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 5]);
                    rnz = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 6]);
                    rnw = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 7]);
                }
                bool active = pvalid && q < qlim;
                double norm = TWO_PI_52;
                if (tk.diag) {
                    active = active && ppos >= qpos;
                    if (ppos == qpos) norm *= 0.5;
                }
                T g[C::NF];
                if (active) {
                    // first Q primitive pair, marked wave-uniform so that the
                    // primitive loop reads its data with scalar loads
                    const int qq0 = __builtin_amdgcn_readfirstlane(tk.prim_q0 + qpos * tk.nprim_q);
                    const double rc[3] = {ubx[sc], uby[sc], ubz[sc]};
                    const double rd[3] = {ubx[sd], uby[sd], ubz[sd]};
                    if constexpr (RYS_LDS)
                        eri_block<LA, LB, LC, LD, T>(g, pd, pp0, tk.nprim_p,
                                                     qq0, tk.nprim_q,
                                                     norm, ra, rb, rc, rd, 0, rys_tab, p_first);
                    else
                        eri_block<LA, LB, LC, LD, T>(g, pd, pp0, tk.nprim_p,
                                                     qq0, tk.nprim_q,
                                                     norm, ra, rb, rc, rd, 0, nullptr, p_first);
                } else {
#pragma unroll
                    for (int n = 0; n < C::NF; ++n) g[n] = (T)0;
                }
#define G(x, y, z, w) ((double)g[(((x) * NB + (y)) * NC + (z)) * ND + (w)])
                if (ROLE != 2 && VJ) {
                    // J(a,b) += g D(c,d)
#pragma unroll
                    for (int z = 0; z < NC; ++z)
#pragma unroll
                        for (int w = 0; w < ND; ++w) {
                            const double dcd = DU[(size_t)(c0 + z) * nao + d0 + w];
#pragma unroll
                            for (int x = 0; x < NA; ++x)
#pragma unroll
                                for (int y = 0; y < NB; ++y) jab[x * NB + y] += G(x, y, z, w) * dcd;
                        }
                    // J(c,d) += g D(a,b), whole wavefront
#pragma unroll
                    for (int z = 0; z < NC; ++z)
#pragma unroll
                        for (int w = 0; w < ND; ++w) {
                            double s = 0.0;
#pragma unroll
                            for (int x = 0; x < NA; ++x)
#pragma unroll
                                for (int y = 0; y < NB; ++y) s += G(x, y, z, w) * dab[x * NB + y];
                            s = wave_sum63(s);
                            if (lane == 63) atomic_add(VJ + (size_t)(c0 + z) * nao + d0 + w, s);
                        }
                }
                if (VK) {
                    // K(a,c) += g D(b,d); K(b,c) += g D(a,d); K(a,d) += g D(b,c); K(b,d) += g D(a,c)
                    double dbd[NB * ND], dad[NA * ND];
#pragma unroll
                    for (int w = 0; w < ND; ++w) {
#pragma unroll
                        for (int y = 0; y < NB; ++y)
                            dbd[y * ND + w] = pvalid ? D[(size_t)(d0 + w) * nao + b0 + y] : 0.0;
#pragma unroll
                        for (int x = 0; x < NA; ++x)
                            dad[x * ND + w] = pvalid ? D[(size_t)(d0 + w) * nao + a0 + x] : 0.0;
                    }
                    if (new_c) {
#pragma unroll
                        for (int z = 0; z < NC; ++z) {
#pragma unroll
                            for (int y = 0; y < NB; ++y)
                                dbc[y * NC + z] = pvalid ? D[(size_t)(c0 + z) * nao + b0 + y] : 0.0;
#pragma unroll
                            for (int x = 0; x < NA; ++x)
                                dac[x * NC + z] = pvalid ? D[(size_t)(c0 + z) * nao + a0 + x] : 0.0;
                        }
                    }
#pragma unroll
                    for (int x = 0; x < NA; ++x)
#pragma unroll
                        for (int y = 0; y < NB; ++y)
#pragma unroll
                            for (int z = 0; z < NC; ++z)
#pragma unroll
                                for (int w = 0; w < ND; ++w) {
                                    const double v = G(x, y, z, w);
                                    if (ROLE != 2) kac[x * NC + z] += v * dbd[y * ND + w];
                                    if (ROLE != 1) kbc[y * NC + z] += v * dad[x * ND + w];
                                }
                    // K(a,d): segmented over runs of equal a, added by the run tails
                    if constexpr (ROLE != 2) {
                        constexpr int NV = NA * ND;
                        double val[NV];
                        int off[NV];
#pragma unroll
                        for (int w = 0; w < ND; ++w)
#pragma unroll
                            for (int x = 0; x < NA; ++x) {
                                double s = 0.0;
#pragma unroll
                                for (int y = 0; y < NB; ++y)
#pragma unroll
                                    for (int z = 0; z < NC; ++z) s += G(x, y, z, w) * dbc[y * NC + z];
                                val[w * NA + x] = run_scan(s, lane, run_start);
                                off[w * NA + x] = tail && pvalid ? (d0 + w) * nao + a0 + x : -1;
                            }
                        if constexpr (pack_nv(NV) > 0) {
                            packed_atomic_add<NV>(VK, val, off, sv, so, lane);
                        } else {
#pragma unroll
                            for (int v = 0; v < NV; ++v)
                                if (off[v] >= 0) atomic_add(VK + off[v], val[v]);
                        }
                    }
                    // K(b,d): per lane (segmented over runs of equal b in the b-side role)
                    if constexpr (ROLE != 1) {
                        constexpr int NV = NB * ND;
                        double val[NV];
                        int off[NV];
#pragma unroll
                        for (int w = 0; w < ND; ++w)
#pragma unroll
                            for (int y = 0; y < NB; ++y) {
                                double s = 0.0;
#pragma unroll
                                for (int x = 0; x < NA; ++x)
#pragma unroll
                                    for (int z = 0; z < NC; ++z) s += G(x, y, z, w) * dac[x * NC + z];
                                if constexpr (ROLE == 2) {
                                    val[w * NB + y] = run_scan(s, lane, run_start);
                                    off[w * NB + y] = tail && pvalid ? (d0 + w) * nao + b0 + y : -1;
                                } else {
                                    val[w * NB + y] = s;
                                    off[w * NB + y] = active ? (d0 + w) * nao + b0 + y : -1;
                                }
                            }
                        if constexpr (pack_nv(NV) > 0) {
                            packed_atomic_add<NV>(VK, val, off, sv, so, lane);
                        } else {
#pragma unroll
                            for (int v = 0; v < NV; ++v)
                                if (off[v] >= 0) atomic_add(VK + off[v], val[v]);
                        }
                    }
                }
#undef G
                // end of a run of equal c: flush K(a,c), K(b,c)
                const int sc_next = q + 1 < q1 ? rny : -1;
                if (VK && sc_next != sc) {
                    if constexpr (ROLE != 2) {
                        constexpr int NV = NA * NC;
                        double val[NV];
                        int off[NV];
#pragma unroll
                        for (int z = 0; z < NC; ++z)
#pragma unroll
                            for (int x = 0; x < NA; ++x) {
                                val[z * NA + x] = run_scan(kac[x * NC + z], lane, run_start);
                                off[z * NA + x] = tail && pvalid ? (c0 + z) * nao + a0 + x : -1;
                                kac[x * NC + z] = 0.0;
                            }
                        if constexpr (pack_nv(NV) > 0) {
                            packed_atomic_add<NV>(VK, val, off, sv, so, lane);
                        } else {
#pragma unroll
                            for (int v = 0; v < NV; ++v)
                                if (off[v] >= 0) atomic_add(VK + off[v], val[v]);
                        }
                    }
                    if constexpr (ROLE != 1) {
                        constexpr int NV = NB * NC;
                        double val[NV];
                        int off[NV];
#pragma unroll
                        for (int z = 0; z < NC; ++z)
#pragma unroll
                            for (int y = 0; y < NB; ++y) {
                                if constexpr (ROLE == 2) {
                                    val[z * NB + y] = run_scan(kbc[y * NC + z], lane, run_start);
                                    off[z * NB + y] =
                                        tail && pvalid ? (c0 + z) * nao + b0 + y : -1;
                                } else {
                                    val[z * NB + y] = kbc[y * NC + z];
                                    off[z * NB + y] = pvalid && kbc[y * NC + z] != 0.0
                                                          ? (c0 + z) * nao + b0 + y
                                                          : -1;
                                }
                                kbc[y * NC + z] = 0.0;
                            }
                        if constexpr (pack_nv(NV) > 0) {
                            packed_atomic_add<NV>(VK, val, off, sv, so, lane);
                        } else {
#pragma unroll
                            for (int v = 0; v < NV; ++v)
                                if (off[v] >= 0) atomic_add(VK + off[v], val[v]);
                        }
                    }
                }
                new_c = sc_next != sc;
                sc = sc_next;
                rqx = rnx;
                rqy = rny;
                rqz = rnz;
                rqw = rnw;
            }
            if (ROLE != 2 && VJ && pvalid) {
#pragma unroll
                for (int x = 0; x < NA; ++x)
#pragma unroll
                    for (int y = 0; y < NB; ++y)
                        if (jab[x * NB + y] != 0.0)
                            atomic_add(VJ + (size_t)(a0 + x) * nao + b0 + y, jab[x * NB + y]);
            }
        }
    };
    // one work item per wavefront, or (persistent grids) items taken from
    // the slice counters until the list is exhausted: a wavefront starts on
    // the slice of its XCD, so the items an XCD runs together are neighbours
    // in the list (nearby shells: their density reads and J/K atomics share
    // that XCD's L2), and moves on to the next slice when its own is empty.
    // One call site keeps the item's code inlined once.
    int slice = blockIdx.x % BQC_TICKET_SLICES, tries = 0;
    auto next_item = [&]() -> int {
        if (!ticket) return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
        while (tries < BQC_TICKET_SLICES) {
            const int lo = (int)((long long)nitems * slice / BQC_TICKET_SLICES);
            const int hi = (int)((long long)nitems * (slice + 1) / BQC_TICKET_SLICES);
            int tk0 = 0;
            if ((threadIdx.x & 63) == 0)
                tk0 = __hip_atomic_fetch_add(ticket + slice, 1, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            const int it = lo + __builtin_amdgcn_readfirstlane(tk0);
            if (it < hi) return it;
            slice = slice + 1 == BQC_TICKET_SLICES ? 0 : slice + 1;
            ++tries;
        }
        return nitems;
    };
    for (int item = next_item(); item < nitems; item = next_item()) {
        run_item(item);
        if (!ticket) break;
    }
}

typedef void (*RowKernel)(const TaskDesc *, const int4 *, int, int *);
// [role][precision][bra pair id][ket pair id]: role 0 the single pass, roles
// 1 and 2 the a-side and b-side passes of split classes (FP64 only)
typedef RowKernel RowKernelTable[3][2][BQC_NPAIRID][BQC_NPAIRID];

// Classes digested with the wavefront-stationary kernel, per precision.
// Up to 108 Cartesian components the FP64 block and the lane accumulators
// fit the 512 registers of one wave per SIMD (with small spills), and this
// kernel beats the wavefront-per-quartet one by ~3x (measured on (dp|ds));
// at 162 the spills make it slower than that kernel.  The FP32 instance has
// the same limit: at 162-216 components (2-3 KB of scratch per lane) it ran
// at 1.2-1.4 TFLOP/s against 2.4-3.5 for the wavefront-per-quartet kernel.
template <int LA, int LB, int LC, int LD, typename T>
constexpr bool use_rows() {
    return Cls<LA, LB, LC, LD>::NF <= 108;
}

// Row classes with an FP32 instance (used by mixed precision).  The others
// are bound by atomic requests, which FP32 integrals do not reduce, while
// splitting a row's Q range by precision adds a flush of its accumulators:
// with every row class split, 64 H2O cc-pVDZ took 168 ms against 158 ms in
// FP64; on 128 H2O cc-pVTZ these six gained 7-105 ms each and the rest
// -3..+2 ms.  Mixed builds run the quartets of the others in FP64.
template <int LA, int LB, int LC, int LD>
constexpr bool row_fp32() {
    constexpr int c = LA * 1000 + LB * 100 + LC * 10 + LD;
    return c == 2010 || c == 2100 || c == 2120 || c == 2210 || c == 3020 || c == 3110;
}

// Row classes compiled with the a-side / b-side pair as well: the small s/p
// classes, whose time goes to atomic requests rather than arithmetic.  Whether
// a build uses the pair is chosen per class at run time (BQCset_row_split).
template <int LA, int LB, int LC, int LD>
constexpr bool row_split_candidate() {
    return Cls<LA, LB, LC, LD>::NF <= 27 && LA <= 2 && LC <= 1;
}

template <int LA, int LB, int LC, int LD>
static void reg_row(RowKernelTable &tab) {
    constexpr int pab = LA * (LA + 1) / 2 + LB, pcd = LC * (LC + 1) / 2 + LD;
    if constexpr (use_rows<LA, LB, LC, LD, double>())
        tab[0][0][pab][pcd] = jk_rows<LA, LB, LC, LD, double>;
    if constexpr (use_rows<LA, LB, LC, LD, float>() && row_fp32<LA, LB, LC, LD>())
        tab[0][1][pab][pcd] = jk_rows<LA, LB, LC, LD, float>;
    if constexpr (use_rows<LA, LB, LC, LD, double>() && row_split_candidate<LA, LB, LC, LD>()) {
        tab[1][0][pab][pcd] = jk_rows<LA, LB, LC, LD, double, 1>;
        tab[2][0][pab][pcd] = jk_rows<LA, LB, LC, LD, double, 2>;
    }
}
