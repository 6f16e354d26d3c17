// Wavefront-per-quartet J/K kernel for classes with large integral blocks.
//
// A 64-lane workgroup owns one P pair (a b) and walks a range of the Q list
// (c d) -- the same bin-major, first-shell-minor lists and work items as the
// row kernel, with one P pair per item instead of 64.  For every quartet:
//   A. lanes t < 3*NR evaluate the 2-D Rys table of direction t%3, root t/3
//      (FP64 roots, T-typed recurrence) into LDS;
//   B. the 64 lanes split the Cartesian components and accumulate the
//      contracted block over roots and primitive quartets in registers;
//   C. the block goes to LDS and the lanes split the elements of the six
//      digestion targets, several lanes per element where the sums are long
//      (digest_split.h):
//        J(a,b) += g D(c,d)  lane registers, flushed once per item
//        K(a,c) += g D(b,d)  lane registers over a run of equal c
//        K(b,c) += g D(a,d)  lane registers over a run of equal c
//        J(c,d), K(a,d), K(b,d): one FP64 atomic per element and quartet.
// LDS per workgroup: 3*NR*GSZ + NF values of type T plus the density blocks.
// The workgroup is one wavefront, so the phases are ordered with wavefront
// fences (LDS operations of one wave complete in order) instead of
// workgroup barriers.
#pragma once
#include "digest_split.h"
#include "jk_v2.h"

template <int LA, int LB, int LC, int LD>
struct WaveCls {
    using C = Cls<LA, LB, LC, LD>;
    static constexpr int NA = C::NFI, NB = C::NFJ, NC = C::NFK, ND = C::NFL;
    static constexpr int NG = (C::NF + 63) / 64;              // block values per lane
    static constexpr int NAB = NA * NB, NCD = NC * ND, NAC = NA * NC, NAD = NA * ND,
                         NBC = NB * NC, NBD = NB * ND;
    // density elements per lane of each block
    static constexpr int EAC = (NAC + 63) / 64, EBC = (NBC + 63) / 64, ECD = (NCD + 63) / 64,
                         EBD = (NBD + 63) / 64, EAD = (NAD + 63) / 64;
};

// Sum of digestion target TG for element e over the outer indices u = s,
// s + G, ... of lane s of its group, added over the group's G lanes (the
// group's lanes all hold the total).  Every lane of the wavefront calls it;
// lanes of groups past the last element pass valid = false.
template <class TG, typename T>
__device__ __forceinline__ double split_sum(const T *sg, const double *sd, int e, int s,
                                            bool valid) {
    using S = Split<TG>;
    double sum = 0.0;
    if (valid) {
        const T *gp = sg + TG::gbase(e) + s * TG::SU;
        const double *dp = sd + s * TG::DU;
#pragma unroll
        for (int i = 0; i < S::IU; ++i) {
            if (TG::NU % S::G == 0 || s + S::G * i < TG::NU) {
#pragma unroll
                for (int v = 0; v < TG::NV; ++v)
                    sum += (double)gp[i * S::G * TG::SU + v * TG::SV] *
                           dp[i * S::G * TG::DU + v * TG::DV];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < S::G; o <<= 1) sum += __shfl_xor(sum, o, 64);
    return sum;
}

// I-table index of component n = ((x*NB + y)*NC + z)*ND + w for direction d
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ int gidx(int n, int d) {
    using W = WaveCls<LA, LB, LC, LD>;
    using C = typename W::C;
    const int w = n % W::ND, z = (n / W::ND) % W::NC, y = (n / (W::ND * W::NC)) % W::NB,
              x = n / (W::ND * W::NC * W::NB);
    auto e = [d](int l, int c) { return d == 0 ? cart_x(l, c) : (d == 1 ? cart_y(l, c) : cart_z(l, c)); };
    return C::g(e(LA, x), e(LB, y), e(LC, z), e(LD, w));
}

// Whether the kernel loads a quartet's density elements before its integral
// phases (hiding their latency behind phases A and B) or in the digestion
// phase.  The early loads hold registers through phases A and B; in the
// classes listed here that cost resident waves and made the kernel slower on
// MI355X ((fd|dp) 106 -> 144 ms, (fp|pp) 153 -> 170 ms), so they load late.
// Elsewhere early loads gain up to 25 % ((dp|dp) 412 -> 328 ms; measured per
// class on 128 H2O cc-pVTZ with every class early and every class late).
template <int LA, int LB, int LC, int LD>
constexpr bool wave_prefetch() {
    constexpr int c = LA * 1000 + LB * 100 + LC * 10 + LD;
    return c != 2211 && c != 3310 && c != 3221 && c != 3111;
}

// Resident waves per SIMD the compiler must allow (register budget 512/N).
// For these classes a budget of 128 registers runs 1-4 % faster than the
// budget the compiler picks by itself; for others it runs up to 3x slower
// ((dd|dp) 40 -> 128 ms), and for (dp|pp) and (fp|ds) it no longer pays
// (measured per class on 128 H2O cc-pVTZ with every wave class hinted and
// with none), so the hint is per class.
template <int LA, int LB, int LC, int LD>
constexpr int wave_min_waves() {
    constexpr int c = LA * 1000 + LB * 100 + LC * 10 + LD;
    return (c == 2220 || c == 3021 || c == 3210 || c == 3310) ? 4 : 1;
}

// Work item: (task, P position, first and end Q position); block b takes
// item b (ticket is unused: the launch table shares the row kernel's
// signature).
template <int LA, int LB, int LC, int LD, typename T>
__global__ void __launch_bounds__(64, (wave_min_waves<LA, LB, LC, LD>()))
    jk_wave(const TaskDesc *__restrict__ tasks, const int4 *__restrict__ items, int nitems,
            int *ticket) {
    using W = WaveCls<LA, LB, LC, LD>;
    using C = typename W::C;
    constexpr int NR = C::NR, NA = W::NA, NB = W::NB, NC = W::NC, ND = W::ND;
    using DT = DigestTargets<NA, NB, NC, ND>;
    using SAB = Split<typename DT::Jab>;
    using SCD = Split<typename DT::Jcd>;
    using SAC = Split<typename DT::Kac>;
    using SBC = Split<typename DT::Kbc>;
    using SAD = Split<typename DT::Kad>;
    using SBD = Split<typename DT::Kbd>;
    __shared__ T sI[3 * NR * C::GSZ];
    __shared__ T sg[C::NF];
    __shared__ double sD[W::NAB + W::NCD + W::NBD + W::NAD + W::NBC + W::NAC];
    double *sDab = sD, *sDcd = sDab + W::NAB, *sDbd = sDcd + W::NCD, *sDad = sDbd + W::NBD,
           *sDbc = sDad + W::NAD, *sDac = sDbc + W::NBC;

    const int lane = threadIdx.x;
    // one work item per workgroup (ticket is unused: a persistent loop over
    // items cost registers that this kernel's tight budget lacks -- (dp|pp)
    // at four waves per SIMD ran 30 % slower with it)
    (void)ticket;
    const int item = blockIdx.x;
    if (item >= nitems) return;
    const int4 it = items[item];
    const TaskDesc &tk = tasks[it.x];
    const PairData &pd = tk.pd;
    const JKArgs jk = {tk.vj, tk.vk, tk.dm, tk.nao, tk.n_dm};
    // everything but the density gathers is the same for all lanes: scalar loads
    const cptr<int> ubra = cst(pd.bra), uket = cst(pd.ket), uao = cst(pd.ao_loc);
    // Q list records {plan position, first shell, second shell, AO offsets}
    const cptr<int> uqrec = cst(tk.qrec);
    const cptr<double> ubx = cst(pd.bas_x), uby = cst(pd.bas_y), ubz = cst(pd.bas_z);
    const cptr<double> ua12 = cst(pd.a12), ue12 = cst(pd.e12), ux12 = cst(pd.x12),
                       uy12 = cst(pd.y12), uz12 = cst(pd.z12);
    const int pidx = it.y;
    // the planner ends every item at or before the Q limit of its P pair's
    // bin (kl bin b meets ij bins [0, nbins - b)), so no bin lookup here
    const int q0 = it.z;
    const int q1 = it.w;
    if (q1 <= q0) return;
    const int ppos = cst(tk.perm_p)[pidx];
    const int gp = tk.bas_p0 + ppos;
    const int sa = ubra[gp], sb = uket[gp];
    const int pp0 = tk.prim_p0 + ppos * tk.nprim_p;
    const int a0 = uao[sa], b0 = uao[sb];
    const double ra[3] = {ubx[sa], uby[sa], ubz[sa]};
    const double rb[3] = {ubx[sb], uby[sb], ubz[sb]};
    const double ab[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
    // I-table indices of this lane's components, x | y << 10 | z << 20 (one
    // register per component instead of three)
    static_assert(C::GSZ <= 1024, "I-table index does not fit 10 bits");
    int gi[W::NG];
#pragma unroll
    for (int m = 0; m < W::NG; ++m) {
        const int n = lane + 64 * m;
        gi[m] = n < C::NF ? gidx<LA, LB, LC, LD>(n, 0) | gidx<LA, LB, LC, LD>(n, 1) << 10 |
                                gidx<LA, LB, LC, LD>(n, 2) << 20
                          : 0;
    }
    const int nao = jk.nao;
    const size_t nao2 = (size_t)nao * nao;
    for (int idm = 0; idm < jk.n_dm; ++idm) {
        const gptr<double> D = glb(jk.dm + idm * nao2);
        gdouble *VJ = jk.vj ? glb_rw(jk.vj + idm * nao2) : nullptr;
        gdouble *VK = jk.vk ? glb_rw(jk.vk + idm * nao2) : nullptr;
        for (int e = lane; e < W::NAB; e += 64)
            sDab[e] = D[(size_t)(a0 + e / NB) * nao + b0 + e % NB];
        // register targets: element p * PER + lane / G, held by the group's
        // first lane (the others keep zeros)
        double jab[SAB::PASSES], kac[SAC::PASSES], kbc[SBC::PASSES];
#pragma unroll
        for (int m = 0; m < SAB::PASSES; ++m) jab[m] = 0.0;
#pragma unroll
        for (int m = 0; m < SAC::PASSES; ++m) kac[m] = 0.0;
#pragma unroll
        for (int m = 0; m < SBC::PASSES; ++m) kbc[m] = 0.0;
        int rqx = __builtin_amdgcn_readfirstlane(uqrec[4 * q0]), rqy = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 1]);
        int rqz = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 2]), rqw = __builtin_amdgcn_readfirstlane(uqrec[4 * q0 + 3]);
        int sc = rqy;
        for (int q = q0; q < q1; ++q) {
            const int qpos = rqx, sd = rqz;
            const int c0 = rqw & 0xffff, d0 = (int)((unsigned)rqw >> 16);
            // the next step's record, in flight during this step
            int rnx = rqx, rny = rqy, rnz = rqz, rnw = rqw;
            if (q + 1 < q1) {
                rnx = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 4]);
                rny = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 5]);
                rnz = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 6]);
                rnw = __builtin_amdgcn_readfirstlane(uqrec[4 * q + 7]);
            }
            const bool skip = tk.diag && ppos < qpos;
            // density elements of this Q pair, stored to LDS for the digestion
            double pcd[W::ECD], pbd[W::EBD], pad[W::EAD], pbc[W::EBC], pac[W::EAC];
            constexpr bool PF = wave_prefetch<LA, LB, LC, LD>();
            if constexpr (PF) {  // loads overlap phases A and B
#pragma unroll
                for (int m = 0; m < W::ECD; ++m) {
                    const int e = lane + 64 * m;
                    if (e < W::NCD) pcd[m] = D[(size_t)(c0 + e / ND) * nao + d0 + e % ND];
                }
#pragma unroll
                for (int m = 0; m < W::EBD; ++m) {
                    const int e = lane + 64 * m;
                    if (e < W::NBD) pbd[m] = D[(size_t)(d0 + e % ND) * nao + b0 + e / ND];
                }
#pragma unroll
                for (int m = 0; m < W::EAD; ++m) {
                    const int e = lane + 64 * m;
                    if (e < W::NAD) pad[m] = D[(size_t)(d0 + e % ND) * nao + a0 + e / ND];
                }
#pragma unroll
                for (int m = 0; m < W::EBC; ++m) {
                    const int e = lane + 64 * m;
                    if (e < W::NBC) pbc[m] = D[(size_t)(c0 + e % NC) * nao + b0 + e / NC];
                }
#pragma unroll
                for (int m = 0; m < W::EAC; ++m) {
                    const int e = lane + 64 * m;
                    if (e < W::NAC) pac[m] = D[(size_t)(c0 + e % NC) * nao + a0 + e / NC];
                }
            }
            double norm = TWO_PI_52;
            if (tk.diag && ppos == qpos) norm *= 0.5;
            // the block accumulates in registers, or for large blocks (more than
            // 16 values per lane) in this lane's own elements of sg, which
            // keeps the registers for the rest and avoids spills
            constexpr bool GL = W::NG > 16;
            T g[GL ? 1 : W::NG];
#pragma unroll
            for (int m = 0; m < W::NG; ++m) {
                if constexpr (GL) {
                    if (lane + 64 * m < C::NF) sg[lane + 64 * m] = (T)0;
                } else {
                    g[m] = (T)0;
                }
            }
            if (!skip) {
                const double rc[3] = {ubx[sc], uby[sc], ubz[sc]};
                const double rd[3] = {ubx[sd], uby[sd], ubz[sd]};
                const double cd[3] = {rc[0] - rd[0], rc[1] - rd[1], rc[2] - rd[2]};
                const int qq0 = tk.prim_q0 + qpos * tk.nprim_q;
                for (int ip = 0; ip < tk.nprim_p; ++ip) {
                    const int p = pp0 + ip;
                    const double aij = ua12[p], eij = ue12[p];
                    const double P[3] = {ux12[p], uy12[p], uz12[p]};
                    for (int iq = 0; iq < tk.nprim_q; ++iq) {
                        const int qq = qq0 + iq;
                        const double akl = ua12[qq], ekl = ue12[qq];
                        const double Q[3] = {ux12[qq], uy12[qq], uz12[qq]};
                        const double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                        const double aijkl = aij + akl, a1 = aij * akl, a0q = a1 / aijkl;
                        const double x = a0q * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
                        // phase A: one (direction, root) table per lane
                        if (lane < 3 * NR) {
                            const int d = lane % 3, r = lane / 3;
                            double ur, wr;
                            rys_root_one<NR>(x, r, ur, wr);
                            const double u2 = a0q * ur;
                            const double tmp4 = 0.5 / (u2 * aijkl + a1);
                            const double b00 = u2 * tmp4;
                            const double b10 = b00 + tmp4 * akl;
                            const double b01 = b00 + tmp4 * aij;
                            const double seed = d == 0 ? norm
                                                : (d == 1 ? eij * ekl / (sqrt(aijkl) * a1) : wr);
                            const double Pd = d == 0 ? P[0] : (d == 1 ? P[1] : P[2]);
                            const double Qd = d == 0 ? Q[0] : (d == 1 ? Q[1] : Q[2]);
                            const double Ad = d == 0 ? ra[0] : (d == 1 ? ra[1] : ra[2]);
                            const double Cd = d == 0 ? rc[0] : (d == 1 ? rc[1] : rc[2]);
                            const double pqd = Pd - Qd;
                            const double abd = d == 0 ? ab[0] : (d == 1 ? ab[1] : ab[2]);
                            const double cdd = d == 0 ? cd[0] : (d == 1 ? cd[1] : cd[2]);
                            rys2d<LA, LB, LC, LD, T>(sI + (r * 3 + d) * C::GSZ, (T)seed,
                                                     (T)((Pd - Ad) - 2.0 * b00 * akl * pqd),
                                                     (T)((Qd - Cd) + 2.0 * b00 * aij * pqd), (T)b10,
                                                     (T)b01, (T)b00, (T)abd, (T)cdd);
                        }
                        wave_sync();
                        // phase B: components
#pragma unroll
                        for (int m = 0; m < W::NG; ++m) {
                            if (lane + 64 * m < C::NF) {
                                T acc = (T)0;
                                const int ix = gi[m] & 1023, iy = (gi[m] >> 10) & 1023,
                                          iz = gi[m] >> 20;
#pragma unroll
                                for (int r = 0; r < NR; ++r) {
                                    const T *I0 = sI + (r * 3) * C::GSZ;
                                    acc += I0[ix] * I0[C::GSZ + iy] * I0[2 * C::GSZ + iz];
                                }
                                if constexpr (GL)
                                    sg[lane + 64 * m] += acc;
                                else
                                    g[m] += acc;
                            }
                        }
                        wave_sync();
                        // primitive pairs sorted by |e12|, negligible ones
                        // zeroed: stop before a zero one (tested after the
                        // work, so single-primitive pairs never wait on it)
                        if (iq + 1 < tk.nprim_q && ue12[qq + 1] == 0.0) break;
                    }
                    if (ip + 1 < tk.nprim_p && ue12[p + 1] == 0.0) break;
                }
            }
            // phase C: digestion
            if constexpr (!GL) {
#pragma unroll
                for (int m = 0; m < W::NG; ++m)
                    if (lane + 64 * m < C::NF) sg[lane + 64 * m] = g[m];
            }
            if constexpr (PF) {
#pragma unroll
                for (int m = 0; m < W::ECD; ++m)
                    if (lane + 64 * m < W::NCD) sDcd[lane + 64 * m] = pcd[m];
#pragma unroll
                for (int m = 0; m < W::EBD; ++m)
                    if (lane + 64 * m < W::NBD) sDbd[lane + 64 * m] = pbd[m];
#pragma unroll
                for (int m = 0; m < W::EAD; ++m)
                    if (lane + 64 * m < W::NAD) sDad[lane + 64 * m] = pad[m];
#pragma unroll
                for (int m = 0; m < W::EBC; ++m)
                    if (lane + 64 * m < W::NBC) sDbc[lane + 64 * m] = pbc[m];
#pragma unroll
                for (int m = 0; m < W::EAC; ++m)
                    if (lane + 64 * m < W::NAC) sDac[lane + 64 * m] = pac[m];
            } else {
                for (int e = lane; e < W::NCD; e += 64)
                    sDcd[e] = D[(size_t)(c0 + e / ND) * nao + d0 + e % ND];
                for (int e = lane; e < W::NBD; e += 64)
                    sDbd[e] = D[(size_t)(d0 + e % ND) * nao + b0 + e / ND];
                for (int e = lane; e < W::NAD; e += 64)
                    sDad[e] = D[(size_t)(d0 + e % ND) * nao + a0 + e / ND];
                for (int e = lane; e < W::NBC; e += 64)
                    sDbc[e] = D[(size_t)(c0 + e % NC) * nao + b0 + e / NC];
                for (int e = lane; e < W::NAC; e += 64)
                    sDac[e] = D[(size_t)(c0 + e % NC) * nao + a0 + e / NC];
            }
            wave_sync();
            if (!skip) {
                if (VJ) {
#pragma unroll
                    for (int m = 0; m < SAB::PASSES; ++m) {
                        const int e = m * SAB::PER + lane / SAB::G, sl = lane % SAB::G;
                        const double v = split_sum<typename DT::Jab>(sg, sDcd, e, sl, e < W::NAB);
                        if (sl == 0) jab[m] += v;
                    }
#pragma unroll
                    for (int m = 0; m < SCD::PASSES; ++m) {
                        const int e = m * SCD::PER + lane / SCD::G, sl = lane % SCD::G;
                        const double v = split_sum<typename DT::Jcd>(sg, sDab, e, sl, e < W::NCD);
                        if (sl == 0 && e < W::NCD)
                            atomic_add(VJ + (size_t)(c0 + e / ND) * nao + d0 + e % ND, v);
                    }
                }
                if (VK) {
#pragma unroll
                    for (int m = 0; m < SAC::PASSES; ++m) {
                        const int e = m * SAC::PER + lane / SAC::G, sl = lane % SAC::G;
                        const double v = split_sum<typename DT::Kac>(sg, sDbd, e, sl, e < W::NAC);
                        if (sl == 0) kac[m] += v;
                    }
#pragma unroll
                    for (int m = 0; m < SBC::PASSES; ++m) {
                        const int e = m * SBC::PER + lane / SBC::G, sl = lane % SBC::G;
                        const double v = split_sum<typename DT::Kbc>(sg, sDad, e, sl, e < W::NBC);
                        if (sl == 0) kbc[m] += v;
                    }
#pragma unroll
                    for (int m = 0; m < SAD::PASSES; ++m) {
                        const int e = m * SAD::PER + lane / SAD::G, sl = lane % SAD::G;
                        const double v = split_sum<typename DT::Kad>(sg, sDbc, e, sl, e < W::NAD);
                        if (sl == 0 && e < W::NAD)
                            atomic_add(VK + (size_t)(d0 + e % ND) * nao + a0 + e / ND, v);
                    }
#pragma unroll
                    for (int m = 0; m < SBD::PASSES; ++m) {
                        const int e = m * SBD::PER + lane / SBD::G, sl = lane % SBD::G;
                        const double v = split_sum<typename DT::Kbd>(sg, sDac, e, sl, e < W::NBD);
                        if (sl == 0 && e < W::NBD)
                            atomic_add(VK + (size_t)(d0 + e % ND) * nao + b0 + e / ND, v);
                    }
                }
            }
            wave_sync();
            // end of a run of equal c: flush K(a,c), K(b,c)
            const int sc_next = q + 1 < q1 ? rny : -1;
            if (VK && sc_next != sc) {
#pragma unroll
                for (int m = 0; m < SAC::PASSES; ++m) {
                    const int e = m * SAC::PER + lane / SAC::G;
                    if (lane % SAC::G == 0 && e < W::NAC && kac[m] != 0.0)
                        atomic_add(VK + (size_t)(c0 + e % NC) * nao + a0 + e / NC, kac[m]);
                    kac[m] = 0.0;
                }
#pragma unroll
                for (int m = 0; m < SBC::PASSES; ++m) {
                    const int e = m * SBC::PER + lane / SBC::G;
                    if (lane % SBC::G == 0 && e < W::NBC && kbc[m] != 0.0)
                        atomic_add(VK + (size_t)(c0 + e % NC) * nao + b0 + e / NC, kbc[m]);
                    kbc[m] = 0.0;
                }
            }
            sc = sc_next;
            rqx = rnx;
            rqy = rny;
            rqz = rnz;
            rqw = rnw;
        }
        if (VJ) {
#pragma unroll
            for (int m = 0; m < SAB::PASSES; ++m) {
                const int e = m * SAB::PER + lane / SAB::G;
                if (lane % SAB::G == 0 && e < W::NAB && jab[m] != 0.0)
                    atomic_add(VJ + (size_t)(a0 + e / NB) * nao + b0 + e % NB, jab[m]);
            }
        }
        wave_sync();
    }
}

typedef RowKernel WaveKernel;
typedef WaveKernel WaveKernelTable[2][BQC_NPAIRID][BQC_NPAIRID];

// Classes digested with the wavefront-per-quartet kernel (LDS budget: up to
// (fd|fd) and (ff|fd), whose ~80 KB leave room for two workgroups per CU;
// the planner refuses a kernel above the device's per-workgroup LDS).
template <int LA, int LB, int LC, int LD>
constexpr bool use_wave() {
    using C = Cls<LA, LB, LC, LD>;
    return C::NF > 81 && (3 * C::NR * C::GSZ + C::NF) * 8 <= 80000;
}

template <int LA, int LB, int LC, int LD>
static void reg_wave(WaveKernelTable &tab) {
    if constexpr (use_wave<LA, LB, LC, LD>()) {
        tab[0][LA * (LA + 1) / 2 + LB][LC * (LC + 1) / 2 + LD] = jk_wave<LA, LB, LC, LD, double>;
        tab[1][LA * (LA + 1) / 2 + LB][LC * (LC + 1) / 2 + LD] = jk_wave<LA, LB, LC, LD, float>;
    }
}

template <int LA, int LB, int LC, int LD>
static void reg_class(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3) {
    reg_kernel<LA, LB, LC, LD>(t1);
    reg_row<LA, LB, LC, LD>(t2);
    reg_wave<LA, LB, LC, LD>(t3);
}

// all ket classes (LC LD) with pair-id(LC,LD) <= pair-id(LA,LB)
template <int LA, int LB>
static void reg_bra(JKKernelTable &t1, RowKernelTable &t2, WaveKernelTable &t3) {
    constexpr int pab = LA * (LA + 1) / 2 + LB;
    reg_class<LA, LB, 0, 0>(t1, t2, t3);
    if constexpr (pab >= 1) reg_class<LA, LB, 1, 0>(t1, t2, t3);
    if constexpr (pab >= 2) reg_class<LA, LB, 1, 1>(t1, t2, t3);
    if constexpr (pab >= 3) reg_class<LA, LB, 2, 0>(t1, t2, t3);
    if constexpr (pab >= 4) reg_class<LA, LB, 2, 1>(t1, t2, t3);
    if constexpr (pab >= 5) reg_class<LA, LB, 2, 2>(t1, t2, t3);
    if constexpr (pab >= 6) reg_class<LA, LB, 3, 0>(t1, t2, t3);
    if constexpr (pab >= 7) reg_class<LA, LB, 3, 1>(t1, t2, t3);
    if constexpr (pab >= 8) reg_class<LA, LB, 3, 2>(t1, t2, t3);
    if constexpr (pab >= 9) reg_class<LA, LB, 3, 3>(t1, t2, t3);
}
