// Host check of the lane maps in byteqc_amd/csrc/digest_split.h: for every
// (la lb | lc ld) up to f, the lane-group sums of each digestion target equal
// the direct four-index sums, every element is produced once and every
// block / density index stays inside its array.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../byteqc_amd/csrc/digest_split.h"

using namespace bqc;

static int failures = 0;

// Emulates 64 lanes running split_sum for target TG; out[e] from lane s == 0.
template <class TG>
static void lanes(const std::vector<double> &g, const std::vector<double> &d, std::vector<double> &out,
                  std::vector<int> &hits, int nf) {
    using S = Split<TG>;
    out.assign(TG::NOUT, 0.0);
    hits.assign(TG::NOUT, 0);
    for (int p = 0; p < S::PASSES; ++p) {
        double part[64];
        for (int lane = 0; lane < 64; ++lane) {
            const int e = p * S::PER + lane / S::G, s = lane % S::G;
            double sum = 0.0;
            if (e < TG::NOUT) {
                for (int i = 0; i < S::IU; ++i) {
                    const int u = s + S::G * i;
                    if (u >= TG::NU) continue;
                    for (int v = 0; v < TG::NV; ++v) {
                        const int gi = TG::gbase(e) + u * TG::SU + v * TG::SV;
                        const int di = u * TG::DU + v * TG::DV;
                        if (gi < 0 || gi >= nf || di < 0 || di >= TG::NU * TG::NV) {
                            std::printf("index out of range\n");
                            ++failures;
                            return;
                        }
                        sum += g[gi] * d[di];
                    }
                }
            }
            part[lane] = sum;
        }
        // butterfly over the group, as the lane shuffles do
        for (int o = 1; o < S::G; o <<= 1) {
            double nxt[64];
            for (int lane = 0; lane < 64; ++lane) nxt[lane] = part[lane] + part[lane ^ o];
            for (int lane = 0; lane < 64; ++lane) part[lane] = nxt[lane];
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int e = p * S::PER + lane / S::G;
            if (lane % S::G == 0 && e < TG::NOUT) {
                out[e] = part[lane];
                hits[e]++;
            }
        }
    }
}

static bool close(const std::vector<double> &a, const std::vector<double> &b) {
    for (size_t i = 0; i < a.size(); ++i)
        if (std::fabs(a[i] - b[i]) > 1e-12 * (1.0 + std::fabs(b[i]))) return false;
    return true;
}

template <int NA, int NB, int NC, int ND>
static void check(std::mt19937 &rng) {
    using DT = DigestTargets<NA, NB, NC, ND>;
    std::uniform_real_distribution<double> U(-1, 1);
    const int nf = NA * NB * NC * ND;
    std::vector<double> g(nf), dab(NA * NB), dcd(NC * ND), dbd(NB * ND), dad(NA * ND),
        dbc(NB * NC), dac(NA * NC);
    for (auto *v : {&g, &dab, &dcd, &dbd, &dad, &dbc, &dac})
        for (double &x : *v) x = U(rng);
    auto G = [&](int x, int y, int z, int w) { return g[((x * NB + y) * NC + z) * ND + w]; };
    std::vector<double> jab(NA * NB, 0), jcd(NC * ND, 0), kac(NA * NC, 0), kbc(NB * NC, 0),
        kad(NA * ND, 0), kbd(NB * ND, 0);
    for (int x = 0; x < NA; ++x)
        for (int y = 0; y < NB; ++y)
            for (int z = 0; z < NC; ++z)
                for (int w = 0; w < ND; ++w) {
                    const double v = G(x, y, z, w);
                    jab[x * NB + y] += v * dcd[z * ND + w];
                    jcd[z * ND + w] += v * dab[x * NB + y];
                    kac[x * NC + z] += v * dbd[y * ND + w];
                    kbc[y * NC + z] += v * dad[x * ND + w];
                    kad[x * ND + w] += v * dbc[y * NC + z];
                    kbd[y * ND + w] += v * dac[x * NC + z];
                }
    std::vector<double> out;
    std::vector<int> hits;
    auto verify = [&](const char *name, const std::vector<double> &ref) {
        bool once = true;
        for (int h : hits) once = once && h == 1;
        if (!close(out, ref) || !once) {
            std::printf("mismatch %s for (%d %d | %d %d)\n", name, NA, NB, NC, ND);
            ++failures;
        }
    };
    lanes<typename DT::Jab>(g, dcd, out, hits, nf);
    verify("J(a,b)", jab);
    lanes<typename DT::Jcd>(g, dab, out, hits, nf);
    verify("J(c,d)", jcd);
    lanes<typename DT::Kac>(g, dbd, out, hits, nf);
    verify("K(a,c)", kac);
    lanes<typename DT::Kbc>(g, dad, out, hits, nf);
    verify("K(b,c)", kbc);
    lanes<typename DT::Kad>(g, dbc, out, hits, nf);
    verify("K(a,d)", kad);
    lanes<typename DT::Kbd>(g, dac, out, hits, nf);
    verify("K(b,d)", kbd);
}

template <int LA, int LB, int LC, int LD>
static void cls(std::mt19937 &rng) {
    constexpr auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
    check<nc(LA), nc(LB), nc(LC), nc(LD)>(rng);
}

template <int LA, int LB>
static void bra(std::mt19937 &rng) {
    cls<LA, LB, 0, 0>(rng);
    cls<LA, LB, 1, 0>(rng);
    cls<LA, LB, 1, 1>(rng);
    cls<LA, LB, 2, 0>(rng);
    cls<LA, LB, 2, 1>(rng);
    cls<LA, LB, 2, 2>(rng);
    cls<LA, LB, 3, 0>(rng);
    cls<LA, LB, 3, 1>(rng);
    cls<LA, LB, 3, 2>(rng);
    cls<LA, LB, 3, 3>(rng);
}

int main() {
    std::mt19937 rng(7);
    bra<0, 0>(rng);
    bra<1, 0>(rng);
    bra<1, 1>(rng);
    bra<2, 0>(rng);
    bra<2, 1>(rng);
    bra<2, 2>(rng);
    bra<3, 0>(rng);
    bra<3, 1>(rng);
    bra<3, 2>(rng);
    bra<3, 3>(rng);
    // the split pays where the sums are long: J(c,d) of (dp|pp) uses 4 lanes
    using T = DigestTargets<6, 3, 3, 3>;
    if (Split<T::Jcd>::G != 4 || Split<T::Jab>::G != 2) {
        std::printf("unexpected lane split for (dp|pp)\n");
        ++failures;
    }
    std::printf("%s\n", failures ? "FAIL" : "OK");
    return failures ? 1 : 0;
}
