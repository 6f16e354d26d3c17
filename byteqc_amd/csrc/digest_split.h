// Digestion targets of one (ab|cd) integral block and how the 64 lanes of a
// wavefront share them.
//
// The block g[x][y][z][w] (Cartesian components of a, b, c, d, flattened as
// ((x*NB + y)*NC + z)*ND + w) meets six density blocks:
//   J(a,b)[x,y] = sum_{z,w} g D(c,d)[z,w]      J(c,d)[z,w] = sum_{x,y} g D(a,b)[x,y]
//   K(a,c)[x,z] = sum_{y,w} g D(b,d)[y,w]      K(b,c)[y,z] = sum_{x,w} g D(a,d)[x,w]
//   K(a,d)[x,w] = sum_{y,z} g D(b,c)[y,z]      K(b,d)[y,w] = sum_{x,z} g D(a,c)[x,z]
// Every output element e is a double sum over an outer index u < NU and an
// inner index v < NV with fixed strides in g (SU, SV) and in its density
// block (DU, DV), starting at g offset gbase(e) and density offset 0; the
// density blocks are stored row-major in the order of their index pairs
// above (D(b,d) as [y][w], D(a,c) as [x][z], ...).
//
// G lanes share an element: lane s of the group takes u = s, s + G, ...,
// and the group's partial sums are added with lane shuffles.  G is chosen at
// compile time to minimise passes x (terms per lane + shuffle steps), so a
// target with few long sums (J(c,d) of a (dp|pp) block: 9 sums of 18 terms)
// keeps more lanes busy than one element per lane would.
//
// Everything here is constexpr so the host compiler can check the maps.
#pragma once

namespace bqc {

template <int NA, int NB, int NC, int ND>
struct DigestTargets {
    static constexpr int NCD_ = NC * ND, NBCD = NB * NC * ND;
    struct Jab {  // out (x,y); u = z, v = w; D(c,d)
        static constexpr int NOUT = NA * NB, NU = NC, NV = ND, SU = ND, SV = 1, DU = ND, DV = 1;
        static constexpr int gbase(int e) { return e * NCD_; }
    };
    struct Jcd {  // out (z,w); u = x, v = y; D(a,b)
        static constexpr int NOUT = NC * ND, NU = NA, NV = NB, SU = NBCD, SV = NCD_, DU = NB, DV = 1;
        static constexpr int gbase(int e) { return e; }
    };
    struct Kac {  // out (x,z); u = y, v = w; D(b,d)
        static constexpr int NOUT = NA * NC, NU = NB, NV = ND, SU = NCD_, SV = 1, DU = ND, DV = 1;
        static constexpr int gbase(int e) { return (e / NC) * NBCD + (e % NC) * ND; }
    };
    struct Kbc {  // out (y,z); u = x, v = w; D(a,d)
        static constexpr int NOUT = NB * NC, NU = NA, NV = ND, SU = NBCD, SV = 1, DU = ND, DV = 1;
        static constexpr int gbase(int e) { return e * ND; }
    };
    struct Kad {  // out (x,w); u = y, v = z; D(b,c)
        static constexpr int NOUT = NA * ND, NU = NB, NV = NC, SU = NCD_, SV = ND, DU = NC, DV = 1;
        static constexpr int gbase(int e) { return (e / ND) * NBCD + e % ND; }
    };
    struct Kbd {  // out (y,w); u = x, v = z; D(a,c)
        static constexpr int NOUT = NB * ND, NU = NA, NV = NC, SU = NBCD, SV = ND, DU = NC, DV = 1;
        static constexpr int gbase(int e) { return (e / ND) * NCD_ + e % ND; }
    };
};

constexpr int split_log2(int g) { return g <= 1 ? 0 : 1 + split_log2(g / 2); }

// Lanes per element for a target of NOUT sums of NU x NV terms: a term costs
// two LDS reads and an FMA (~3 issue slots), a shuffle step of a double ~4.
constexpr int split_lanes(int nout, int nu, int nv) {
    int best = 1, best_cost = 1 << 30;
    for (int g = 1; g <= 16 && (g == 1 || g <= nu); g *= 2) {
        const int per = 64 / g, passes = (nout + per - 1) / per, iu = (nu + g - 1) / g;
        const int cost = passes * (3 * iu * nv + 4 * split_log2(g));
        if (cost < best_cost) {
            best_cost = cost;
            best = g;
        }
    }
    return best;
}

template <class TG>
struct Split {
    static constexpr int G = split_lanes(TG::NOUT, TG::NU, TG::NV);
    static constexpr int PER = 64 / G;                        // elements per pass
    static constexpr int PASSES = (TG::NOUT + PER - 1) / PER;
    static constexpr int IU = (TG::NU + G - 1) / G;           // outer terms per lane
};

}  // namespace bqc
