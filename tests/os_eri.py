"""Independent primitive ERIs by the Obara-Saika / Head-Gordon-Pople recursion.

Used only to pin the oracle's Rys-quadrature integrals.  Works in mpmath
(40 digits) with unnormalised Cartesian Gaussians x^a y^b z^c exp(-alpha r^2):

  (00|00)^(m) = 2 pi^(5/2) / (zeta eta sqrt(zeta+eta)) K_AB K_CD F_m(T)
  vertical recursion on a (bra) and c (ket) in the (a0|c0)^(m) integrals,
  then horizontal transfer (a, b+1_i| = (a+1_i, b| + (A-B)_i (a, b| and the
  same on the ket.
"""
from functools import lru_cache

import mpmath

mpmath.mp.dps = 40


def boys(m, t):
    t = mpmath.mpf(t)
    if t == 0:
        return mpmath.mpf(1) / (2 * m + 1)
    a = m + mpmath.mpf(1) / 2
    return mpmath.gammainc(a, 0, t) / (2 * t ** a)


def primitive_eri(la, lb, lc, ld, alpha, beta, gamma, delta, A, B, C, D, subset=None):
    """Nested lists [a][b][c][d] of the integrals over all Cartesian
    components (storage order), or, with ``subset`` a list of component index
    quartets (x, y, z, w), the list of those integrals only."""
    mp = mpmath.mpf
    alpha, beta, gamma, delta = map(mp, (alpha, beta, gamma, delta))
    A, B, C, D = ([mp(x) for x in v] for v in (A, B, C, D))
    zeta = alpha + beta
    eta = gamma + delta
    P = [(alpha * A[i] + beta * B[i]) / zeta for i in range(3)]
    Q = [(gamma * C[i] + delta * D[i]) / eta for i in range(3)]
    rho = zeta * eta / (zeta + eta)
    W = [(zeta * P[i] + eta * Q[i]) / (zeta + eta) for i in range(3)]
    ab2 = sum((A[i] - B[i]) ** 2 for i in range(3))
    cd2 = sum((C[i] - D[i]) ** 2 for i in range(3))
    pq2 = sum((P[i] - Q[i]) ** 2 for i in range(3))
    kab = mpmath.exp(-alpha * beta / zeta * ab2)
    kcd = mpmath.exp(-gamma * delta / eta * cd2)
    pre = 2 * mpmath.pi ** mpmath.mpf(2.5) / (zeta * eta * mpmath.sqrt(zeta + eta)) * kab * kcd
    T = rho * pq2
    mmax = la + lb + lc + ld
    fm = [pre * boys(m, T) for m in range(mmax + 1)]

    @lru_cache(maxsize=None)
    def vrr(a, c, m):
        if any(x < 0 for x in a) or any(x < 0 for x in c):
            return mp(0)
        if sum(a) == 0 and sum(c) == 0:
            return fm[m]
        if sum(a) > 0:
            i = next(k for k in range(3) if a[k] > 0)
            a1 = tuple(a[k] - (k == i) for k in range(3))
            a2 = tuple(a1[k] - (k == i) for k in range(3))
            c1 = tuple(c[k] - (k == i) for k in range(3))
            v = (P[i] - A[i]) * vrr(a1, c, m) + (W[i] - P[i]) * vrr(a1, c, m + 1)
            if a1[i] > 0:
                v += a1[i] / (2 * zeta) * (vrr(a2, c, m) - rho / zeta * vrr(a2, c, m + 1))
            if c[i] > 0:
                v += c[i] / (2 * (zeta + eta)) * vrr(a1, c1, m + 1)
            return v
        i = next(k for k in range(3) if c[k] > 0)
        c1 = tuple(c[k] - (k == i) for k in range(3))
        c2 = tuple(c1[k] - (k == i) for k in range(3))
        v = (Q[i] - C[i]) * vrr(a, c1, m) + (W[i] - Q[i]) * vrr(a, c1, m + 1)
        if c1[i] > 0:
            v += c1[i] / (2 * eta) * (vrr(a, c2, m) - rho / eta * vrr(a, c2, m + 1))
        return v

    @lru_cache(maxsize=None)
    def hrr(a, b, c, d):
        if sum(b) > 0:
            i = next(k for k in range(3) if b[k] > 0)
            b1 = tuple(b[k] - (k == i) for k in range(3))
            a1 = tuple(a[k] + (k == i) for k in range(3))
            return hrr(a1, b1, c, d) + (A[i] - B[i]) * hrr(a, b1, c, d)
        if sum(d) > 0:
            i = next(k for k in range(3) if d[k] > 0)
            d1 = tuple(d[k] - (k == i) for k in range(3))
            c1 = tuple(c[k] + (k == i) for k in range(3))
            return hrr(a, b, c1, d1) + (C[i] - D[i]) * hrr(a, b, c, d1)
        return vrr(a, c, 0)

    def comps(l):
        return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]

    if subset is not None:
        ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
        return [hrr(ca[x], cb[y], cc[z], cd[w]) for x, y, z, w in subset]
    return [[[[hrr(a, b, c, d) for d in comps(ld)] for c in comps(lc)]
             for b in comps(lb)] for a in comps(la)]
