"""CPU oracle for the SCF caller of the J/K path -- TEST INFRASTRUCTURE ONLY.

Only tests may import this module.  It provides, for small molecules:

* one-electron integrals S, T, V by the Obara-Saika recursions (overlap and
  kinetic from 1-D Gaussian products; nuclear attraction from the auxiliary
  integrals (a|A|b)^(m) with Boys functions F_m from the oracle library), in
  the decontracted Cartesian basis of ``oracle.seg_contract`` mapped to the
  spherical contracted AOs -- the integrals PySCF/libcint supplies to the
  reference's SCF loop (cuobc/scf/hf.py:272-274 keeps PySCF's hcore/ovlp);
* a plain RHF loop (core guess, Pulay DIIS) on top of ``oracle.get_jk``.
"""
from functools import lru_cache
import math

import numpy

from oracle import oracle as orc

FAC_SP = {0: 0.282094791773878143, 1: 0.488602511902919921}


def _comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _prim_1e(la, lb, a, b, A, B, charges, centres):
    """Primitive Cartesian S, T, V blocks (unnormalised Gaussians)."""
    p = a + b
    P = (a * A + b * B) / p
    kab = math.exp(-a * b / p * float(numpy.sum((A - B) ** 2)))

    def s1(i, j, d):
        # 1-D overlap by Obara-Saika
        @lru_cache(maxsize=None)
        def s(i, j):
            if i < 0 or j < 0:
                return 0.
            if i == 0 and j == 0:
                return math.sqrt(math.pi / p)
            if i > 0:
                return (P[d] - A[d]) * s(i - 1, j) + ((i - 1) * s(i - 2, j) + j * s(i - 1, j - 1)) / (2 * p)
            return (P[d] - B[d]) * s(i, j - 1) + (i * s(i - 1, j - 1) + (j - 1) * s(i, j - 2)) / (2 * p)
        return s(i, j)

    def t1(i, j, d):
        v = b * (2 * j + 1) * s1(i, j, d) - 2 * b * b * s1(i, j + 2, d)
        if j >= 2:
            v -= 0.5 * j * (j - 1) * s1(i, j - 2, d)
        return v

    ca, cb = _comps(la), _comps(lb)
    S = numpy.zeros((len(ca), len(cb)))
    T = numpy.zeros_like(S)
    for x, ea in enumerate(ca):
        for y, eb in enumerate(cb):
            sd = [s1(ea[d], eb[d], d) for d in range(3)]
            td = [t1(ea[d], eb[d], d) for d in range(3)]
            S[x, y] = kab * sd[0] * sd[1] * sd[2]
            T[x, y] = kab * (td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2] + sd[0] * sd[1] * td[2])
    V = numpy.zeros_like(S)
    mmax = la + lb
    for z, C in zip(charges, centres):
        if z == 0:
            continue
        PC = P - C
        f = orc.boys(mmax, p * float(numpy.sum(PC ** 2)))
        pre = 2 * math.pi / p * kab

        @lru_cache(maxsize=None)
        def th(ea, eb, m):
            if min(ea) < 0 or min(eb) < 0:
                return 0.
            if sum(ea) == 0 and sum(eb) == 0:
                return pre * f[m]
            if sum(ea) > 0:
                i = next(k for k in range(3) if ea[k] > 0)
                a1 = tuple(ea[k] - (k == i) for k in range(3))
                a2 = tuple(a1[k] - (k == i) for k in range(3))
                b1 = tuple(eb[k] - (k == i) for k in range(3))
                v = (P[i] - A[i]) * th(a1, eb, m) - PC[i] * th(a1, eb, m + 1)
                if a1[i] > 0:
                    v += a1[i] / (2 * p) * (th(a2, eb, m) - th(a2, eb, m + 1))
                if eb[i] > 0:
                    v += eb[i] / (2 * p) * (th(a1, b1, m) - th(a1, b1, m + 1))
                return v
            i = next(k for k in range(3) if eb[k] > 0)
            b1 = tuple(eb[k] - (k == i) for k in range(3))
            b2 = tuple(b1[k] - (k == i) for k in range(3))
            v = (P[i] - B[i]) * th(ea, b1, m) - PC[i] * th(ea, b1, m + 1)
            if b1[i] > 0:
                v += b1[i] / (2 * p) * (th(ea, b2, m) - th(ea, b2, m + 1))
            return v

        for x, ea in enumerate(ca):
            for y, eb in enumerate(cb):
                V[x, y] -= z * th(tuple(ea), tuple(eb), 0)
    return S, T, V


def int1e(mol):
    """S, T, V in the spherical contracted AO basis of mol."""
    bas, env, coeff = orc.seg_contract(mol)
    atm = numpy.asarray(mol._atm)
    ao_loc = orc.cart_ao_loc(bas)
    n = ao_loc[-1]
    out = [numpy.zeros((n, n)) for _ in range(3)]
    coords = mol.atom_coords()
    charges = mol.atom_charges()
    for I in range(len(bas)):
        for J in range(I + 1):
            li, lj = int(bas[I, 1]), int(bas[J, 1])
            ai, aj = env[bas[I, 5]:bas[I, 5] + bas[I, 2]], env[bas[J, 5]:bas[J, 5] + bas[J, 2]]
            ci, cj = env[bas[I, 6]:bas[I, 6] + bas[I, 2]], env[bas[J, 6]:bas[J, 6] + bas[J, 2]]
            A = env[atm[bas[I, 0], 1]:atm[bas[I, 0], 1] + 3]
            B = env[atm[bas[J, 0], 1]:atm[bas[J, 0], 1] + 3]
            blk = [0., 0., 0.]
            for p, a in enumerate(ai):
                for q, b in enumerate(aj):
                    s, t, v = _prim_1e(li, lj, a, b, A, B, charges, coords)
                    c = ci[p] * cj[q] * FAC_SP.get(li, 1.) * FAC_SP.get(lj, 1.)
                    blk = [blk[0] + c * s, blk[1] + c * t, blk[2] + c * v]
            for m, b_ in zip(out, blk):
                m[ao_loc[I]:ao_loc[I + 1], ao_loc[J]:ao_loc[J + 1]] = b_
                m[ao_loc[J]:ao_loc[J + 1], ao_loc[I]:ao_loc[I + 1]] = b_.T
    return tuple(coeff.T @ m @ coeff for m in out)


def rhf(mol, conv_tol=1e-10, max_cycle=100):
    """Plain RHF with Pulay DIIS; returns (e_tot, dm)."""
    s, t, v = int1e(mol)
    h = t + v
    e, u = numpy.linalg.eigh(s)
    x = u[:, e > 1e-9] / numpy.sqrt(e[e > 1e-9])
    nocc = mol.nelectron // 2
    plan = orc.build_plan(mol)

    def dm_of(f):
        _, c = numpy.linalg.eigh(x.T @ f @ x)
        c = x @ c[:, :nocc]
        return 2 * c @ c.T

    dm = dm_of(h)
    fs, es, e_old = [], [], None
    for _ in range(max_cycle):
        j, k = orc.get_jk(mol, dm, plan)
        f = h + j - 0.5 * k
        e_tot = 0.5 * numpy.sum(dm * (h + f)) + mol.energy_nuc()
        err = x.T @ (f @ dm @ s - s @ dm @ f) @ x
        if e_old is not None and abs(e_tot - e_old) < conv_tol and abs(err).max() < 1e-7:
            return e_tot, dm
        e_old = e_tot
        fs.append(f)
        es.append(err)
        fs, es = fs[-8:], es[-8:]
        n = len(fs)
        if n > 1:
            b = -numpy.ones((n + 1, n + 1))
            b[n, n] = 0
            b[:n, :n] = [[numpy.sum(a * c) for c in es] for a in es]
            rhs = numpy.zeros(n + 1)
            rhs[n] = -1
            cdiis = numpy.linalg.lstsq(b, rhs, rcond=None)[0]
            f = sum(c * fi for c, fi in zip(cdiis[:n], fs))
        dm = dm_of(f)
    raise RuntimeError('oracle RHF did not converge')
