"""Pins for the CPU oracle (oracle/): every oracle function is checked here
against mathematics it must satisfy, independent of how it is coded.

* Boys / Rys: 80-digit mpmath vectors (tests/golden/rys_golden.npz), the
  moment identities sum_i w_i s_i^k = F_k(x), and the x = 0 limit, where the
  Rys rule is the positive half of Gauss-Legendre of order 2n.
* Primitive ERIs: an independent Obara-Saika recursion in mpmath, the
  closed-form (ss|ss), and 8-fold permutational symmetry.
* Plan + J/K assembly (unique quartets, pair and quartet weights, bins,
  symmetrisation, back-transform): brute force from the full ERI tensor.
"""
import itertools
import math
import os

import numpy
import pytest

from oracle import oracle as orc
from byteqc_amd.gto import M, gto_norm
from tests import os_eri

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'rys_golden.npz')


@pytest.fixture(scope='module')
def gold():
    return numpy.load(GOLD)


def test_boys_matches_mpmath(gold):
    for ix, x in enumerate(gold['x']):
        f = orc.boys(19, x)
        assert numpy.allclose(f, gold['boys'][ix], rtol=1e-14, atol=0), x


@pytest.mark.parametrize("n", range(1, 12))
def test_rys_matches_mpmath(gold, n):
    for ix, x in enumerate(gold['x']):
        u, w = orc.rys(n, x)
        s = u / (1 + u)
        assert numpy.allclose(s, gold[f's_{n}'][ix], rtol=1e-14, atol=0), (n, x)
        assert numpy.allclose(w, gold[f'w_{n}'][ix], rtol=1e-14, atol=0), (n, x)


@pytest.mark.parametrize('n', [1, 3, 5, 7, 9, 11])
def test_rys_reproduces_boys_moments(n):
    rng = numpy.random.default_rng(n)
    for x in numpy.concatenate([rng.uniform(0, 100, 30), [0., 1e-9, 47.5, 500.]]):
        u, w = orc.rys(n, x)
        s = u / (1 + u)
        f = orc.boys(2 * n - 1, x)
        for k in range(2 * n):
            assert abs(numpy.sum(w * s ** k) - f[k]) <= 1e-13 * f[k], (n, x, k)


@pytest.mark.parametrize('n', range(1, 12))
def test_rys_small_x_is_gauss_legendre(n):
    t, wt = numpy.polynomial.legendre.leggauss(2 * n)
    u, w = orc.rys(n, 0.)
    s = u / (1 + u)
    assert numpy.allclose(s, numpy.sort(t[t > 0] ** 2), rtol=1e-14)
    order = numpy.argsort(t[t > 0] ** 2)
    assert numpy.allclose(w, wt[t > 0][order], rtol=1e-14)


def _prim_shell_mol(shells):
    """One primitive shell per 'atom' (a ghost H at each centre)."""
    atoms = [('H', tuple(r)) for _, _, r in shells]
    mol = M(atom=atoms, basis='cc-pvdz', unit='Bohr')
    env = list(mol._env[:20 + 4 * len(shells)])
    bas = []
    for ia, (l, a, _) in enumerate(shells):
        pe = len(env)
        env += [a, float(gto_norm(l, a))]
        bas.append([ia, l, 1, 1, 0, pe, pe + 1, 0])
    return (mol._atm, numpy.asarray(bas, dtype=numpy.int32),
            numpy.asarray(env))


def _fac_sp(l):
    return {0: 0.282094791773878143, 1: 0.488602511902919921}.get(l, 1.)


@pytest.mark.parametrize('ls', [(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 1, 0), (2, 0, 1, 1),
                                (2, 2, 1, 0), (2, 2, 2, 2), (3, 1, 2, 0), (3, 3, 2, 1),
                                (3, 3, 3, 3), (0, 2, 1, 3)])
def test_eri_matches_obara_saika(ls):
    rng = numpy.random.default_rng(sum(l * 7 ** i for i, l in enumerate(ls)))
    shells = [(l, float(rng.uniform(0.2, 3.0)), rng.uniform(-1.5, 1.5, 3)) for l in ls]
    atm, bas, env = _prim_shell_mol(shells)
    got = orc.eri_shells(0, 1, 2, 3, atm, bas, env)
    ref = os_eri.primitive_eri(*ls, *[s[1] for s in shells], *[s[2] for s in shells])
    scale = 1.
    for l, a, _ in shells:
        scale *= _fac_sp(l) * float(gto_norm(l, a))
    ref = numpy.array(ref, dtype=float) * scale
    assert got.shape == ref.shape
    # block-relative: HRR cancellation leaves ~1e-13 of the block maximum on
    # elements that are orders of magnitude smaller than it
    assert numpy.abs(got - ref).max() <= 2e-13 * numpy.abs(ref).max()


@pytest.mark.parametrize('ls', [(4, 0, 0, 0), (4, 1, 2, 0), (5, 0, 1, 0), (4, 4, 0, 0),
                                (5, 2, 1, 1), (3, 4, 0, 5)])
def test_eri_matches_obara_saika_g_h_shells(ls):
    """g and h shells (8-11 Rys roots in the larger classes), whole blocks."""
    rng = numpy.random.default_rng(97 + sum(l * 7 ** i for i, l in enumerate(ls)))
    shells = [(l, float(rng.uniform(0.3, 2.0)), rng.uniform(-1.2, 1.2, 3)) for l in ls]
    atm, bas, env = _prim_shell_mol(shells)
    got = orc.eri_shells(0, 1, 2, 3, atm, bas, env)
    ref = os_eri.primitive_eri(*ls, *[s[1] for s in shells], *[s[2] for s in shells])
    scale = 1.
    for l, a, _ in shells:
        scale *= _fac_sp(l) * float(gto_norm(l, a))
    ref = numpy.array(ref, dtype=float) * scale
    assert got.shape == ref.shape
    assert numpy.abs(got - ref).max() <= 1e-12 * numpy.abs(ref).max()


@pytest.mark.parametrize('ls', [(4, 4, 4, 4), (5, 5, 5, 5), (5, 4, 5, 3)])
def test_eri_matches_obara_saika_largest_classes(ls):
    """(gg|gg), (hh|hh), (hg|hf): 9-11 roots; a random sample of components
    (the mpmath recursion over every component would take minutes)."""
    rng = numpy.random.default_rng(11 + sum(ls))
    shells = [(l, float(rng.uniform(0.4, 1.8)), rng.uniform(-1.0, 1.0, 3)) for l in ls]
    atm, bas, env = _prim_shell_mol(shells)
    got = orc.eri_shells(0, 1, 2, 3, atm, bas, env)
    nf = [(l + 1) * (l + 2) // 2 for l in ls]
    subset = [tuple(int(rng.integers(n)) for n in nf) for _ in range(12)]
    ref = os_eri.primitive_eri(*ls, *[s[1] for s in shells], *[s[2] for s in shells],
                               subset=subset)
    scale = 1.
    for l, a, _ in shells:
        scale *= _fac_sp(l) * float(gto_norm(l, a))
    ref = numpy.array([float(v) for v in ref]) * scale
    vals = numpy.array([got[q] for q in subset])
    big = numpy.abs(got).max()
    assert numpy.abs(vals - ref).max() <= 1e-12 * big


def test_ssss_closed_form():
    rng = numpy.random.default_rng(7)
    for _ in range(20):
        a = rng.uniform(0.05, 8, 4)
        r = rng.uniform(-1.5, 1.5, (4, 3))
        atm, bas, env = _prim_shell_mol([(0, a[i], r[i]) for i in range(4)])
        got = orc.eri_shells(0, 1, 2, 3, atm, bas, env).ravel()[0]
        p, q = a[0] + a[1], a[2] + a[3]
        P = (a[0] * r[0] + a[1] * r[1]) / p
        Q = (a[2] * r[2] + a[3] * r[3]) / q
        kab = math.exp(-a[0] * a[1] / p * numpy.sum((r[0] - r[1]) ** 2))
        kcd = math.exp(-a[2] * a[3] / q * numpy.sum((r[2] - r[3]) ** 2))
        t = p * q / (p + q) * numpy.sum((P - Q) ** 2)
        f0 = 1. if t == 0 else 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * kab * kcd * f0
        for i in range(4):
            ref *= _fac_sp(0) * float(gto_norm(0, a[i]))
        assert abs(got - ref) <= 1e-13 * abs(ref)


def test_eri_permutation_symmetry():
    rng = numpy.random.default_rng(3)
    ls = (2, 1, 3, 0)
    shells = [(l, float(rng.uniform(0.3, 2.0)), rng.uniform(-1, 1, 3)) for l in ls]
    atm, bas, env = _prim_shell_mol(shells)
    ref = orc.eri_shells(0, 1, 2, 3, atm, bas, env)
    for perm, axes in [((1, 0, 2, 3), (1, 0, 2, 3)), ((0, 1, 3, 2), (0, 1, 3, 2)),
                       ((2, 3, 0, 1), (2, 3, 0, 1)), ((3, 2, 1, 0), (3, 2, 1, 0))]:
        got = orc.eri_shells(*perm, atm, bas, env)
        assert numpy.allclose(got.transpose(numpy.argsort(axes)), ref, rtol=1e-13,
                              atol=1e-15)


def _bruteforce_jk(mol, dm):
    bas, env, coeff = orc.seg_contract(mol)
    eri = orc.eri_tensor(numpy.asarray(mol._atm), bas, env)
    dmc = coeff @ dm @ coeff.T
    j, k = orc.jk_from_tensor(eri, dmc)
    return coeff.T @ j @ coeff, coeff.T @ k @ coeff


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-svp'])
def test_jk_matches_full_tensor(basis):
    mol = M(atom='O 0 0 0.58589194; H 0 0.75668992 0; H 0 -0.75668992 0', basis=basis)
    rng = numpy.random.default_rng(11)
    a = rng.standard_normal((mol.nao, mol.nao))
    dm = a + a.T
    jref, kref = _bruteforce_jk(mol, dm)
    j, k = orc.get_jk(mol, dm)
    assert numpy.abs(j - jref).max() < 1e-11
    assert numpy.abs(k - kref).max() < 1e-11
    # J only / K only paths
    j2, _ = orc.get_jk(mol, dm, with_k=False)
    _, k2 = orc.get_jk(mol, dm, with_j=False)
    assert numpy.abs(j2 - j).max() < 1e-13 and numpy.abs(k2 - k).max() < 1e-13


def test_seg_contract_reproduces_contracted_overlap():
    """C^T S_c C (decontracted Cartesian overlap mapped back) has unit
    diagonal: the spherical contracted AOs are normalised."""
    mol = M(atom='O 0 0 0; H 0 0.8 0.6; H 0 -0.8 0.6', basis='cc-pvtz')
    bas, env, coeff = orc.seg_contract(mol)
    s = _cart_overlap(numpy.asarray(mol._atm), bas, env)
    ssph = coeff.T @ s @ coeff
    assert numpy.allclose(numpy.diag(ssph), 1., atol=1e-12)


def _cart_overlap(atm, bas, env):
    """Cartesian overlap of primitive shells (closed-form 1-D Gaussian moments)."""
    ao_loc = orc.cart_ao_loc(bas)
    n = ao_loc[-1]
    s = numpy.zeros((n, n))

    def comps(l):
        return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]

    def one_d(i, j, a, b, xa, xb):
        p = a + b
        xp = (a * xa + b * xb) / p
        tot = 0.
        for u in range(i + 1):
            for v in range(j + 1):
                t = u + v
                if t % 2:
                    continue
                mom = math.gamma((t + 1) / 2) / p ** ((t + 1) / 2)
                tot += math.comb(i, u) * math.comb(j, v) * (xp - xa) ** (i - u) * \
                    (xp - xb) ** (j - v) * mom
        return tot

    for I in range(len(bas)):
        for J in range(len(bas)):
            li, lj = bas[I, 1], bas[J, 1]
            ai, aj = env[bas[I, 5]], env[bas[J, 5]]
            ci, cj = env[bas[I, 6]], env[bas[J, 6]]
            ri = env[atm[bas[I, 0], 1]:atm[bas[I, 0], 1] + 3]
            rj = env[atm[bas[J, 0], 1]:atm[bas[J, 0], 1] + 3]
            kab = math.exp(-ai * aj / (ai + aj) * numpy.sum((ri - rj) ** 2))
            for x, ci_ in enumerate(comps(li)):
                for y, cj_ in enumerate(comps(lj)):
                    v = kab * ci * cj * _fac_sp(li) * _fac_sp(lj)
                    for d in range(3):
                        v *= one_d(ci_[d], cj_[d], ai, aj, ri[d], rj[d])
                    s[ao_loc[I] + x, ao_loc[J] + y] = v
    return s


def test_oracle_rhf_known_answer():
    """RHF/cc-pVDZ of water at O 0 0 0; H 0 -+0.757 0.587 (Angstrom):
    -76.0267656731 Eh is the energy PySCF's examples print for this input; it
    pins the transcribed basis, the normalisation and Cartesian-to-spherical
    conventions, the one-electron integrals and the J/K path together."""
    from oracle import scf_oracle as so
    mol = M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587', basis='cc-pvdz')
    e, _ = so.rhf(mol)
    assert abs(e - (-76.0267656731)) < 2e-10


def test_oracle_int1e_normalised_and_symmetric():
    from oracle import scf_oracle as so
    mol = M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587', basis='def2-svp')
    s, t, v = so.int1e(mol)
    assert numpy.allclose(numpy.diag(s), 1, atol=1e-13)
    for m in (s, t, v):
        assert numpy.allclose(m, m.T, atol=1e-14)
    assert numpy.all(numpy.linalg.eigvalsh(t) > 0)
