"""CPU tests of the host side: basis/Mole conventions, the plan helpers that
mirror the reference's Python (checked against the oracle's independent
restatement), the Rys table data, and the C-ABI library's exports."""
import ctypes
import math
import os
import re

import numpy
import pytest

from byteqc_amd.gto import M, cart2sph, gto_norm, load_basis
from byteqc_amd import rys_table
from byteqc_amd.scf import hf
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cart2sph_matches_libcint_constants():
    """Real solid harmonics in libcint's convention (no Condon-Shortley
    phase, m = -l..l), checked against the closed-form normalisation
    constants sqrt(15/4pi), sqrt(5/16pi), sqrt(15/16pi), sqrt(35/32pi),
    sqrt(105/4pi), sqrt(21/32pi), sqrt(7/16pi), sqrt(105/16pi)."""
    pi = math.pi
    d = cart2sph(2)  # rows xx xy xz yy yz zz
    assert d[1, 0] == pytest.approx(math.sqrt(15 / (4 * pi)))          # xy
    assert d[4, 1] == pytest.approx(math.sqrt(15 / (4 * pi)))          # yz
    assert d[5, 2] == pytest.approx(2 * math.sqrt(5 / (16 * pi)))      # zz
    assert d[0, 2] == pytest.approx(-math.sqrt(5 / (16 * pi)))
    assert d[2, 3] == pytest.approx(math.sqrt(15 / (4 * pi)))          # xz
    assert d[0, 4] == pytest.approx(math.sqrt(15 / (16 * pi)))         # xx - yy
    assert d[3, 4] == pytest.approx(-math.sqrt(15 / (16 * pi)))
    f = cart2sph(3)  # rows xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    c = math.sqrt(35 / (32 * pi))
    assert f[1, 0] == pytest.approx(3 * c) and f[6, 0] == pytest.approx(-c)
    assert f[4, 1] == pytest.approx(math.sqrt(105 / (4 * pi)))
    c = math.sqrt(21 / (32 * pi))
    assert f[8, 2] == pytest.approx(4 * c) and f[1, 2] == pytest.approx(-c)
    c = math.sqrt(7 / (16 * pi))
    assert f[9, 3] == pytest.approx(2 * c) and f[2, 3] == pytest.approx(-3 * c)
    c = math.sqrt(105 / (16 * pi))
    assert f[2, 5] == pytest.approx(c) and f[7, 5] == pytest.approx(-c)
    c = math.sqrt(35 / (32 * pi))
    assert f[0, 6] == pytest.approx(c) and f[3, 6] == pytest.approx(-3 * c)


@pytest.mark.parametrize('l', [2, 3])
def test_cart2sph_columns_are_orthonormal_harmonics(l):
    """On the unit sphere the columns are orthonormal (quadrature over a
    product Gauss-Legendre x uniform-phi grid)."""
    from byteqc_amd.gto import cart_components
    t, wt = numpy.polynomial.legendre.leggauss(20)
    phi = numpy.linspace(0, 2 * math.pi, 40, endpoint=False)
    ct, ph = numpy.meshgrid(t, phi, indexing='ij')
    st = numpy.sqrt(1 - ct ** 2)
    x, y, z = st * numpy.cos(ph), st * numpy.sin(ph), ct
    w = numpy.outer(wt, numpy.full(phi.size, 2 * math.pi / phi.size))
    mono = numpy.stack([x ** a * y ** b * z ** c for a, b, c in cart_components(l)], -1)
    ylm = mono @ cart2sph(l)
    gram = numpy.einsum('ij,ijm,ijn->mn', w, ylm, ylm)
    assert numpy.allclose(gram, numpy.eye(2 * l + 1), atol=1e-12)


def test_basis_sizes_match_workloads():
    water = 'O 0 0 0; H 0 0.757 0.587; H 0 -0.757 0.587'
    assert M(atom=water, basis='cc-pvdz').nao == 24
    assert M(atom=water, basis='cc-pvtz').nao == 58
    assert M(atom=water, basis='def2-svp').nao == 24
    # general contractions kept as one shell: O cc-pVDZ s is 9 primitives x 3
    o = load_basis('cc-pvdz', 'O')
    assert o[0][0] == 0 and len(o[0]) - 1 == 9 and len(o[0][1]) == 4


def test_contracted_functions_are_normalised():
    mol = M(atom='O 0 0 0', basis='cc-pvtz')
    for ib in range(mol.nbas):
        l = mol.bas_angular(ib)
        es = mol.bas_exp(ib)
        cs = mol.bas_ctr_coeff(ib)
        ee = es[:, None] + es[None, :]
        s = numpy.vectorize(lambda a: math.gamma(l + 1.5) / (2 * a ** (l + 1.5)))(ee)
        assert numpy.allclose(numpy.einsum('pi,pq,qi->i', cs, s, cs), 1.)


@pytest.mark.parametrize('basis', ['cc-pvdz', 'cc-pvtz', 'def2-svp'])
def test_seg_contraction_matches_oracle(basis):
    mol = M(atom='O 0 0 0; H 0 0.757 0.587; H 0 -0.757 0.587; O 3 0 0', basis=basis)
    pmol, c = hf.basis_seg_contraction(mol)
    bas, env, c2 = orc.seg_contract(mol)
    assert numpy.array_equal(pmol._bas, bas)
    assert numpy.allclose(pmol._env, env, rtol=1e-15)
    assert numpy.allclose(c, c2, rtol=1e-14, atol=1e-15)


def test_s_index_offsets_match_oracle():
    rng = numpy.random.default_rng(1)
    for _ in range(20):
        lq = -numpy.sort(rng.exponential(8, rng.integers(1, 500)))
        lq[lq > 0] = 0
        for nbins in (1, 10, 17):
            cut = 10 ** rng.uniform(-16, -6)
            assert numpy.array_equal(hf._make_s_index_offsets(lq, nbins, cut),
                                     orc.make_s_index_offsets(lq, nbins, cut))


def test_split_l_ctr_groups():
    uniq = numpy.array([[0, 1], [1, 1], [2, 1]])
    counts = numpy.array([10, 7, 3])
    u, c = hf._split_l_ctr_groups(uniq, counts, group_size=9)
    assert c.sum() == counts.sum()
    for (l, _), n in zip(u, c):
        assert n * (l + 1) * (l + 2) // 2 <= max(9, 2 * (l + 1) * (l + 2) // 2)


def test_rys_table_matches_mpmath():
    """Both compiled tables (n = 1..7 and 8..11 roots) against the 80-digit
    mpmath vectors."""
    gold = numpy.load(os.path.join(ROOT, 'tests', 'golden', 'rys_golden.npz'))
    for n in range(1, 12):
        for ix, x in enumerate(gold['x']):
            u, w = rys_table.rys(n, x)
            assert numpy.allclose(u, gold[f'u_{n}'][ix], rtol=2e-14, atol=0), (n, x)
            assert numpy.allclose(w, gold[f'w_{n}'][ix], rtol=2e-14, atol=0), (n, x)


def test_library_exports_header_symbols():
    from byteqc_amd.lib import library_path
    path = library_path()
    if not os.path.exists(path):
        pytest.skip('libbqcjk.so not built')
    lib = ctypes.CDLL(path)
    names = set()
    for fn in os.listdir(os.path.join(ROOT, 'include')):
        text = open(os.path.join(ROOT, 'include', fn)).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names.update(re.findall(r'^\s*(?:int|void|double)\s+\**(\w+)\s*\(', text, re.M))
    assert len(names) >= 10
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_workload_generator_is_seeded_and_respects_min_distance():
    from byteqc_amd.systems import water_cluster, random_density
    a = water_cluster(20, 12.0, 3)
    b = water_cluster(20, 12.0, 3)
    assert a == b
    o = numpy.array([xyz for s, xyz in a if s == 'O'])
    d = numpy.linalg.norm(o[:, None] - o[None], axis=-1) + numpy.eye(20) * 99
    assert d.min() >= 2.7
    h = numpy.array([xyz for s, xyz in a if s == 'H']).reshape(20, 2, 3)
    assert numpy.allclose(numpy.linalg.norm(h - o[:, None], axis=-1), 0.9572)
    dm = random_density(30, 5, 1)
    assert numpy.allclose(dm @ dm, 2 * dm)


@pytest.mark.parametrize('prescreen', [True, False])
def test_task_planner_matches_oracle(prescreen):
    """BQCplan_tasks (C++) gives the reference's task list, bins and bounds."""
    from byteqc_amd.lib import library_path, load_library
    if not os.path.exists(library_path()):
        pytest.skip('libbqcjk.so not built')
    from byteqc_amd.systems import water_cluster
    mol = M(atom=water_cluster(3, 7.0, seed=5), basis='cc-pvdz')
    plan = orc.build_plan(mol)
    ng = len(plan.uniq_l_ctr)
    cp_i, cp_j = numpy.tril_indices(ng)
    cp_grp = numpy.ascontiguousarray(numpy.stack([cp_i, cp_j], 1), dtype=numpy.int32)
    log_q = numpy.ascontiguousarray(numpy.concatenate(plan.log_qs))
    rng = numpy.random.default_rng(7)
    ncp = len(plan.log_qs)
    lib = load_library()
    for trial in range(6):
        cond = 10 ** rng.uniform(-16, 0, (ng, ng))
        cond = numpy.ascontiguousarray(numpy.maximum(cond, cond.T))
        if trial == 0:
            cond[:] = 0.0
        nbins = (10, 3, 17)[trial % 3]
        ref_cp, ref_bins, ref_sub = orc.task_list(plan, cond, nbins, prescreen)
        nmax = ncp * (ncp + 1) // 2
        tcp = numpy.zeros((nmax, 2), dtype=numpy.int32)
        tb = numpy.zeros((nmax, 2 * (nbins + 1)), dtype=numpy.int32)
        td = numpy.zeros(nmax)
        nt = ctypes.c_int(0)
        assert lib.BQCplan_tasks(ncp, hf._ptr(cp_grp), hf._ptr(plan.bas_pairs_locs),
                                 hf._ptr(log_q), hf._ptr(cond), ng, plan.cutoff, nbins,
                                 int(prescreen), hf._ptr(tcp), hf._ptr(tb), hf._ptr(td),
                                 ctypes.byref(nt)) == 0
        n = nt.value
        assert n == len(ref_cp), trial
        assert numpy.array_equal(tcp[:n], ref_cp)
        assert numpy.array_equal(tb[:n], ref_bins)
        assert numpy.array_equal(td[:n], ref_sub)
        if trial == 0:
            assert n == (0 if prescreen else nmax)


def test_poisson_workload_and_footprint_estimate():
    from byteqc_amd.scf.serve import poisson_workload, estimate_footprint
    from byteqc_amd.systems import water_mole, box_for_density
    a = poisson_workload(2.0, 200.0, seed=4)
    assert a == poisson_workload(2.0, 200.0, seed=4)
    t = numpy.array([x[0] for x in a])
    n = numpy.array([x[1] for x in a])
    assert numpy.all(numpy.diff(t) > 0) and t[-1] < 200.0
    assert abs(len(a) / 200.0 - 2.0) < 0.3             # Poisson rate
    assert n.min() >= 5 and n.max() <= 120
    assert abs(numpy.median(numpy.log(n)) - numpy.log(numpy.sqrt(5 * 120))) < 0.3
    fp = [estimate_footprint(water_mole(k, box_for_density(k), 1, 'def2-svp'))
          for k in (1, 2, 4, 8)]
    assert all(b > a_ for a_, b in zip(fp, fp[1:]))
    assert fp[3] > 3 * fp[1]


def test_fractional_aufbau_occupations():
    """Spin-summed aufbau with a shared Fermi-level set: O (8 electrons) has
    1s2 2s2 and 4/3 in each 2p; H has 1 in 1s; degenerate sets below the
    Fermi level are full."""
    from byteqc_amd.scf.rhf import _frac_occ
    e = numpy.array([0.5, -20.0, -1.2, -0.6, -0.6 + 1e-9, -0.6 - 1e-9, 0.9])
    occ = _frac_occ(e, 8)
    assert numpy.allclose(occ, [0, 2, 2, 4 / 3, 4 / 3, 4 / 3, 0])
    assert numpy.allclose(_frac_occ(numpy.array([-0.5, 0.2]), 1), [1, 0])
    occ = _frac_occ(e, 10)
    assert numpy.allclose(occ, [0, 2, 2, 2, 2, 2, 0]) and occ.sum() == 10
    assert numpy.allclose(_frac_occ(e, 11).sum(), 11)



def test_default_tiles_follow_molecule_size():
    """The tile rule is host logic in the library (no GPU call): Q pairs per
    item by Cartesian AO count, below 2048 / from 2048 / from 8192: row
    classes up to 81 components 128 / 64 / 256 (below 2048: 32 for (ss|ss),
    16 for (pp|pp)), larger row classes 128 / 64 / 128, wave classes 32 /
    32 / 64; a tile set for all classes of a kind overrides the rule and a
    per-class tile overrides both; kind 0 restores the rule."""
    from byteqc_amd.lib import library_path
    path = library_path()
    if not os.path.exists(path):
        pytest.skip('libbqcjk.so not built')
    lib = ctypes.CDLL(path)
    get = lib.BQCget_tile
    assert [get(1, 1000, 0, n) for n in (500, 2047, 2048, 8191, 8192, 10368)] == \
        [128, 128, 64, 64, 256, 256]
    assert [get(1, 0, 0, n) for n in (500, 2624, 10368)] == [32, 64, 256]
    assert [get(1, 1111, 0, n) for n in (500, 2624, 10368)] == [16, 64, 256]
    assert [get(1, 2120, 1, n) for n in (500, 2624, 10368)] == [128, 64, 128]
    assert [get(4, 2111, 0, n) for n in (500, 2624, 10368)] == [32, 32, 64]
    assert get(2, 0, 0, 100) == -1
    try:
        assert lib.BQCset_tile(1, -1, 0, 128) == 0
        assert lib.BQCset_tile(1, 1000, 0, 16) == 0
        assert get(1, 0, 0, 3000) == 128 and get(1, 2120, 1, 100) == 128
        assert get(1, 1000, 0, 3000) == 16 and get(1, 1000, 1, 3000) == 128
        assert get(4, 2111, 0, 3000) == 32
        assert lib.BQCset_tile(1, 1000, 0, 0) == 1 and lib.BQCset_tile(3, 1000, 0, 8) == 1
    finally:
        lib.BQCset_tile(0, -1, 0, 0)
    assert get(1, 1000, 0, 3000) == 64


def test_cartesian_ao_limit_is_refused():
    """nao_c**2 must fit the kernels' 32-bit element offsets: the host plan
    refuses more than 46340 Cartesian AOs, and the C entry points return 4
    before touching the device."""
    hf.check_nao(46340)
    with pytest.raises(ValueError, match='46340'):
        hf.check_nao(46341)
    from byteqc_amd.lib import load_library
    lib = load_library()
    assert lib.BQCbuild_jk(None, None, None, None, 46341, 1, 0, None, None, 10, None, 0,
                           None, None, 0., None) == 4
    assert lib.GINTbuild_jk(None, None, None, None, 46341, 1, None, None, 10, 0, 0,
                            None) == 4


def test_build_entries_refuse_bad_arguments_before_the_device():
    """BQCbuild_jk checks its arguments before planning or any device call:
    a precision other than 0 (FP64) or 1 (mixed) returns 1, more than 4096
    Schwarz bins returns 3; an empty task list returns 0 without work."""
    from byteqc_amd.lib import load_library
    lib = load_library()
    args = lambda nao, ntask, nbins, prec: (None, None, None, None, nao, 1, ntask, None, None,
                                            nbins, None, prec, None, None, 0., None)
    assert lib.BQCbuild_jk(*args(100, 0, 10, 2)) == 1
    assert lib.BQCbuild_jk(*args(100, 0, 10, -1)) == 1
    assert lib.BQCbuild_jk(*args(100, 0, 4097, 0)) == 3
    assert lib.BQCbuild_jk(*args(100, 0, 10, 0)) == 0


def test_pair_builder_sorts_and_screens_primitive_pairs():
    """GINTinit_basis_prod_cpu (host code): inside every shell pair the
    primitive pairs hold the same (a12, e12, P) as a direct evaluation of
    e12 = fac ci cj exp(-ai aj r^2 / (ai + aj)), sorted by |e12| from the
    largest; those below 1e-20 of the largest |e12| of the cache are zero,
    and no larger one is."""
    from byteqc_amd.lib import library_path
    path = library_path()
    if not os.path.exists(path):
        pytest.skip('libbqcjk.so not built')
    from byteqc_amd.systems import water_mole, box_for_density
    lib = ctypes.CDLL(path)
    mol = water_mole(4, box_for_density(4), 3, 'def2-svp')
    pmol, _ = hf.basis_seg_contraction(mol)
    bas = numpy.ascontiguousarray(pmol._bas, dtype=numpy.int32)
    atm = numpy.ascontiguousarray(pmol._atm, dtype=numpy.int32)
    env = numpy.ascontiguousarray(pmol._env, dtype=numpy.float64)
    nbas = len(bas)
    # one pair type per (shell group, shell group), groups by (l, nprim)
    key = [(int(b[1]), int(b[2])) for b in bas]
    groups = sorted(set(key))
    bra, ket = [], []
    for gi, g1 in enumerate(groups):
        for g2 in groups[:gi + 1]:
            sh1 = [i for i in range(nbas) if key[i] == g1]
            sh2 = [j for j in range(nbas) if key[j] == g2]
            pr = [(i, j) for i in sh1 for j in sh2 if g1 != g2 or i >= j]
            bra.append([p[0] for p in pr])
            ket.append([p[1] for p in pr])
    pair2shls = numpy.ascontiguousarray(numpy.vstack([numpy.hstack(bra), numpy.hstack(ket)]),
                                        dtype=numpy.int32)
    locs = numpy.append(0, numpy.cumsum([len(x) for x in bra])).astype(numpy.int32)
    ncp = len(bra)
    coords = numpy.zeros(3 * nbas)
    bp = ctypes.POINTER(hf.BasisProdCache)()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.GINTinit_basis_prod_cpu(ctypes.byref(bp), ctypes.c_double(0.5), p(coords), p(pair2shls),
                                p(locs), ctypes.c_int(ncp), p(atm), ctypes.c_int(len(atm)),
                                p(bas), ctypes.c_int(nbas), p(env))
    c = bp.contents
    nprim_pairs = c.primitive_pairs_locs[ncp]
    aexyz = numpy.ctypeslib.as_array(c.aexyz, (5 * nprim_pairs,)).reshape(5, nprim_pairs).copy()
    fac_sp = {0: 0.282094791773878143, 1: 0.488602511902919921}
    ref = []
    for n in range(ncp):
        for k in range(locs[n], locs[n + 1]):
            i, j = pair2shls[0, k], pair2shls[1, k]
            bi, bj = bas[i], bas[j]
            ai, aj = env[bi[5]:bi[5] + bi[2]], env[bj[5]:bj[5] + bj[2]]
            ci, cj = env[bi[6]:bi[6] + bi[2]], env[bj[6]:bj[6] + bj[2]]
            ri, rj = env[atm[bi[0], 1]:atm[bi[0], 1] + 3], env[atm[bj[0], 1]:atm[bj[0], 1] + 3]
            f = fac_sp.get(int(bi[1]), 1.0) * fac_sp.get(int(bj[1]), 1.0) * (0.5 if i == j else 1.0)
            a12 = numpy.add.outer(ai, aj).ravel()
            e12 = (f * numpy.outer(ci, cj) *
                   numpy.exp(-((ri - rj) ** 2).sum() * numpy.outer(ai, aj) /
                             numpy.add.outer(ai, aj))).ravel()
            pxyz = [(numpy.add.outer(ai * ri[d], aj * rj[d])).ravel() / a12 for d in range(3)]
            ref.append(numpy.vstack([a12, e12] + pxyz))
    emax = max(numpy.abs(r[1]).max() for r in ref)
    off, nzero = 0, 0
    for r in ref:
        k = r.shape[1]
        got = aexyz[:, off:off + k]
        order = numpy.argsort(-numpy.abs(r[1]), kind='stable')
        want = r[:, order].copy()
        small = numpy.abs(want[1]) < 1e-20 * emax
        want[1, small] = 0.0
        nzero += int(small.sum())
        assert numpy.all(numpy.diff(numpy.abs(got[1])) <= 0)
        # equal |e12| (the (ip, jp) and (jp, ip) pairs of a diagonal shell
        # pair) may come in either order: compare the columns as sets
        canon = lambda x: x[:, numpy.lexsort([numpy.round(x[i], 9) for i in (4, 3, 2, 1, 0)])]
        # exp(-mu r^2) of arguments near -40: the two operation orders agree
        # to ~4e-14
        assert numpy.allclose(canon(got), canon(want), rtol=1e-12, atol=0)
        off += k
    assert off == nprim_pairs and nzero > 0
    lib.GINTdel_basis_prod(ctypes.byref(bp))


def test_host_controls_check_their_arguments():
    """Process-wide controls of the library (host code, no GPU call): the
    split rule takes 0/1 per class and -1/0/1 for all classes, the planning
    threads 0..256, the launch mode 0..2; each returns 1 on a bad value and
    leaves the setting alone."""
    from byteqc_amd.lib import library_path
    path = library_path()
    if not os.path.exists(path):
        pytest.skip('libbqcjk.so not built')
    lib = ctypes.CDLL(path)
    try:
        assert lib.BQCset_row_split(1000, 2) == 1 and lib.BQCset_row_split(1000, -1) == 1
        assert lib.BQCset_row_split(-1, 2) == 1
        assert lib.BQCset_row_split(1000, 0) == 0 and lib.BQCset_row_split(-1, 1) == 0
        assert lib.BQCset_row_split(-1, -1) == 0
        assert lib.BQCset_plan_threads(-1) == 1 and lib.BQCset_plan_threads(257) == 1
        assert lib.BQCset_plan_threads(4) == 0 and lib.BQCset_plan_threads(0) == 0
        assert lib.BQCset_launch_mode(3) == 1 and lib.BQCset_launch_mode(-1) == 1
        assert lib.BQCset_launch_mode(1) == 0
    finally:
        lib.BQCset_row_split(-1, -1)
        lib.BQCset_plan_threads(0)
        lib.BQCset_launch_mode(2)
