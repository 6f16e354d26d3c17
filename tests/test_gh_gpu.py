"""g and h shells on the GPU (run-time-class kernel, 8-11 Rys roots): J/K
and Schwarz bounds against the CPU oracle on molecules with primitive and
contracted g/h shells (no basis file holds them here, so the shells are
given explicitly in PySCF's [[l, [exp, coeff, ...], ...], ...] form).
Bar: 1e-10 for FP64 J/K."""
import numpy
import pytest

pytestmark = pytest.mark.gpu

GH_BASIS = {
    # s, general-contracted p, primitive g
    'H': [[0, [1.3, 1.0]], [1, [2.1, 0.6, 0.3], [0.6, 0.5, 1.0]], [4, [0.9, 1.0]]],
    # contracted d, primitive h, contracted g (two primitives)
    'O': [[2, [1.7, 0.7], [0.5, 0.4]], [5, [0.8, 1.0]], [4, [1.5, 0.5], [0.45, 0.6]]],
}


def _mol():
    from byteqc_amd.gto import M
    return M(atom='O 0 0 0; H 0.3 0.9 0.5; H -0.8 -0.2 0.7', basis=GH_BASIS)


def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def test_gh_shells_match_oracle():
    _gpu()
    from byteqc_amd.scf import hf
    from oracle import oracle as orc
    mol = _mol()
    rng = numpy.random.default_rng(17)
    a = rng.standard_normal((mol.nao, mol.nao))
    dm = a + a.T
    opt = hf._VHFOpt(mol).build()
    assert max(l for l, _ in opt.uniq_l_ctr) == 5
    j, k = hf.get_jk(mol, dm, vhfopt=opt)
    jref, kref = orc.get_jk(mol, dm)
    dj = numpy.abs(j - jref).max()
    dk = numpy.abs(k - kref).max()
    print(f'g/h shells: max|dJ| {dj:.3e} max|dK| {dk:.3e} (|J| {numpy.abs(jref).max():.2e})',
          flush=True)
    assert dj < 1e-10 and dk < 1e-10
    # mixed precision keeps these classes in FP64
    jm, km = hf.get_jk(mol, dm, vhfopt=opt, precision='mixed')
    assert numpy.abs(jm - j).max() < 1e-8 and numpy.abs(km - k).max() < 1e-8


def test_gh_schwarz_bounds_match_oracle():
    _gpu()
    from byteqc_amd.scf import hf
    from oracle import oracle as orc
    opt = hf._VHFOpt(_mol()).build()
    pm = opt.mol
    qref = orc.q_cond(pm._atm, pm._bas, pm._env)
    big = qref > 1e-20
    assert numpy.allclose(opt.q_cond[big], qref[big], rtol=1e-12)


def test_gh_two_molecules_dense_density():
    """Two g/h-shell molecules side by side: every class with l = 4, 5 on
    both bra and ket, tasks across groups, density prescreen active."""
    _gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf
    from oracle import oracle as orc
    mol = M(atom='O 0 0 0; H 0.3 0.9 0.5; O 2.5 0.4 -0.3; H 2.9 -0.5 0.2',
            basis=GH_BASIS)
    c = numpy.linalg.qr(numpy.random.default_rng(3).standard_normal((mol.nao, 12)))[0]
    dm = 2 * c @ c.T
    j, k = hf.get_jk(mol, dm)
    jref, kref = orc.get_jk(mol, dm)
    dj = numpy.abs(j - jref).max()
    dk = numpy.abs(k - kref).max()
    print(f'g/h shells, 4 atoms: max|dJ| {dj:.3e} max|dK| {dk:.3e}', flush=True)
    assert dj < 1e-10 and dk < 1e-10
