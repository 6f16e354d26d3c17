"""GPU tests of the four-index ERI fill (GINTfill_int2e) against the oracle's
ERI tensor and against the J/K path."""
import numpy
import pytest

pytestmark = pytest.mark.gpu

WATER = 'O 0 0 0.58589194; H 0 0.75668992 0; H 0 -0.75668992 0'


def _torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-svp'])
def test_fill_matches_oracle_tensor(basis):
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import int4c
    from oracle import oracle as orc
    mol = M(atom=WATER, basis=basis)
    eri, opt = int4c.get_int4c(mol, cutoff=1e-20)
    pm = opt.mol
    ref = orc.eri_tensor(pm._atm, pm._bas, pm._env)
    d = float(numpy.abs(eri.cpu().numpy() - ref).max())
    print(f'1 H2O {basis}: {ref.shape[0]} Cartesian AOs, max|d eri| {d:.2e}', flush=True)
    assert d < 1e-12


def test_fill_contracted_with_density_equals_jk_path():
    """cc-pVTZ (f shells): J, K from the filled tensor equal get_jk."""
    torch = _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf, int4c
    mol = M(atom=WATER, basis='cc-pvtz')
    eri, opt = int4c.get_int4c(mol, cutoff=1e-20)
    rng = numpy.random.default_rng(3)
    a = rng.standard_normal((mol.nao, mol.nao))
    dm = a + a.T
    c = opt.coeff_dev
    dmc = c @ torch.as_tensor(dm, device='cuda') @ c.T
    n = dmc.shape[0]
    e2 = eri.reshape(n * n, n * n)
    jc = (e2 @ dmc.reshape(-1)).reshape(n, n)
    kc = torch.einsum('ikjl,kl->ij', eri, dmc.contiguous())
    j = (c.T @ jc @ c).cpu().numpy()
    k = (c.T @ kc @ c).cpu().numpy()
    jr, kr = hf.get_jk(mol, dm)
    dj, dk = numpy.abs(j - jr).max(), numpy.abs(k - kr).max()
    print(f'cc-pVTZ fill vs J/K path: max|dJ| {dj:.2e} max|dK| {dk:.2e}', flush=True)
    assert dj < 1e-10 and dk < 1e-10


def test_fill_screened_bins_and_errors():
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import int4c
    from byteqc_amd.systems import water_mole
    mol = water_mole(3, 7.0, 21, 'cc-pvdz')
    opt = int4c.build_plan(mol)
    ncp = len(opt.log_qs)
    for a, b in [(0, 0), (ncp - 1, 0), (1, ncp - 1)]:
        full = int4c.get_int4c_slice(a, b, opt, nbins=1)[-1]
        screened = int4c.get_int4c_slice(a, b, opt, nbins=10)[-1]
        # the bins drop only quartets with q_ij q_kl below the cutoff scale
        assert float((full - screened).abs().max()) < 1e-10
    # the C entry point rejects a pair type outside the plan (return code 1)
    import torch
    from byteqc_amd.lib import load_library
    from byteqc_amd.scf import hf
    buf = torch.zeros(16, dtype=torch.float64, device='cuda')
    four = numpy.zeros(4, dtype=numpy.int32)
    bins = numpy.array([0, 1], dtype=numpy.int32)
    assert load_library().GINTfill_int2e(opt.bpcache, hf._ptr(buf), 1, hf._ptr(four),
                                         hf._ptr(four), hf._ptr(bins), hf._ptr(bins), 1, ncp, 0,
                                         hf._stream_ptr()) == 1
