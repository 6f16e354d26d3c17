"""GPU parity of the HIP J/K path (libbqcjk.so through its C-ABI) with the
CPU oracle.  Bar: max |dJ|, max |dK| <= 1e-10 for the FP64 path.
"""
import os

import numpy
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10
GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _report(tag, j, k, jref, kref):
    dj = numpy.abs(j - jref).max()
    dk = numpy.abs(k - kref).max()
    print(f'{tag}: max|dJ| {dj:.3e}  max|dK| {dk:.3e}', flush=True)
    return dj, dk


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-svp', 'cc-pvtz'])
def test_one_water_matches_oracle(basis):
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf
    from oracle import oracle as orc
    mol = M(atom='O 0 0 0.58589194; H 0 0.75668992 0; H 0 -0.75668992 0', basis=basis)
    rng = numpy.random.default_rng(5)
    a = rng.standard_normal((mol.nao, mol.nao))
    dm = a + a.T
    j, k = hf.get_jk(mol, dm)
    jref, kref = orc.get_jk(mol, dm)
    dj, dk = _report(f'1 H2O {basis}', j, k, jref, kref)
    assert dj < TOL and dk < TOL


def test_primitive_pair_screening_keeps_oracle_parity():
    """6 waters def2-SVP: the pair builder zeroes the primitive pairs below
    1e-20 of the largest |e12| (some of them here) and the kernels skip them;
    the oracle keeps every primitive pair.  FP64 and mixed J/K against it."""
    torch = _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, box_for_density
    from oracle import oracle as orc
    mol = water_mole(6, box_for_density(6), 11, 'def2-svp')
    opt = hf._VHFOpt(mol).build()
    e12 = opt._dev['aexyz'].reshape(5, -1)[1]
    nzero = int((e12 == 0).sum())
    print(f'6 H2O def2-svp: {nzero} of {e12.numel()} primitive pairs screened', flush=True)
    assert nzero > 0
    rng = numpy.random.default_rng(12)
    c = numpy.linalg.qr(rng.standard_normal((mol.nao_nr(), 30)))[0]
    dm = 2 * c @ c.T
    jref, kref = orc.get_jk(mol, dm)
    for prec in ('fp64', 'mixed'):
        j, k = hf.get_jk(mol, dm, vhfopt=opt, precision=prec)
        dj, dk = _report(f'6 H2O def2-svp {prec}', j, k, jref, kref)
        assert dj < (TOL if prec == 'fp64' else 1e-8) and dk < (TOL if prec == 'fp64' else 1e-8)


def test_c1_matches_golden():
    """8 waters cc-pVDZ (192 AOs) against the committed oracle fixture."""
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    mol, dm, _ = workload('c1')
    gold = numpy.load(os.path.join(GOLD, 'jk_c1.npz'))
    j, k = hf.get_jk(mol, dm)
    dj, dk = _report('C1', j, k, gold['vj'], gold['vk'])
    assert dj < TOL and dk < TOL


def test_multiple_densities_and_partial_outputs():
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    mol = water_mole(3, 6.0, 7, 'def2-svp')
    dms = numpy.stack([random_density(mol.nao, 15, s) for s in (1, 2)])
    opt = hf._VHFOpt(mol).build()
    j, k = hf.get_jk(mol, dms, vhfopt=opt)
    assert j.shape == dms.shape and k.shape == dms.shape
    for i in range(2):
        j1, k1 = hf.get_jk(mol, dms[i], vhfopt=opt)
        assert numpy.abs(j1 - j[i]).max() < 1e-12
        assert numpy.abs(k1 - k[i]).max() < 1e-12
    j2, none = hf.get_jk(mol, dms, vhfopt=opt, with_k=False)
    assert none is None and numpy.abs(j2 - j).max() < 1e-12
    none, k2 = hf.get_jk(mol, dms, vhfopt=opt, with_j=False)
    assert none is None and numpy.abs(k2 - k).max() < 1e-12


def test_schwarz_bounds_match_oracle():
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole
    from oracle import oracle as orc
    mol = water_mole(2, 5.0, 3, 'cc-pvdz')
    opt = hf._VHFOpt(mol).build()
    pm = opt.mol
    qref = orc.q_cond(pm._atm, pm._bas, pm._env)
    big = qref > 1e-20
    assert numpy.allclose(opt.q_cond[big], qref[big], rtol=1e-12)


def test_hermi_and_omega_errors():
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf
    mol = M(atom='H 0 0 0; H 0 0 0.74', basis='cc-pvdz')
    dm = numpy.eye(mol.nao)
    with pytest.raises(NotImplementedError):
        hf.get_jk(mol, dm, hermi=0)

    class MF:
        pass
    with pytest.raises(NotImplementedError):
        hf._get_jk(MF(), mol, dm, omega=0.3)


def test_water_dimer_cc_pvtz_matches_oracle():
    """f shells on both centres: every class up to (ff|ff)."""
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    from oracle import oracle as orc
    mol = water_mole(2, 4.0, 11, 'cc-pvtz')
    dm = random_density(mol.nao, 10, 3)
    j, k = hf.get_jk(mol, dm)
    jref, kref = orc.get_jk(mol, dm)
    dj, dk = _report('2 H2O cc-pVTZ', j, k, jref, kref)
    assert dj < TOL and dk < TOL


def test_mixed_precision_c1_energy_bar():
    """Mixed precision (FP32 for quartets with bound < 1e-7, FP64 roots and
    accumulators) against the FP64 oracle fixture: |dE| <= 1e-8 Eh with
    E = 1/2 Tr[D (J - K/2)], and the J/K deviations stated."""
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    mol, dm, _ = workload('c1')
    gold = numpy.load(os.path.join(GOLD, 'jk_c1.npz'))
    j, k = hf.get_jk(mol, dm, precision='mixed', fp32_cut=1e-7)
    fl = hf.last_build_flops()
    e = 0.5 * numpy.einsum('ij,ij', dm, j - 0.5 * k)
    eref = 0.5 * numpy.einsum('ij,ij', dm, gold['vj'] - 0.5 * gold['vk'])
    dj, dk = _report('C1 mixed', j, k, gold['vj'], gold['vk'])
    print(f'C1 mixed: |dE| {abs(e - eref):.3e} Eh, FP32 share of FLOPs {fl[1] / fl.sum():.3f}',
          flush=True)
    assert fl[1] > 0
    assert abs(e - eref) <= 1e-8
    assert dj <= 1e-8 and dk <= 1e-8


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-svp', 'cc-pvtz'])
def test_int1e_matches_oracle(basis):
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf, rhf
    from oracle import scf_oracle as so
    mol = M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587; O 2.1 0.3 -0.4', basis=basis)
    opt = hf._VHFOpt(mol).build()
    got = [m.cpu().numpy() for m in rhf.int1e(opt)]
    ref = so.int1e(mol)
    for name, g, r in zip('STV', got, ref):
        d = numpy.abs(g - r).max()
        print(f'int1e {basis} {name}: max diff {d:.2e}', flush=True)
        assert d < 1e-11


def test_rhf_known_answer_and_mixed_schedule():
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import rhf
    mol = M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587', basis='cc-pvdz')
    e64, mf = rhf.scf_energy(mol, conv_tol=1e-10)
    print(f'RHF fp64 {e64:.10f} in {mf.cycles} cycles', flush=True)
    assert mf.converged and abs(e64 - (-76.0267656731)) < 1e-9
    emx, mf2 = rhf.scf_energy(mol, conv_tol=1e-10, precision='mixed_then_fp64')
    print(f'RHF mixed->fp64 {emx:.10f} builds {mf2.jk_builds}', flush=True)
    assert mf2.converged and abs(emx - e64) < 1e-8
    eh, mf3 = rhf.scf_energy(mol, conv_tol=1e-10, init_guess='hcore')
    print(f'RHF core guess {eh:.10f} in {mf3.cycles} cycles', flush=True)
    assert mf3.converged and abs(eh - e64) < 1e-9 and mf.cycles < mf3.cycles


def test_incremental_builds_match_full_rebuilds():
    """J/K of the density change added to the last build's (PySCF's direct-SCF
    get_veff): the same SCF energy as rebuilding from the whole density every
    cycle (<= 1e-9 Eh), in the same cycles, and the density prescreen makes
    the late builds cheaper than the first."""
    _torch_gpu()
    from byteqc_amd.scf import rhf
    from byteqc_amd.systems import water_mole, box_for_density
    mol = water_mole(6, box_for_density(6), 3, 'def2-svp')
    for prec in ('fp64', 'mixed_then_fp64'):
        full = rhf.RHF(mol, conv_tol=1e-9, init_guess='hcore', precision=prec,
                       incremental=False)
        e_full = full.kernel()
        incr = rhf.RHF(mol, conv_tol=1e-9, init_guess='hcore', precision=prec)
        e_incr = incr.kernel()
        print(f'{prec}: full {e_full:.10f} in {full.cycles} cycles, incremental '
              f'{e_incr:.10f} in {incr.cycles}; FLOPs first {incr.build_flops[0]:.3g}, '
              f'last {incr.build_flops[-1]:.3g}', flush=True)
        assert full.converged and incr.converged
        assert abs(e_incr - e_full) <= 1e-9
        assert abs(incr.cycles - full.cycles) <= 2
        # 6 compact waters: 0.59x measured; larger molecules, whose distant
        # pairs sit near the cutoff, drop far more
        assert incr.build_flops[-1] < 0.75 * incr.build_flops[0]


def test_parallel_planning_matches_one_thread():
    """A batch planned by 8 host threads (blocks of tasks merged in order)
    gives the work list of one thread: the same FLOP count to the last digit,
    and J/K equal to the atomics' summation-order noise, eager (planned class
    by class) and captured in a graph (planned as one list)."""
    torch = _torch_gpu()
    from byteqc_amd.lib import load_library
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    from byteqc_amd.systems import water_mole, box_for_density
    lib = load_library()
    mols = [water_mole(n, box_for_density(n), 40 + i, 'def2-svp')
            for i, n in enumerate([3, 4, 5, 6] * 6)]
    rng = numpy.random.default_rng(8)
    dms = []
    for m in mols:
        c = numpy.linalg.qr(rng.standard_normal((m.nao_nr(), m.nelectron // 2)))[0]
        dms.append(torch.as_tensor(2 * c @ c.T, device='cuda'))
    bj = BatchJK(mols)
    out = {}
    try:
        for nth in (1, 8):
            assert lib.BQCset_plan_threads(nth) == 0
            eager = [(j.cpu().numpy(), k.cpu().numpy()) for j, k in bj.build(dms)]
            fl = list(hf.last_build_flops())
            bj.capture(dms)
            graph = [(j.cpu().numpy(), k.cpu().numpy()) for j, k in bj.replay()]
            out[nth] = (eager, graph, fl)
    finally:
        lib.BQCset_plan_threads(0)
    assert out[1][2] == out[8][2]
    for which in (0, 1):
        d = max(numpy.abs(a - b).max() for x, y in zip(out[1][which], out[8][which])
                for a, b in zip(x, y))
        print(f'{"graph" if which else "eager"}: FLOPs {out[1][2]}, max diff 1 vs 8 threads '
              f'{d:.1e}', flush=True)
        assert d < 1e-12
    assert lib.BQCset_plan_threads(-1) == 1


def test_atomic_densities_hold_the_atoms_electrons():
    """Spherically averaged atomic densities: Tr(D S) is the electron count,
    D is symmetric and its natural occupations lie in [0, 2]."""
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf, rhf
    for sym, basis, ne in (('O', 'cc-pvdz', 8), ('H', 'def2-svp', 1), ('O', 'cc-pvtz', 8)):
        dm = rhf.atomic_density(sym, basis)
        atom = M(atom=f'{sym} 0 0 0', basis=basis)
        s = rhf.int1e(hf._VHFOpt(atom).build())[0].cpu().numpy()
        assert abs(numpy.trace(dm @ s) - ne) < 1e-10
        assert numpy.abs(dm - dm.T).max() < 1e-12
        w, u = numpy.linalg.eigh(s)
        half = u @ numpy.diag(numpy.sqrt(w)) @ u.T
        occ = numpy.linalg.eigvalsh(half @ dm @ half)
        assert occ.min() > -1e-10 and occ.max() < 2 + 1e-10


def test_fp32_everywhere_cc_pvtz_stays_at_fp32_accuracy():
    """All quartets of every class with an FP32 instance in FP32 (cut 1e30),
    classes to (ff|ff): J and K stay within FP32 rounding of the FP64 result
    (catches a broken FP32 kernel that the 1e-7 split would hide behind small
    integrals)."""
    _torch_gpu()
    from byteqc_amd.gto import M
    from byteqc_amd.scf import hf
    mol = M(atom='O 0 0 0.58589194; H 0 0.75668992 0; H 0 -0.75668992 0', basis='cc-pvtz')
    rng = numpy.random.default_rng(5)
    a = rng.standard_normal((mol.nao, mol.nao))
    dm = a + a.T
    opt = hf._VHFOpt(mol).build()
    j0, k0 = hf.get_jk(mol, dm, vhfopt=opt)
    j1, k1 = hf.get_jk(mol, dm, vhfopt=opt, precision='mixed', fp32_cut=1e30)
    dj, dk = _report('1 H2O cc-pVTZ all FP32', j1, k1, j0, k0)
    scale = max(abs(j0).max(), abs(k0).max())
    assert dj < 1e-5 * scale and dk < 1e-5 * scale


def test_scf_hook_caches_plan_and_matches_oracle():
    """_get_jk (cuobc/scf/hf.py:230-250): the plan is built once per SCF
    object and reused; J/K equal the oracle's."""
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    from oracle import oracle as orc
    mol = water_mole(3, 6.0, 31, 'cc-pvdz')

    class MF:
        direct_scf_tol = 1e-13

        def __init__(self, mol):
            self.mol = mol

    mf = MF(mol)
    dms = [random_density(mol.nao, 15, s) for s in (41, 42)]
    plans = []
    for dm in dms:
        j, k = hf._get_jk(mf, dm=dm)
        plans.append(mf._opt_gpu)
        jref, kref = orc.get_jk(mol, dm)
        dj, dk = _report('_get_jk 3 H2O cc-pVDZ', j, k, jref, kref)
        assert dj < TOL and dk < TOL
    assert plans[0] is plans[1]
    j, none = hf._get_jk(mf, mol, dms[0], with_k=False)
    assert none is None and numpy.abs(j - orc.get_jk(mol, dms[0])[0]).max() < TOL


@pytest.mark.parametrize('group_size', [8, 12])
def test_group_size_build_matches_oracle(group_size):
    """Pair groups capped at group_size AOs (_split_l_ctr_groups,
    cuobc/scf/hf.py:581-606): more, smaller tasks, the same J/K."""
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    from oracle import oracle as orc
    mol = water_mole(4, 6.5, 37, 'def2-svp')
    dm = random_density(mol.nao, 20, 43)
    opt = hf._VHFOpt(mol).build(group_size=group_size)
    assert len(opt.uniq_l_ctr) > len(hf._VHFOpt(mol).build().uniq_l_ctr)
    j, k = hf.get_jk(mol, dm, vhfopt=opt)
    jref, kref = orc.get_jk(mol, dm)
    dj, dk = _report(f'group_size {group_size}', j, k, jref, kref)
    assert dj < TOL and dk < TOL
    j2, k2 = hf.get_jk(mol, dm, group_size=group_size)
    assert numpy.abs(j2 - j).max() < 1e-12 and numpy.abs(k2 - k).max() < 1e-12
