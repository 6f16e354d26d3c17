"""The reference-ABI entry GINTbuild_jk driven the way the reference's host
code drives it: one ctypes call per (cp_ij, cp_kl) task with that task's
Schwarz bins, on the current stream (cuobc/scf/hf.py:111-166, call at
hf.py:147-156), then J = 2(V + V^T), K = V + V^T and the back-transform
(hf.py:171-181).  Bar: 1e-10 against the oracle fixture."""
import os
import time

import numpy
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def per_task_get_jk(mol, dm, opt, nbins=None):
    """(vj, vk) on the device from one GINTbuild_jk call per task."""
    torch = _torch()
    from byteqc_amd.lib import load_library
    from byteqc_amd.scf import hf
    lib = load_library()
    coeff = opt.coeff_dev
    nao, nao0 = coeff.shape
    dms = torch.as_tensor(dm, dtype=torch.float64, device=coeff.device).reshape(-1, nao0, nao0)
    n_dm = dms.shape[0]
    dmc = torch.matmul(torch.matmul(coeff, dms), coeff.T).contiguous()
    if nbins is None:
        nbins = opt.default_nbins()
    cond = numpy.max(hf.dm_cond_batch(list(dmc), [opt.group_ao_loc()] * n_dm), axis=0)
    task_cp, task_bins, _ = opt.tasks(cond, nbins)
    vj = torch.zeros((n_dm, nao, nao), dtype=torch.float64, device=coeff.device)
    vk = torch.zeros_like(vj)
    stream = hf._stream_ptr()
    for t in range(len(task_cp)):
        bi = numpy.ascontiguousarray(task_bins[t, :nbins + 1])
        bk = numpy.ascontiguousarray(task_bins[t, nbins + 1:])
        err = lib.GINTbuild_jk(opt.bpcache, hf._ptr(vj), hf._ptr(vk), hf._ptr(dmc), nao, n_dm,
                               hf._ptr(bi), hf._ptr(bk), nbins, int(task_cp[t, 0]),
                               int(task_cp[t, 1]), stream)
        assert err == 0, (t, err)
    lib.BQCsymmetrize_jk(hf._ptr(vj), hf._ptr(vk), nao, n_dm, stream)
    out = [torch.matmul(torch.matmul(coeff.T, v), coeff) for v in (vj, vk)]
    if numpy.ndim(dm) == 2:
        out = [v[0] for v in out]
    return out[0], out[1]


def test_per_task_loop_matches_golden_c1():
    _torch()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    mol, dm, _ = workload('c1')
    gold = numpy.load(os.path.join(GOLD, 'jk_c1.npz'))
    opt = hf._VHFOpt(mol).build()
    j, k = per_task_get_jk(mol, dm, opt)
    dj = numpy.abs(j.cpu().numpy() - gold['vj']).max()
    dk = numpy.abs(k.cpu().numpy() - gold['vk']).max()
    print(f'GINTbuild_jk per task, C1: max|dJ| {dj:.3e} max|dK| {dk:.3e}', flush=True)
    assert dj < 1e-10 and dk < 1e-10


def test_per_task_loop_many_bins_and_ff_classes():
    """cc-pVTZ water dimer (every class to (ff|ff), whose quartets take the
    generic kernel) with 40 bins: more rectangles than one generic launch
    carries, and more bins than the earlier limit of 32."""
    _torch()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    from oracle import oracle as orc
    mol = water_mole(2, 4.0, 11, 'cc-pvtz')
    dm = random_density(mol.nao, 10, 3)
    opt = hf._VHFOpt(mol).build()
    j, k = per_task_get_jk(mol, dm, opt, nbins=40)
    jref, kref = orc.get_jk(mol, dm, nbins=40)
    dj = numpy.abs(j.cpu().numpy() - jref).max()
    dk = numpy.abs(k.cpu().numpy() - kref).max()
    print(f'GINTbuild_jk per task, 2 H2O cc-pVTZ, 40 bins: max|dJ| {dj:.3e} max|dK| {dk:.3e}',
          flush=True)
    assert dj < 1e-10 and dk < 1e-10
    j2, k2 = hf.get_jk(mol, dm, vhfopt=opt, nbins=40)
    assert numpy.abs(j2 - jref).max() < 1e-10 and numpy.abs(k2 - kref).max() < 1e-10


def test_per_task_loop_speed_on_c2():
    """64 waters cc-pVDZ: the per-task loop against one BQCbuild_jk call
    (same kernels; results equal to 1e-10)."""
    torch = _torch()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    mol, dm, _ = workload('c2')
    opt = hf._VHFOpt(mol).build()
    dm_d = torch.as_tensor(dm, device='cuda')

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best, out

    t_task, (j1, k1) = timed(lambda: per_task_get_jk(mol, dm_d, opt))
    t_one, (j0, k0) = timed(lambda: hf.get_jk(mol, dm_d, vhfopt=opt, return_device=True))
    dj = (j1 - j0).abs().max().item()
    dk = (k1 - k0).abs().max().item()
    print(f'C2: per-task GINTbuild_jk {t_task * 1e3:.1f} ms, BQCbuild_jk {t_one * 1e3:.1f} ms '
          f'(ratio {t_task / t_one:.3f}); max|dJ| {dj:.2e} max|dK| {dk:.2e}', flush=True)
    assert dj < 1e-10 and dk < 1e-10
    assert t_task < 1.5 * t_one
