"""BQCbuild_jk_noalloc (work list in a static per-GPU array, no allocation)
against BQCbuild_jk through the C-ABI, on two calls in a row, the second on a
smaller molecule; bar 1e-10 (FP64, atomics reorder the sums)."""
import numpy
import pytest

pytestmark = pytest.mark.gpu


def _build(entry, opt, dm_d, torch):
    from byteqc_amd.lib import load_library
    from byteqc_amd.scf import hf
    lib = load_library()
    coeff = opt.coeff_dev
    nao = coeff.shape[0]
    dmc = torch.matmul(torch.matmul(coeff, dm_d), coeff.T).contiguous().reshape(1, nao, nao)
    nbins = opt.default_nbins()
    cond = hf.dm_cond_batch(list(dmc), [opt.group_ao_loc()])[0]
    task_cp, task_bins, task_dmax = opt.tasks(cond, nbins)
    vj = torch.zeros((1, nao, nao), dtype=torch.float64, device=dm_d.device)
    vk = torch.zeros_like(vj)
    fn = getattr(lib, entry)
    err = fn(opt.bpcache, hf._ptr(vj), hf._ptr(vk), hf._ptr(dmc), nao, 1, len(task_cp),
             hf._ptr(numpy.ascontiguousarray(task_cp)),
             hf._ptr(numpy.ascontiguousarray(task_bins)), nbins, hf._ptr(opt.bra_order), 0,
             hf._ptr(opt.log_q_prec), hf._ptr(numpy.ascontiguousarray(task_dmax)), 1e-7,
             hf._stream_ptr())
    assert err == 0, (entry, err)
    torch.cuda.synchronize()
    return vj, vk


def test_noalloc_matches_build_jk_twice():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import water_mole, random_density
    for nw, seed in ((8, 61), (3, 62)):
        mol = water_mole(nw, 7.0, seed, 'cc-pvdz')
        opt = hf._VHFOpt(mol).build()
        dm_d = torch.as_tensor(random_density(mol.nao, 5 * nw, seed), device='cuda')
        ref = _build('BQCbuild_jk', opt, dm_d, torch)
        got = _build('BQCbuild_jk_noalloc', opt, dm_d, torch)
        dj = (got[0] - ref[0]).abs().max().item()
        dk = (got[1] - ref[1]).abs().max().item()
        print(f'no-alloc vs BQCbuild_jk, {nw} H2O: max|dJ| {dj:.2e} max|dK| {dk:.2e}',
              flush=True)
        assert ref[0].abs().max().item() > 0
        assert dj < 1e-10 and dk < 1e-10
