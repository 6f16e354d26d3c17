"""Parity at the benchmark sizes.

* 64 waters cc-pVDZ (2624 decontracted Cartesian AOs, the default
  large-molecule tiles): J and K from the GPU equal the CPU oracle's to
  1e-10 (the oracle takes about half a minute on 16 host threads).
* 128 waters cc-pVTZ (10368 decontracted Cartesian AOs, f shells): too large
  for the oracle, so size-independent properties: the mixed-precision build
  agrees with the FP64 build to |dE| <= 1e-8 Eh (E = 1/2 Tr D(J - K/2)) and
  max|dJ|, max|dK| <= 1e-8; J and K are symmetric; Tr(DJ) > 0 and K is
  positive semidefinite for the positive semidefinite D.
"""
import numpy
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def test_c2_matches_oracle():
    _torch()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    from oracle import oracle as orc
    mol, dm, _ = workload('c2')
    print('C2: GPU build', flush=True)
    j, k = hf.get_jk(mol, dm)
    print('C2: oracle build', flush=True)
    jref, kref = orc.get_jk(mol, dm)
    dj = numpy.abs(j - jref).max()
    dk = numpy.abs(k - kref).max()
    print(f'C2 64 H2O cc-pVDZ: max|dJ| {dj:.3e} max|dK| {dk:.3e}', flush=True)
    assert dj < 1e-10 and dk < 1e-10


def test_c4_mixed_against_fp64_and_invariants():
    torch = _torch()
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    mol, dm, _ = workload('c4')
    print('C4: plan', flush=True)
    opt = hf._VHFOpt(mol).build()
    d = torch.as_tensor(dm, device='cuda')
    print('C4: FP64 build', flush=True)
    j0, k0 = hf.get_jk(mol, d, vhfopt=opt, return_device=True)
    print('C4: mixed build', flush=True)
    j1, k1 = hf.get_jk(mol, d, vhfopt=opt, return_device=True, precision='mixed',
                       fp32_cut=1e-7)
    fl = hf.last_build_flops()

    def energy(j, k):
        return 0.5 * float(torch.sum(d * (j - 0.5 * k)))

    de = abs(energy(j1, k1) - energy(j0, k0))
    dj = (j1 - j0).abs().max().item()
    dk = (k1 - k0).abs().max().item()
    print(f'C4 mixed vs FP64: |dE| {de:.3e} Eh, max|dJ| {dj:.3e}, max|dK| {dk:.3e}, '
          f'FP32 share of FLOPs {fl[1] / fl.sum():.3f}', flush=True)
    assert fl[1] > 0
    assert de <= 1e-8 and dj <= 1e-8 and dk <= 1e-8
    for m in (j0, k0):
        assert (m - m.T).abs().max().item() <= 1e-12 * m.abs().max().item()
    assert float(torch.sum(d * j0)) > 0
    w = torch.linalg.eigvalsh(k0)
    print(f'C4 K eigenvalues: min {w[0].item():.3e} max {w[-1].item():.3e}', flush=True)
    assert w[0].item() >= -1e-9 * w[-1].item()
