"""GPU tests of the batched J/K builds (many molecules per launch set) and of
their HIP-graph capture, against single-molecule builds and the oracle."""
import numpy
import pytest

pytestmark = pytest.mark.gpu


def _torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _batch_inputs():
    from byteqc_amd.systems import water_mole, random_density
    specs = [(1, 4.0, 11, 'cc-pvdz'), (2, 6.0, 12, 'def2-svp'), (3, 7.0, 13, 'cc-pvdz'),
             (2, 6.0, 14, 'cc-pvtz')]
    mols = [water_mole(n, box, seed, basis) for n, box, seed, basis in specs]
    dms = [random_density(m.nao, m.nelectron // 2, 200 + i) for i, m in enumerate(mols)]
    return mols, dms


def _maxdiff(a, b):
    return float(numpy.abs(numpy.asarray(a) - numpy.asarray(b)).max())


def test_batch_matches_single_builds_and_oracle():
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    from oracle import oracle as orc
    mols, dms = _batch_inputs()
    out = BatchJK(mols).build(dms)
    for i, (m, d) in enumerate(zip(mols, dms)):
        j1, k1 = hf.get_jk(m, d)
        dj, dk = _maxdiff(out[i][0].cpu(), j1), _maxdiff(out[i][1].cpu(), k1)
        print(f'mol {i} ({m.nao} AOs): batch vs single max|dJ| {dj:.2e} max|dK| {dk:.2e}',
              flush=True)
        assert dj < 1e-11 and dk < 1e-11
    jr, kr = orc.get_jk(mols[0], dms[0])
    assert _maxdiff(out[0][0].cpu(), jr) < 1e-10
    assert _maxdiff(out[0][1].cpu(), kr) < 1e-10


def test_batch_j_only_and_k_only():
    _torch_gpu()
    from byteqc_amd.scf.batch import BatchJK
    mols, dms = _batch_inputs()
    both = BatchJK(mols).build(dms)
    jo = BatchJK(mols, with_k=False).build(dms)
    ko = BatchJK(mols, with_j=False).build(dms)
    for b, j, k in zip(both, jo, ko):
        assert j[1] is None and k[0] is None
        assert _maxdiff(b[0].cpu(), j[0].cpu()) < 1e-11
        assert _maxdiff(b[1].cpu(), k[1].cpu()) < 1e-11


def test_graph_replays_match_eager():
    """Capture once, replay several times with other GPU work in between."""
    torch = _torch_gpu()
    from byteqc_amd.scf.batch import BatchJK
    mols, dms = _batch_inputs()
    eager = BatchJK(mols).build(dms)
    bj = BatchJK(mols).capture(dms)
    for rep in range(4):
        x = torch.randn(300, 300, dtype=torch.float64, device='cuda')
        (x @ x).sum().item()
        torch.cuda.synchronize()
        got = bj.replay()
        torch.cuda.synchronize()
        for i, (e, g) in enumerate(zip(eager, got)):
            dj, dk = _maxdiff(e[0].cpu(), g[0].cpu()), _maxdiff(e[1].cpu(), g[1].cpu())
            print(f'replay {rep} mol {i}: max|dJ| {dj:.2e} max|dK| {dk:.2e}', flush=True)
            assert dj < 1e-11 and dk < 1e-11
    bj.release()


def test_batch_mixed_precision_energy():
    """Mixed batched build: |E_J - E_K/2| change <= 1e-8 Eh against FP64."""
    _torch_gpu()
    from byteqc_amd.scf.batch import BatchJK
    mols, dms = _batch_inputs()
    f64 = BatchJK(mols).build(dms)
    mix = BatchJK(mols, precision='mixed').build(dms)
    for d, a, b in zip(dms, f64, mix):
        ea = 0.5 * numpy.sum(d * (a[0].cpu().numpy() - 0.5 * a[1].cpu().numpy()))
        eb = 0.5 * numpy.sum(d * (b[0].cpu().numpy() - 0.5 * b[1].cpu().numpy()))
        assert abs(ea - eb) < 1e-8
        assert _maxdiff(a[0].cpu(), b[0].cpu()) < 1e-8
        assert _maxdiff(a[1].cpu(), b[1].cpu()) < 1e-8


def test_autotuned_tiles_keep_results():
    """Any tile choice gives the same J/K (only the accumulation order moves)."""
    _torch_gpu()
    from byteqc_amd.scf import autotune
    from byteqc_amd.scf.batch import BatchJK
    mols, dms = _batch_inputs()
    bj = BatchJK(mols)
    try:
        base = bj.build(dms)
        table = autotune.tune(lambda: bj.build(dms), candidates={1: (16, 256), 4: (8, 128)})
        assert table and all(rec['ck'] in rec['ms'] for rec in table.values())
        for (kind, cls, prec), rec in table.items():
            # the tuned tile is installed; reset() brings back the default one
            assert autotune.current_tile(kind, cls, prec, 100) == rec['ck']
            assert rec['builtin'] in (16, 32, 64, 128, 256)
        got = bj.build(dms)
        for a, b in zip(base, got):
            assert _maxdiff(a[0].cpu(), b[0].cpu()) < 1e-11
            assert _maxdiff(a[1].cpu(), b[1].cpu()) < 1e-11
    finally:
        autotune.reset()
    from byteqc_amd.lib import load_library
    lib = load_library()
    # default tiles by molecule size: s/p rows, 108-component rows, waves
    assert [lib.BQCget_tile(1, 1010, 0, n) for n in (500, 2624, 10368)] == [128, 64, 256]
    assert [lib.BQCget_tile(1, 2120, 1, n) for n in (500, 2624, 10368)] == [128, 64, 128]
    assert [lib.BQCget_tile(4, 2111, 0, n) for n in (500, 2624, 10368)] == [32, 32, 64]
    assert lib.BQCget_tile(2, 0, 0, 100) == -1
    lib.BQCset_tile(1, -1, 0, 128)          # a tile for all classes of a kind beats the defaults
    lib.BQCset_tile(1, 1000, 0, 16)         # and a per-class tile beats that
    try:
        assert all(lib.BQCget_tile(1, c, p, 3000) == 128 for c in (0, 2120) for p in (0, 1))
        assert lib.BQCget_tile(1, 1000, 0, 3000) == 16
    finally:
        autotune.reset()


def test_persistent_and_per_item_launches_agree_on_c3_subbatch():
    """Molecules drawn like the 256-molecule batch (4-40 waters, def2-SVP,
    density 1 g/cm^3): the persistent work-queue launches (default), the
    one-workgroup-per-item launches and a replayed graph of the persistent
    launches give the same J/K, equal to single builds and, for the smallest
    molecule, to the oracle."""
    torch = _torch_gpu()
    from byteqc_amd.lib import load_library
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    from byteqc_amd.systems import batch_specs, water_mole, random_density
    from oracle import oracle as orc
    specs = batch_specs(256)[:6]
    mols = [water_mole(s['nwater'], s['box'], s['seed'], s['basis']) for s in specs]
    dms = [random_density(m.nao, 5 * s['nwater'], s['dm_seed']) for m, s in zip(mols, specs)]
    lib = load_library()
    bj = BatchJK(mols)
    try:
        assert lib.BQCset_launch_mode(2) == 0      # persistent for every molecule size
        pers = bj.build(dms)
        assert lib.BQCset_launch_mode(0) == 0
        stat = bj.build(dms)
        assert lib.BQCset_launch_mode(2) == 0
        bj.capture(dms)
    finally:
        lib.BQCset_launch_mode(2)             # the default
    assert lib.BQCset_launch_mode(3) == 1
    for rep in range(3):
        x = torch.randn(200, 200, dtype=torch.float64, device='cuda')
        (x @ x).sum().item()
        torch.cuda.synchronize()
        rep_out = bj.replay()
        torch.cuda.synchronize()
        for a, b in zip(pers, rep_out):
            assert _maxdiff(a[0].cpu(), b[0].cpu()) < 1e-11
            assert _maxdiff(a[1].cpu(), b[1].cpu()) < 1e-11
    bj.release()
    for i, (m, d) in enumerate(zip(mols, dms)):
        dj = _maxdiff(pers[i][0].cpu(), stat[i][0].cpu())
        dk = _maxdiff(pers[i][1].cpu(), stat[i][1].cpu())
        j1, k1 = hf.get_jk(m, d)
        print(f'{specs[i]["nwater"]} H2O: persistent vs per-item max|dJ| {dj:.2e} '
              f'max|dK| {dk:.2e}', flush=True)
        assert dj < 1e-11 and dk < 1e-11
        assert _maxdiff(pers[i][0].cpu(), j1) < 1e-11
        assert _maxdiff(pers[i][1].cpu(), k1) < 1e-11
    small = min(range(len(mols)), key=lambda i: mols[i].nao)
    jr, kr = orc.get_jk(mols[small], dms[small])
    assert _maxdiff(pers[small][0].cpu(), jr) < 1e-10
    assert _maxdiff(pers[small][1].cpu(), kr) < 1e-10
