"""GPU tests of the continuous-batching request loop: a request's result is
the same alone and in a mix (|dE| <= 1e-9 Eh, max|dD| <= 1e-7), admission
follows the HBM budget, J/K requests match single builds."""
import numpy
import pytest

pytestmark = pytest.mark.gpu


def _torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _mols():
    from byteqc_amd.systems import water_mole, box_for_density
    return [water_mole(n, box_for_density(n), 40 + n, 'def2-svp') for n in (2, 4, 1, 3)]


def test_scf_result_same_alone_and_in_mix():
    _torch_gpu()
    from byteqc_amd.scf.serve import Server
    mols = _mols()
    alone = Server()
    ra = alone.submit(mols[0])
    alone.drain()
    mix = Server()
    rm = mix.submit(mols[0])
    others = [mix.submit(m) for m in mols[1:3]]
    mix.step()
    mix.step()
    late = mix.submit(mols[3])          # joins while the others are in flight
    mix.drain()
    assert mix.stats['max_active'] >= 3
    for r in [ra, rm, late] + others:
        assert r.error is None and r.result['converged'], r.rid
    de = abs(ra.result['e_tot'] - rm.result['e_tot'])
    dd = numpy.abs(ra.result['dm'] - rm.result['dm']).max()
    print(f'alone vs mix: |dE| {de:.2e}  max|dD| {dd:.2e}  cycles {ra.result["cycles"]} '
          f'{rm.result["cycles"]}', flush=True)
    assert de <= 1e-9 and dd <= 1e-7
    assert ra.result['jk_builds']['mixed'] > 0 and ra.result['jk_builds']['fp64'] > 0


def test_serving_matches_standalone_rhf():
    _torch_gpu()
    from byteqc_amd.scf.rhf import RHF
    from byteqc_amd.scf.serve import Server
    mol = _mols()[1]
    srv = Server()
    r = srv.submit(mol)
    srv.submit(_mols()[2])
    srv.drain()
    ref = RHF(mol, precision='mixed_then_fp64', diis_space=srv.scf_defaults['diis_space']).kernel()
    assert abs(r.result['e_tot'] - ref) <= 1e-9


def test_admission_respects_hbm_budget():
    _torch_gpu()
    from byteqc_amd.scf.serve import Server, estimate_footprint
    mols = _mols()
    fp = [estimate_footprint(m, diis_space=12) for m in mols]
    budget = max(fp) + min(fp) // 2          # never two of the larger ones at once
    srv = Server(hbm_budget=budget, diis_space=12)
    reqs = [srv.submit(m) for m in mols]
    srv.drain()
    assert all(r.error is None and r.result['converged'] for r in reqs)
    assert srv.stats['max_in_use'] <= budget
    # FIFO: admission order is submission order
    adm = [r.admitted for r in reqs]
    assert adm == sorted(adm)
    tiny = Server(hbm_budget=4096)
    r = tiny.submit(mols[0])
    assert r.error is not None and r.done and tiny.idle()


def test_jk_requests_in_mix_match_single_builds():
    _torch_gpu()
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.serve import Server
    from byteqc_amd.systems import random_density
    mols = _mols()
    srv = Server()
    dms = [random_density(m.nao_nr(), 5, 300 + i) for i, m in enumerate(mols[:2])]
    scf = srv.submit(mols[2])
    jk = [srv.submit(m, kind='jk', dm=d) for m, d in zip(mols[:2], dms)]
    srv.drain()
    assert scf.result['converged']
    for r, m, d in zip(jk, mols, dms):
        j, k = hf.get_jk(m, d)
        assert numpy.abs(r.result['vj'] - j).max() < 1e-11
        assert numpy.abs(r.result['vk'] - k).max() < 1e-11
        assert r.jk_builds == 1


def test_poisson_load_completes_with_latencies():
    _torch_gpu()
    from byteqc_amd.scf.serve import Server, poisson_workload, run_load
    from byteqc_amd.systems import water_mole, box_for_density
    arr = poisson_workload(qps=4.0, duration=1.5, seed=9, lo=1, hi=4)
    assert arr
    srv = Server()
    reqs, wall = run_load(srv, arr, lambda n, s: water_mole(n, box_for_density(n), s,
                                                             'def2-svp'))
    assert len(reqs) == len(arr) and all(r.result['converged'] for r in reqs)
    lat = [r.latency for r in reqs]
    assert min(lat) > 0 and wall >= arr[-1][0]


def test_size_lanes_run_concurrently_and_keep_results():
    """Two lanes (split at 60 AOs): a small request iterates in its own lane
    while a larger one builds; both give their standalone results."""
    _torch_gpu()
    from byteqc_amd.scf.serve import Server
    from byteqc_amd.systems import water_mole, box_for_density
    big = water_mole(6, box_for_density(6), 7, 'def2-svp')       # 144 AOs
    small = _mols()[2]                                            # 24 AOs
    ref = {}
    for name, m in (('big', big), ('small', small)):
        alone = Server(lanes=(60,))
        r = alone.submit(m)
        alone.drain()
        ref[name] = r.result
    srv = Server(lanes=(60,))
    rb = srv.submit(big)
    rs = srv.submit(small)
    srv.drain()
    assert srv.stats['lane_builds'][0] > 0 and srv.stats['lane_builds'][1] > 0
    for name, r in (('big', rb), ('small', rs)):
        assert r.error is None and r.result['converged']
        assert abs(r.result['e_tot'] - ref[name]['e_tot']) <= 1e-9
        assert numpy.abs(r.result['dm'] - ref[name]['dm']).max() <= 1e-7
    print(f'lanes: builds {srv.stats["lane_builds"]}, latency small {rs.latency:.2f} s '
          f'big {rb.latency:.2f} s', flush=True)
    srv.close()
    assert all(lane.thread is None for lane in srv.lanes)


def test_lane_error_reaches_the_caller():
    """An exception in a lane's host thread (here a density of the wrong
    shape, rejected before any launch) is raised by the caller's step()."""
    _torch_gpu()
    from byteqc_amd.scf.serve import Server
    mol = _mols()[2]
    srv = Server()
    srv.submit(mol, kind='jk', dm=numpy.zeros((mol.nao_nr() + 1, mol.nao_nr() + 1)))
    with pytest.raises(RuntimeError, match='lane failed'):
        srv.drain()
