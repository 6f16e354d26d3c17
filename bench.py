"""Benchmark of the direct-SCF J/K build on one MI355X.

Default workload (``--workload c2``): 64 waters (18 A cube, seed 2),
cc-pVDZ, 1536 spherical AOs, D = 2 C C^T from a QR-orthonormalised seeded
1536 x 320 Gaussian (seed 102), Schwarz cutoff 1e-13, FP64 J and K.  A step
is one full ``get_jk`` build (D transform, density bounds, task planning, all
quartet kernels, symmetrisation, back-transform) with the density already on
the GPU; the plan (pair lists, pair data, Schwarz bounds) is built once
before timing, as the reference caches it across SCF iterations
(cuobc/scf/hf.py:237-246).

Other workloads: ``c1`` (8 waters cc-pVDZ), ``c4`` (128 waters cc-pVTZ,
7424 AOs, f shells) and ``c3``: 256 molecules of 4-40 waters (seeded), def2-SVP,
built together by one batched launch set per integral class and replayed
from a captured HIP graph; a step is one build of all 256 molecules and
``value`` is molecule builds/s.  ``--precision mixed`` computes quartets
whose Schwarz x density bound is below 1e-7 in FP32 (FP64 roots, prefactors
and accumulators) and reports |dE| (E = 1/2 Tr D(J - K/2)), max|dJ|, max|dK|
against an FP64 build of the same input.  Work-item tiles default by
molecule size; ``--autotune`` tunes them per class before timing.  ``c5``
feeds Poisson arrivals of RHF requests (5-120 waters, def2-SVP) into the
serving loop (``--qps``, ``--duration``, ``--lanes``) and reports latency
percentiles.

The default c1/c2 run adds two objects to the line: ``mixed``, the same
workload timed in mixed precision with its ratio to FP64 and |dE|, max|dJ|,
max|dK| against the FP64 build, and ``parity_vs_oracle``, the FP64 and mixed
results against a full CPU-oracle build of the same input (about half a
minute on 16 host threads at c2; ``--no-mixed``, ``--no-oracle`` skip them).

Prints one JSON line.  ``roofline`` is for the kernel class that takes the
most time: its algorithmic FLOPs per launch (quartet count x the per-quartet
FLOP model) over its mean launch time, measured with HIP events on the
build's stream during the timed steps.  ``cpu_baseline`` times the CPU
oracle (C + OpenMP restatement of the same algorithm) on a sample and scales
to the metric's unit.
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed; unpacked issue is half)
HBM_PEAK_GBS = 8000.0
METRIC = 'J/K builds/s, p50/p99 latency; mixed vs FP64 at |dE|<=1e-8 Eh'


def _percentile(x, p):
    return float(numpy.percentile(numpy.asarray(x), p))


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def _threads():
    return int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))


def cpu_baseline(mol, dm, stride, stride1=64):
    """Oracle J/K on every stride-th kl pair of every task, scaled to builds/s,
    on all host threads; and on one thread over every stride1-th kl pair."""
    import ctypes
    from oracle import oracle as orc
    plan = orc.build_plan(mol)
    t0 = time.perf_counter()
    _, _, nq = orc.get_jk(mol, dm, plan, return_count=True, kl_stride=stride)
    dt = time.perf_counter() - t0
    lib = orc.lib()
    nthr = lib.oracle_get_threads()
    lib.oracle_set_threads(ctypes.c_int(1))
    try:
        t1 = time.perf_counter()
        _, _, nq1 = orc.get_jk(mol, dm, plan, return_count=True, kl_stride=stride1)
        dt1 = time.perf_counter() - t1
    finally:
        lib.oracle_set_threads(ctypes.c_int(nthr))
    return {'value': 1.0 / (dt * stride), 'unit': 'J/K builds/s', 'cores': nthr,
            'kind': 'port',
            'sample': f'oracle on every {stride}th kl pair of every task '
                      f'({nq} quartets, {dt:.1f} s on {nthr} threads of {_cpu_model()}); '
                      f'scaled x{stride} to one build',
            'one_thread': {'value': 1.0 / (dt1 * stride1), 'cores': 1,
                           'sample': f'every {stride1}th kl pair ({nq1} quartets, {dt1:.1f} s '
                                     f'on 1 thread); scaled x{stride1}'}}


def cpu_baseline_batch(mols, dms, nsample, stride, frac):
    """Oracle on the first nsample molecules of the batch (every stride-th kl
    pair), scaled to molecule builds/s of the whole batch: frac is the sample's
    share of the batch's algorithmic FLOPs."""
    from oracle import oracle as orc
    t0 = time.perf_counter()
    nq = 0
    for m, d in zip(mols[:nsample], dms[:nsample]):
        _, _, n = orc.get_jk(m, d, return_count=True, kl_stride=stride)
        nq += n
    dt = time.perf_counter() - t0
    batch_s = dt * stride / frac
    return {'value': len(mols) / batch_s, 'unit': 'molecule J/K builds/s',
            'cores': _threads(), 'kind': 'port',
            'sample': f'oracle on molecules 0..{nsample - 1} of the batch, every {stride}th kl '
                      f'pair ({nq} quartets, {dt:.1f} s on {_threads()} threads of '
                      f'{_cpu_model()}); scaled x{stride} and by the sample\'s share '
                      f'({frac:.4f}) of the batch FLOPs'}


KIND_NAMES = {0: 'jk_quartets', 1: 'jk_rows', 4: 'jk_wave'}


def kernel_name(cls, kind, split=False):
    """Template name of a profiled launch: class code li*1000+lj*100+lk*10+ll,
    kind 0 quartet / 1 row / 4 wavefront kernel, +2 for the FP32 instance, +8
    for the b-side pass of a split row class.  split names the a-side and
    b-side row kernels of a class together."""
    ls = (cls // 1000, cls // 100 % 10, cls // 10 % 10, cls % 10)
    t = 'float' if kind & 2 else 'double'
    if split:
        return ' + '.join('%s<%d,%d,%d,%d,%s,%d>' % (KIND_NAMES[1], *ls, t, r) for r in (1, 2))
    return '%s<%d,%d,%d,%d,%s>' % (KIND_NAMES[kind & ~10], *ls, t)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the same
    counters' reading for a streaming kernel of known size), or None."""
    path = os.path.join(ROOT, 'profiles', f'traffic_{workload}.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    kernels = rec.get('kernels', {})
    # the profiler names a one-pass row kernel with its role argument (0)
    k = kernels.get(kernel)
    if k is None and kernel.startswith('jk_rows<') and kernel.count(',') == 4:
        k = kernels.get(kernel[:-1] + ',0>')
    return None if k is None else float(k['hbm_bytes_per_launch'])


def copy_probe(torch, nbytes=1 << 29, reps=3):
    """A streaming kernel that reads nbytes and writes nbytes (dst = -src, an
    elementwise kernel rather than a DMA copy), for calibrating HBM byte
    counters."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst


def roofline(recs, steps, workload):
    """Dominant (class, kind) over the records of the timed steps."""
    by_cls = {}
    for r in recs:
        # the two passes of a split row class (kind 1 and 9) count as one
        # kernel: their launches together evaluate and digest the class
        d = by_cls.setdefault((r['cls'], r['kind'] & ~8),
                              {'ms': 0.0, 'n': 0, 'flops': 0.0, 'split': False})
        d['ms'] += r['ms']
        d['n'] += 0 if r['kind'] & 8 else 1
        d['flops'] += r['flops']
        d['split'] = d['split'] or bool(r['kind'] & 8)
    (cls, kind), dom = max(by_cls.items(), key=lambda kv: kv[1]['ms'])
    avg_ms = dom['ms'] / dom['n']
    flops_per_launch = dom['flops'] / dom['n']
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if kind & 2 else FP64_PEAK_TFLOPS
    name = kernel_name(cls, kind, dom['split'])
    kern_ms = sum(v['ms'] for v in by_cls.values()) / steps
    top = sorted(by_cls.items(), key=lambda kv: -kv[1]['ms'])[:12]
    classes = {kernel_name(c, k, v['split']): {
        'ms_per_build': v['ms'] / steps,
        'tflops': v['flops'] / (v['ms'] * 1e-3) / 1e12 if v['ms'] > 0 else 0.0}
        for (c, k), v in top}
    return {'bound': 'valu-fp32' if kind & 2 else 'valu-fp64', 'kernel': name,
            'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s', 'frac': achieved / peak,
            'traffic': measured_traffic(workload, name), 'launches': dom['n'],
            'avg_launch_ms': avg_ms, 'flops_per_launch': flops_per_launch,
            'top_classes': classes}, kern_ms


def log(msg):
    """Progress line on stderr (the JSON result goes to stdout)."""
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def energy_terms(dm, vj, vk):
    return 0.5 * float(numpy.sum(dm * (vj - 0.5 * vk)))


def run_single(args, torch):
    from byteqc_amd.scf import hf, autotune
    from byteqc_amd.systems import workload
    mol, dm, spec = workload(args.workload)
    t0 = time.perf_counter()
    opt = hf._VHFOpt(mol).build()
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    dm_d = torch.as_tensor(dm, device='cuda')

    def step(prec=args.precision):
        return hf.get_jk(mol, dm_d, vhfopt=opt, return_device=True, precision=prec)

    log(f'plan {plan_s:.2f} s')
    tuned = None
    if args.autotune:
        table = autotune.tune(step, nao=int(opt.coeff.shape[0]))
        tuned = dict(zip(('default_ms', 'tuned_ms'), autotune.gain(table)))
        log(f'autotuned {len(table)} (kind, class, precision) tiles: {tuned}')
    ref = None
    if args.precision != 'fp64':
        t1 = time.perf_counter()
        rj, rk = step('fp64')
        ref = (rj.cpu().numpy(), rk.cpu().numpy())
        log(f'FP64 reference build {time.perf_counter() - t1:.2f} s')
    for i in range(args.warmup):
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f'warm-up build {i}: {time.perf_counter() - t1:.3f} s')
    torch.cuda.synchronize()
    hf.profile_read()
    hf.profile_enable(True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        vj, vk = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        if lat[-1] > 2.0:
            log(f'build {len(lat) - 1}: {lat[-1]:.3f} s')
    total = time.perf_counter() - t_start
    hf.profile_enable(False)
    recs = hf.profile_read()
    flops = hf.last_build_flops()
    roof, kern_ms = roofline(recs, args.steps, args.workload)
    prec_txt = 'FP64' if args.precision == 'fp64' else 'mixed FP32/FP64 (FP32 below 1e-7)'
    out = {
        'metric': METRIC, 'value': args.steps / total, 'unit': 'J/K builds/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * total / args.steps,
        'p50_ms': 1e3 * _percentile(lat, 50), 'p99_ms': 1e3 * _percentile(lat, 99),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64' if args.precision == 'fp64' else 'f64+f32',
        'data': 'synthetic',
        'config': {'workload': f"{args.workload}: {spec['nwater']} H2O {spec['basis']} "
                               f"({mol.nao} AOs), {prec_txt} J+K, Schwarz 1e-13",
                   'nao': mol.nao, 'nao_cart_decontracted': int(opt.coeff.shape[0]),
                   'seed': spec['seed'], 'dm_seed': spec['dm_seed'],
                   'precision': args.precision, 'parallelism': 'single GPU'},
        'build_tflops': (flops[0] + flops[1]) / (total / args.steps) / 1e12,
        'algorithmic_flops_per_build': {'fp64': flops[0], 'fp32': flops[1]},
        'kernel_ms_per_build': kern_ms,
        'plan_s': plan_s,
        'roofline': roof,
    }
    if tuned:
        out['autotune'] = tuned
    if ref is not None:
        j, k = vj.cpu().numpy(), vk.cpu().numpy()
        out['mixed_vs_fp64'] = {
            'dE': abs(energy_terms(dm, j, k) - energy_terms(dm, *ref)),
            'max_dJ': float(numpy.abs(j - ref[0]).max()),
            'max_dK': float(numpy.abs(k - ref[1]).max()),
            'fp32_flop_share': flops[1] / max(flops[0] + flops[1], 1.0)}
    elif not args.no_mixed:
        out['mixed'] = mixed_leg(args, torch, step, dm, out['ms_per_step'])
    if args.copy_probe:
        copy_probe(torch)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(mol, dm, args.cpu_stride)
    if not args.no_oracle and args.precision == 'fp64':
        out['parity_vs_oracle'] = oracle_parity(mol, dm, step)
    return out


def mixed_leg(args, torch, step, dm, fp64_ms):
    """The same workload in mixed precision (FP32 below a Schwarz x density
    bound of 1e-7): timed like the FP64 steps, with |dE| and max|dJ|, max|dK|
    against an FP64 build of the same input."""
    from byteqc_amd.scf import hf
    rj, rk = step('fp64')
    ref = (rj.cpu().numpy(), rk.cpu().numpy())
    for _ in range(args.warmup):
        step('mixed')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vj, vk = step('mixed')
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    fl = hf.last_build_flops()
    j, k = vj.cpu().numpy(), vk.cpu().numpy()
    out = {'ms_per_step': ms, 'builds_s': 1e3 / ms, 'fp64_ms_per_step': fp64_ms,
           'speedup_vs_fp64': fp64_ms / ms,
           'fp32_flop_share': fl[1] / max(fl[0] + fl[1], 1.0),
           'vs_fp64': {'dE': abs(energy_terms(dm, j, k) - energy_terms(dm, *ref)),
                       'max_dJ': float(numpy.abs(j - ref[0]).max()),
                       'max_dK': float(numpy.abs(k - ref[1]).max())}}
    log(f"mixed: {ms:.1f} ms per build ({out['speedup_vs_fp64']:.3f}x FP64), "
        f"|dE| {out['vs_fp64']['dE']:.2e} Eh")
    return out


def oracle_parity(mol, dm, step):
    """FP64 and mixed J/K of the benchmark workload against the CPU oracle's
    full build (max |dJ|, max |dK|, |dE| with E = 1/2 Tr D(J - K/2))."""
    from oracle import oracle as orc
    t0 = time.perf_counter()
    jref, kref = orc.get_jk(mol, dm)
    dt = time.perf_counter() - t0
    eref = energy_terms(dm, jref, kref)
    out = {'oracle_s': dt, 'oracle_threads': _threads()}
    for prec in ('fp64', 'mixed'):
        vj, vk = step(prec)
        j, k = vj.cpu().numpy(), vk.cpu().numpy()
        out[prec] = {'max_dJ': float(numpy.abs(j - jref).max()),
                     'max_dK': float(numpy.abs(k - kref).max()),
                     'dE': abs(energy_terms(dm, j, k) - eref)}
    log(f"oracle parity ({dt:.1f} s): FP64 max|dJ| {out['fp64']['max_dJ']:.2e} "
        f"max|dK| {out['fp64']['max_dK']:.2e}; mixed |dE| {out['mixed']['dE']:.2e} Eh")
    return out


def run_batch(args, torch):
    from byteqc_amd.scf import hf, autotune
    from byteqc_amd.scf.batch import BatchJK
    from byteqc_amd.systems import batch_specs, water_mole, random_density
    specs = batch_specs(args.nmol)
    mols = [water_mole(s['nwater'], s['box'], s['seed'], s['basis']) for s in specs]
    dms = [torch.as_tensor(random_density(m.nao_nr(), 5 * s['nwater'], s['dm_seed']),
                           device='cuda') for m, s in zip(mols, specs)]
    log(f'{len(mols)} molecules, {sum(m.nao_nr() for m in mols)} AOs in total')
    t0 = time.perf_counter()
    opts = []
    for i, m in enumerate(mols):
        opts.append(hf._VHFOpt(m).build())
        if (i + 1) % 32 == 0:
            log(f'planned {i + 1} molecules, {time.perf_counter() - t0:.1f} s')
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    bj = BatchJK(mols, vhfopts=opts, precision=args.precision)
    t0 = time.perf_counter()
    bj.build(dms)
    torch.cuda.synchronize()
    log(f'first batched build {time.perf_counter() - t0:.2f} s')
    tuned = None
    if args.autotune:
        # tiles are per class, so a sub-batch with the same size mix suffices
        sub = BatchJK(mols[:args.tune_nmol], vhfopts=opts[:args.tune_nmol],
                      precision=args.precision)
        table = autotune.tune(lambda: sub.build(dms[:args.tune_nmol]), nao=max(sub.nao_c))
        del sub
        tuned = dict(zip(('default_ms', 'tuned_ms'), autotune.gain(table)))
        tuned['tuned_on_molecules'] = args.tune_nmol
        log(f'autotuned {len(table)} (kind, class, precision) tiles: {tuned}')
    for _ in range(args.warmup):
        bj.build(dms)
    torch.cuda.synchronize()
    # eager batched builds (planning included), profiled per launch
    hf.profile_read()
    hf.profile_enable(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        bj.build(dms)
    torch.cuda.synchronize()
    eager = time.perf_counter() - t1
    hf.profile_enable(False)
    recs = hf.profile_read()
    flops = hf.last_build_flops()
    roof, kern_ms = roofline(recs, args.steps, args.workload)
    log(f'eager: {eager / args.steps:.3f} s per batched build')
    # graph replays: the captured kernels of the same build
    bj.capture(dms)
    bj.replay()
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t2 = time.perf_counter()
        bj.replay()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t2)
    total = time.perf_counter() - t_start
    log(f'graph: {total / args.steps:.3f} s per batched build')
    bj.release()
    nw = [s['nwater'] for s in specs]
    out = {
        'metric': METRIC, 'value': len(mols) * args.steps / total,
        'unit': 'molecule J/K builds/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': 1e3 * total / args.steps,
        'p50_ms': 1e3 * _percentile(lat, 50), 'p99_ms': 1e3 * _percentile(lat, 99),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64' if args.precision == 'fp64' else 'f64+f32', 'data': 'synthetic',
        'config': {'workload': f'c3: {len(mols)} molecules of {min(nw)}-{max(nw)} H2O '
                               f'(total {sum(nw)}), def2-svp, one batched launch set, '
                               f'HIP-graph replay',
                   'nmol': len(mols), 'nwater_total': sum(nw),
                   'nao_total': sum(m.nao_nr() for m in mols), 'precision': args.precision,
                   'parallelism': 'single GPU'},
        'eager_molecule_builds_s': len(mols) * args.steps / eager,
        'build_tflops': (flops[0] + flops[1]) / (total / args.steps) / 1e12,
        'algorithmic_flops_per_build': {'fp64': flops[0], 'fp32': flops[1]},
        'kernel_ms_per_build': kern_ms, 'plan_s': plan_s, 'roofline': roof,
    }
    if tuned:
        out['autotune'] = tuned
    if args.copy_probe:
        copy_probe(torch)
    if not args.no_cpu:
        ns = args.cpu_sample
        sample_flops = 0.0
        for m, d, o in zip(mols[:ns], dms[:ns], opts[:ns]):
            hf.get_jk(m, d, vhfopt=o, return_device=True, precision=args.precision)
            sample_flops += sum(hf.last_build_flops())
        frac = sample_flops / (flops[0] + flops[1])
        out['cpu_baseline'] = cpu_baseline_batch(mols, [d.cpu().numpy() for d in dms],
                                                 ns, args.cpu_stride, frac)
    return out


def run_serve(args, torch):
    """Poisson arrivals of RHF requests (log-uniform 5-120 waters, def2-SVP)
    into the continuous-batching loop for --duration seconds of arrivals."""
    from byteqc_amd.scf.serve import Server, poisson_workload, run_load
    from byteqc_amd.systems import water_mole, box_for_density
    arr = poisson_workload(args.qps, args.duration, seed=5, lo=args.min_water,
                           hi=args.max_water)
    log(f'{len(arr)} requests over {args.duration} s at {args.qps} QPS, '
        f'{min(a[1] for a in arr)}-{max(a[1] for a in arr)} waters')
    srv = Server(log=log)
    # warm-up request (kernel loading), not counted
    srv.submit(water_mole(3, box_for_density(3), 1, 'def2-svp'))
    srv.drain()
    srv.close()
    lanes = tuple(int(x) for x in args.lanes.split(',')) if args.lanes else (800, 1600)
    srv = Server(log=log, lanes=lanes, init_guess=args.guess)
    reqs, wall = run_load(srv, arr, lambda n, sd: water_mole(n, box_for_density(n), sd,
                                                            'def2-svp'), log=log)
    srv.close()
    lat = [r.latency for r in reqs]
    ok = [r for r in reqs if r.result and r.result.get('converged')]
    st = srv.stats
    return {
        'metric': METRIC, 'value': st['molecule_builds'] / wall,
        'unit': 'molecule J/K builds/s (serving)', 'n_gpus': 1,
        'steps': st['iterations'], 'warmup': 1, 'ms_per_step': 1e3 * wall / st['iterations'],
        'p50_ms': 1e3 * _percentile(lat, 50), 'p99_ms': 1e3 * _percentile(lat, 99),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64+f32', 'data': 'synthetic',
        'config': {'workload': f'c5: RHF requests, Poisson {args.qps} QPS for '
                               f'{args.duration} s, {args.min_water}-{args.max_water} H2O '
                               f'log-uniform, def2-svp, {args.guess} initial guess, mixed J/K '
                               'until |dE|<1e-5 then FP64, converged to 1e-8 Eh',
                   'requests': len(reqs), 'lanes': list(lanes), 'parallelism': 'single GPU'},
        'serving': {'offered_qps': args.qps, 'completed_qps': len(reqs) / wall,
                    'wall_s': wall, 'converged': len(ok), 'requests': len(reqs),
                    'latency_p50_s': _percentile(lat, 50), 'latency_p99_s': _percentile(lat, 99),
                    'latency_max_s': max(lat), 'iterations': st['iterations'],
                    'batched_builds': st['batched_builds'],
                    'mean_batch': float(numpy.mean(st['batch_sizes'])),
                    'max_active': st['max_active'],
                    'max_hbm_in_use_gib': st['max_in_use'] / 2**30,
                    'hbm_budget_gib': srv.hbm_budget / 2**30,
                    'mean_scf_cycles': float(numpy.mean([r.jk_builds for r in reqs])),
                    'loop_seconds': {k: round(v, 3) for k, v in st['seconds'].items()}},
        'roofline': None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='c2', choices=('c1', 'c2', 'c3', 'c4', 'c5'))
    ap.add_argument('--qps', type=float, default=0.5, help='c5: request rate')
    ap.add_argument('--duration', type=float, default=60.0, help='c5: seconds of arrivals')
    ap.add_argument('--lanes', default='', help='c5: AO counts separating the size lanes')
    ap.add_argument('--guess', default='hcore', choices=('hcore', 'atom'),
                    help='c5: initial guess of the SCF requests (core Hamiltonian by default)')
    ap.add_argument('--min-water', type=int, default=5)
    ap.add_argument('--max-water', type=int, default=120)
    ap.add_argument('--precision', default='fp64', choices=('fp64', 'mixed'))
    ap.add_argument('--autotune', action='store_true')
    ap.add_argument('--nmol', type=int, default=256, help='molecules of the c3 batch')
    ap.add_argument('--tune-nmol', type=int, default=32,
                    help='c3: molecules of the sub-batch the autotuner times')
    ap.add_argument('--cpu-stride', type=int, default=8)
    ap.add_argument('--cpu-sample', type=int, default=4,
                    help='c3: molecules the CPU baseline runs')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--row-split', type=int, default=-1, choices=(-1, 0, 1),
                    help='two-pass (a-side / b-side) digestion of the small s/p row classes: '
                         '-1 the built-in per-class rule, 0 none, 1 every candidate')
    ap.add_argument('--persistent', type=int, default=2, choices=(0, 1, 2),
                    help='persistent work-queue grids for the row kernels: '
                         '0 never, 1 for molecules from 2048 Cartesian AOs, 2 always (default)')
    ap.add_argument('--no-mixed', action='store_true',
                    help='skip the mixed-precision leg of the c1/c2/c4 runs')
    ap.add_argument('--no-oracle', action='store_true',
                    help='skip the full CPU-oracle parity check of the c1/c2 runs')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run a 512 MiB streaming kernel (HBM counter calibration)')
    args = ap.parse_args()
    if args.workload == 'c4':
        args.no_oracle = True  # the CPU oracle needs hours for 128 waters cc-pVTZ
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); '
              'multi-GPU sharding of the quartet list is not part of this build.',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from byteqc_amd.lib import load_library
    lib = load_library()
    lib.BQCset_row_split(-1, args.row_split)
    lib.BQCset_launch_mode(args.persistent)
    torch.cuda.set_device(0)
    if args.workload == 'c3':
        out = run_batch(args, torch)
    elif args.workload == 'c5':
        out = run_serve(args, torch)
    else:
        out = run_single(args, torch)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
