"""Benchmark of the direct-SCF J/K build on one MI355X.

Workload (default): 64 waters (18 A cube, seed 2), cc-pVDZ, 1536 spherical
AOs, D = 2 C C^T from a QR-orthonormalised seeded 1536 x 320 Gaussian
(seed 102), Schwarz cutoff 1e-13, FP64 J and K.  A step is one full
``get_jk`` build (D transform, all quartet kernels, symmetrisation,
back-transform) with the density already on the GPU; the plan (pair lists,
pair data, Schwarz bounds) is built once before timing, as the reference
caches it across SCF iterations (cuobc/scf/hf.py:237-246).

Prints one JSON line.  ``value`` is J/K builds per second; p50/p99 build
latency are reported beside it.  ``roofline`` is for the kernel class that
takes the most time: its algorithmic FP64 FLOPs per launch (quartet count x
the per-quartet FLOP model) over its mean launch time, measured with HIP
events on the build's stream during the timed steps.  ``cpu_baseline`` times
the CPU oracle (C + OpenMP restatement of the same algorithm) on a uniform
1/S sample of every task's kl pairs and scales to builds/s.
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def _percentile(x, p):
    return float(numpy.percentile(numpy.asarray(x), p))


def cpu_baseline(mol, dm, stride):
    from oracle import oracle as orc
    plan = orc.build_plan(mol)
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    t0 = time.perf_counter()
    _, _, nq = orc.get_jk(mol, dm, plan, return_count=True, kl_stride=stride)
    dt = time.perf_counter() - t0
    build_s = dt * stride
    cpu = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    cpu = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return {'value': 1.0 / build_s, 'unit': 'J/K builds/s', 'cores': threads,
            'kind': 'port',
            'sample': f'oracle on every {stride}th kl pair of every task '
                      f'({nq} quartets, {dt:.1f} s on {threads} threads of {cpu}); '
                      f'scaled x{stride} to one build'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='c2')
    ap.add_argument('--cpu-stride', type=int, default=8)
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); '
              'multi-GPU sharding of the quartet list is not part of this build.',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from byteqc_amd.scf import hf
    from byteqc_amd.systems import workload
    from byteqc_amd.lib import load_library
    lib = load_library()
    torch.cuda.set_device(0)
    mol, dm, spec = workload(args.workload)
    t0 = time.perf_counter()
    opt = hf._VHFOpt(mol).build()
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    dm_d = torch.as_tensor(dm, device='cuda')

    def step():
        return hf.get_jk(mol, dm_d, vhfopt=opt, return_device=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    hf.profile_enable(True)
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        vj, vk = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    hf.profile_enable(False)
    recs = hf.profile_read()
    flops = hf.last_build_flops()

    # dominant kernel class over the timed steps
    by_cls = {}
    for r in recs:
        key = (r['cls'], r['kind'])
        d = by_cls.setdefault(key, {'ms': 0.0, 'n': 0, 'flops': 0.0})
        d['ms'] += r['ms']
        d['n'] += 1
        d['flops'] += r['flops']
    (cls, kind), dom = max(by_cls.items(), key=lambda kv: kv[1]['ms'])
    avg_ms = dom['ms'] / dom['n']
    flops_per_launch = dom['flops'] / dom['n']
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    kern_ms_total = sum(v['ms'] for v in by_cls.values()) / args.steps
    cname = 'jk_%s<%d,%d,%d,%d>' % ('rows' if kind else 'quartets', cls // 1000,
                                     cls // 100 % 10, cls // 10 % 10, cls % 10)

    builds_s = args.steps / total
    out = {
        'metric': 'J/K builds/s, p50/p99 latency; mixed vs FP64 at |dE|<=1e-8 Eh',
        'value': builds_s,
        'unit': 'J/K builds/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * total / args.steps,
        'p50_ms': 1e3 * _percentile(lat, 50),
        'p99_ms': 1e3 * _percentile(lat, 99),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': f"{args.workload}: {spec['nwater']} H2O {spec['basis']} "
                               f"({mol.nao} AOs), FP64 J+K, Schwarz 1e-13",
                   'nao': mol.nao, 'nao_cart_decontracted': int(opt.coeff.shape[0]),
                   'seed': spec['seed'], 'dm_seed': spec['dm_seed'],
                   'parallelism': 'single GPU'},
        'build_tflops': flops[0] / (total / args.steps) / 1e12,
        'algorithmic_flops_per_build': flops[0],
        'kernel_ms_per_build': kern_ms_total,
        'plan_s': plan_s,
        'roofline': {'bound': 'valu-fp64', 'kernel': cname,
                     'achieved': achieved, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP64_PEAK_TFLOPS, 'traffic': None,
                     'launches': dom['n'], 'avg_launch_ms': avg_ms,
                     'flops_per_launch': flops_per_launch},
    }
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(mol, dm, args.cpu_stride)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
