"""Tile autotuner for the J/K kernels.

The tunable tile is the number of Q pairs per work item (``ck``) of the
row kernels (kind 1: 64 P pairs x ck Q pairs per wavefront) and of the
wavefront-per-P-pair kernels (kind 4: 1 P pair x ck Q pairs per workgroup),
separately for every integral class and precision.  A small ck gives more
work items (better fill of the 256 CUs for classes with few pairs) at the
cost of more flushes of the lane accumulators (J(a,b), K(a,c), K(b,c));
a large ck gives fewer flushes and fewer items.

``tune(run)`` times every candidate with the library's per-launch HIP-event
records (one launch per class, kind and precision per build), keeps the
fastest ck per (kind, class, precision) and installs it with
``BQCset_tile``.  Without tuning the library uses default tiles chosen by
molecule size (large molecules want long Q ranges per item, batches of
small ones short), measured the same way over the benchmark workloads.  Any tile gives the same quartet set and the same sums
(only the order of floating-point accumulation differs).

The reference has fixed launch shapes per class (nr_jk_driver.cu:46-225).
"""
from __future__ import annotations

import json
from collections import defaultdict

import torch

from byteqc_amd.lib import load_library
from byteqc_amd.scf import hf

ROWS, WAVE = 1, 4
CANDIDATES = {ROWS: (16, 32, 64, 128, 256), WAVE: (8, 16, 32, 64, 128)}
DEFAULTS = {ROWS: 64, WAVE: 32}     # reference point for a table without recorded defaults


def reset():
    """Drop every tuned tile and restore the library's default tiles."""
    load_library().BQCset_tile(0, -1, 0, 0)


def current_tile(kind, cls, prec, nao):
    """Tile a build of a molecule with nao Cartesian AOs planned now uses for
    (kind, cls, prec): a tuned tile if one is installed, else the default."""
    return int(load_library().BQCget_tile(int(kind), int(cls), int(prec), int(nao)))


def _timed(run, reps):
    run()                       # plans with the current tiles, warms caches
    torch.cuda.synchronize()
    hf.profile_read()
    hf.profile_enable(True)
    try:
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
    finally:
        hf.profile_enable(False)
    return hf.profile_read()


def tune(run, candidates=None, reps=1, nao=0):
    """Tune tiles for the builds done by ``run()`` (one eager J/K build per
    call, any number of molecules; nao: Cartesian AO count of the largest,
    which picks the default tiles recorded as 'builtin').  Returns
    {(kind, cls, prec): {'ck': best, 'ms': {ck: ms per build}, 'builtin':
    default ck}} and leaves the best tiles installed."""
    lib = load_library()
    candidates = candidates or CANDIDATES
    reset()
    times = defaultdict(dict)
    for kind, cks in candidates.items():
        for ck in cks:
            if lib.BQCset_tile(kind, -1, 0, int(ck)):
                raise ValueError(f'bad tile {ck}')
            agg = defaultdict(float)
            for r in _timed(run, reps):
                # kind: 1 row / 4 wavefront kernel, + 2 FP32, + 8 b-side pass
                # of a split class (both passes of a class share its tile)
                if (r['kind'] & ~10) != kind:
                    continue
                agg[(r['cls'], (r['kind'] >> 1) & 1)] += r['ms']
            for (cls, prec), ms in agg.items():
                times[(kind, cls, prec)][int(ck)] = ms / reps
        reset()
    table = {}
    for key, per in times.items():
        best = min(per, key=per.get)
        table[key] = {'ck': best, 'ms': per, 'builtin': current_tile(*key, nao)}  # after reset()
    install(table)
    return table


def install(table):
    """Install a tuned table (as returned by ``tune`` or ``load``)."""
    lib = load_library()
    reset()
    for (kind, cls, prec), rec in table.items():
        if lib.BQCset_tile(int(kind), int(cls), int(prec), int(rec['ck'])):
            raise ValueError(f'bad tile entry {(kind, cls, prec)}: {rec}')


def gain(table):
    """(ms per build with the default tiles, ms with tuned tiles) over the
    table."""
    base = sum(rec['ms'].get(rec.get('builtin', DEFAULTS[k]), min(rec['ms'].values()))
               for (k, _, _), rec in table.items())
    tuned = sum(rec['ms'][rec['ck']] for rec in table.values())
    return base, tuned


def save(table, path):
    rows = [{'kind': k, 'cls': c, 'prec': p, 'ck': rec['ck'], 'builtin': rec.get('builtin'),
             'ms': {str(ck): ms for ck, ms in rec['ms'].items()}}
            for (k, c, p), rec in sorted(table.items())]
    with open(path, 'w') as f:
        json.dump(rows, f, indent=1)


def load(path):
    with open(path) as f:
        rows = json.load(f)
    out = {}
    for r in rows:
        rec = {'ck': r['ck'], 'ms': {int(k): v for k, v in r['ms'].items()}}
        if r.get('builtin') is not None:
            rec['builtin'] = r['builtin']
        out[(r['kind'], r['cls'], r['prec'])] = rec
    return out

