"""Request loop for J/K and SCF requests with continuous batching.

The reference has no serving layer: its callers run one PySCF SCF at a time
and each ``get_jk`` call launches per task from Python
(cuobc/scf/hf.py:53-227, 253-274).  Here many requests share the GPU:

* A request is either one J/K build for a given density ("jk") or a full
  RHF run to convergence ("scf": atomic-density guess, DIIS, mixed-
  precision J/K until |dE| < switch_de, then FP64).
* Admission by HBM footprint: a request leaves the FIFO queue only when its
  estimated device footprint (plan, pair data, SCF matrices, DIIS history,
  J/K buffers) fits in the budget beside the requests in flight; the oldest
  request goes first and is never overtaken, so large molecules do not starve.
* Initial guess: superposition of atomic densities (cached per element and
  basis), so an SCF request needs few iterations.
* Lanes by size: a request goes to a lane by its number of AOs; each lane
  has its own HIP stream and host thread.  A lane builds J/K for *all* its
  requests in flight in one batched launch set (``BQCbuild_jk_batch``: one
  launch per integral class and kernel kind covering every molecule), then
  advances each SCF by one step, and repeats.  The lanes run concurrently on
  the device and on the host, so small molecules iterate many times while
  one build of large molecules runs and their latency does not follow the
  largest request in flight.
* Continuous batching: requests that arrive while others run join their
  lane at its next build; finished ones leave and free their memory at once.
* Per-request state (plan, pair lists, pair data, S, H, X, D, DIIS history)
  stays on the device between iterations.

Requests in one batched build share the number of Schwarz bins and the
precision; the loop groups requests by both, so a request's quartet set and
its result do not depend on which other requests are in flight.
"""
from __future__ import annotations

import bisect
import collections
import itertools
import threading
import time

import numpy
import torch

from byteqc_amd.scf import hf
from byteqc_amd.scf.batch import BatchJK
from byteqc_amd.scf.rhf import RHF

# device matrices of one SCF request besides the J/K build buffers: S, H, X,
# D, F, MO coefficients, orbital-basis work matrices, J, K and D of the last
# build (incremental builds), DIIS history (F and error per slot)
_SCF_MATS = 13
_JK_CART_MATS = 5   # C D C^T, J, K, two transform temporaries (Cartesian size)


class Request:
    """One J/K or SCF request and its result."""

    def __init__(self, rid, mol, kind='scf', dm=None, arrival=None, **scf_kwargs):
        if kind not in ('scf', 'jk'):
            raise ValueError(f'unknown request kind {kind!r}')
        if kind == 'jk' and dm is None:
            raise ValueError('a jk request needs a density')
        self.rid = rid
        self.mol = mol
        self.kind = kind
        self.dm = dm
        self.scf_kwargs = scf_kwargs
        self.arrival = time.perf_counter() if arrival is None else arrival
        self.admitted = None
        self.finished = None
        self.footprint = None
        self.opt = None
        self.scf = None
        self.result = None
        self.error = None
        self.jk_builds = 0

    @property
    def done(self):
        return self.finished is not None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def estimate_footprint(mol, diis_space=8):
    """Device bytes of one request, estimated from the basis before planning:
    Cartesian (decontracted) and spherical matrices, pair data of all shell
    pairs of the decontracted basis, Schwarz bounds."""
    pmol, coeff = hf.basis_seg_contraction(mol)
    nao_c, nao = coeff.shape
    nbas = pmol.nbas
    nprim = numpy.asarray(pmol._bas)[:, 2]
    npairs = nbas * (nbas + 1) // 2
    prim_pairs = float(numpy.sum(nprim)) ** 2 / 2
    mats = 8 * (_JK_CART_MATS * nao_c * nao_c + (_SCF_MATS + 2 * diis_space) * nao * nao
                + nao_c * nao)
    pairs = 8 * nbas * nbas + 12 * npairs + 40 * prim_pairs
    return int(mats + pairs)


class _Lane:
    """Requests of one size range, their stream and their host thread."""

    def __init__(self):
        self.stream = None
        self.reqs = []
        self.thread = None


class Server:
    """Continuous-batching request loop on one GPU.

    lanes: AO counts separating the size lanes (ascending); a request with
    nao AOs joins lane bisect_right(lanes, nao).  Each lane runs its builds
    and SCF steps in a host thread of its own on its own stream, so the host
    work of one lane (planning, diagonalisation) overlaps the builds of the
    others; ``step()`` admits requests from the calling thread and returns
    the requests that finished since the previous call."""

    def __init__(self, hbm_budget=None, max_active=64, precision='mixed_then_fp64',
                 conv_tol=1e-8, switch_de=1e-5, max_cycle=80, fp32_cut=1e-7, log=None,
                 lanes=(800, 1600), init_guess='atom', diis_space=12):
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.lane_bounds = tuple(sorted(lanes))
        self.lanes = [_Lane() for _ in range(len(self.lane_bounds) + 1)]
        self._setup_stream = None
        self.max_active = max_active
        self.scf_defaults = dict(precision=precision, conv_tol=conv_tol, switch_de=switch_de,
                                 max_cycle=max_cycle, fp32_cut=fp32_cut, init_guess=init_guess,
                                 diis_space=diis_space)
        self.fp32_cut = fp32_cut
        self.queue = collections.deque()
        self.active = []
        self.completed = []
        self.in_use = 0
        self._ids = itertools.count()
        self.log = log
        self._cv = threading.Condition()
        self._done = []        # finished since the last step()
        self._error = None     # first exception raised in a lane thread
        self._closed = False
        # seconds spent admitting (plans, one-electron integrals, initial
        # guess), planning and launching batched J/K builds, waiting for them
        # on the device and in the SCF steps; lane times add up over threads
        self.stats = {'iterations': 0, 'batched_builds': 0, 'molecule_builds': 0,
                      'max_active': 0, 'max_in_use': 0, 'batch_sizes': [],
                      'seconds': {'admit': 0.0, 'launch': 0.0, 'wait': 0.0, 'scf': 0.0},
                      'lane_builds': [0] * len(self.lanes)}

    # -- requests ----------------------------------------------------------------
    def submit(self, mol, kind='scf', dm=None, arrival=None, **scf_kwargs):
        r = Request(next(self._ids), mol, kind, dm, arrival, **scf_kwargs)
        r.footprint = estimate_footprint(mol, self.scf_defaults['diis_space'])
        if r.footprint > self.hbm_budget:
            r.error = (f'request needs ~{r.footprint / 2**30:.2f} GiB of HBM, above the '
                       f'budget of {self.hbm_budget / 2**30:.2f} GiB')
            r.finished = time.perf_counter()
            with self._cv:
                self.completed.append(r)
            return r
        self.queue.append(r)
        return r

    def _admit(self):
        if self._closed:
            raise RuntimeError('server is closed')
        while self.queue:
            with self._cv:
                if len(self.active) >= self.max_active:
                    break
                r = self.queue[0]
                if self.active and self.in_use + r.footprint > self.hbm_budget:
                    break
            self.queue.popleft()
            r.admitted = time.perf_counter()
            li = bisect.bisect_right(self.lane_bounds, r.mol.nao_nr())
            lane = self.lanes[li]
            if lane.stream is None:
                lane.stream = torch.cuda.Stream()
            # set-up runs on a stream of its own, so it does not queue behind
            # a build in flight in the request's lane
            if self._setup_stream is None:
                self._setup_stream = torch.cuda.Stream()
            with hf.on_stream(self._setup_stream):
                r.opt = hf._VHFOpt(r.mol).build()
                if r.kind == 'scf':
                    kw = dict(self.scf_defaults)
                    kw.update(r.scf_kwargs)
                    r.scf = RHF(r.mol, vhfopt=r.opt, stream=self._setup_stream, **kw).build()
                    r.scf.stream = lane.stream
                else:
                    r.dm = torch.as_tensor(r.dm, dtype=torch.float64, device=hf._device())
            self._setup_stream.synchronize()
            with self._cv:
                self.in_use += r.footprint
                self.active.append(r)
                lane.reqs.append(r)
                self.stats['max_active'] = max(self.stats['max_active'], len(self.active))
                self.stats['max_in_use'] = max(self.stats['max_in_use'], self.in_use)
                if lane.thread is None:
                    lane.thread = threading.Thread(target=self._lane_loop, args=(li,),
                                                   daemon=True)
                    lane.thread.start()
                self._cv.notify_all()

    def _retire(self, r):
        """Record a finished request's result (lane thread, on its stream)."""
        r.finished = time.perf_counter()
        if r.opt is not None:
            r.opt.clear()
        r.opt = None
        if r.scf is not None:
            s = r.scf
            r.result = {'e_tot': s.e_tot, 'converged': s.converged, 'cycles': s.cycles,
                        'jk_builds': dict(s.jk_builds), 'dm': s.dm.cpu().numpy(),
                        'mo_energy': s.mo_energy.cpu().numpy()}
            r.scf = None
        if self.log:
            self.log(f'request {r.rid} ({r.kind}, {r.mol.nao_nr()} AOs) done in '
                     f'{r.latency:.2f} s, {r.jk_builds} J/K builds')

    # -- one lane --------------------------------------------------------------------
    def _build(self, lane, reqs):
        """Queue the batched J/K builds of reqs on the lane's stream."""
        groups = collections.defaultdict(list)
        for r in reqs:
            prec = r.scf.next_precision() if r.kind == 'scf' else 'fp64'
            groups[(prec, r.opt.default_nbins())].append(r)
        builds = []
        for (prec, nbins), rs in groups.items():
            dms = [r.scf.next_density(prec) if r.kind == 'scf' else r.dm for r in rs]
            bj = BatchJK(None, vhfopts=[r.opt for r in rs], precision=prec, nbins=nbins,
                         fp32_cut=self.fp32_cut, stream=lane.stream)
            builds.append((prec, rs, bj, bj.build(dms)))
        return builds

    def _advance(self, builds):
        """SCF step of every request of finished builds; returns the
        requests that are done."""
        finished = []
        for prec, rs, _, out in builds:
            for r, (j, k) in zip(rs, out):
                r.jk_builds += 1
                if r.kind == 'jk':
                    r.result = {'vj': j.cpu().numpy(), 'vk': k.cpu().numpy()}
                    finished.append(r)
                elif r.scf.absorb(j, k, prec) or r.scf.cycles >= r.scf.max_cycle:
                    finished.append(r)
        return finished

    def _lane_loop(self, li):
        lane = self.lanes[li]
        sec = self.stats['seconds']
        try:
            with hf.on_stream(lane.stream):
                while True:
                    with self._cv:
                        while not lane.reqs and not self._closed:
                            self._cv.wait()
                        if not lane.reqs:
                            return
                        reqs = list(lane.reqs)
                    t0 = time.perf_counter()
                    builds = self._build(lane, reqs)
                    t1 = time.perf_counter()
                    lane.stream.synchronize()
                    t2 = time.perf_counter()
                    finished = self._advance(builds)
                    for r in finished:
                        self._retire(r)
                    t3 = time.perf_counter()
                    with self._cv:
                        sec['launch'] += t1 - t0
                        sec['wait'] += t2 - t1
                        sec['scf'] += t3 - t2
                        self.stats['iterations'] += 1
                        self.stats['lane_builds'][li] += 1
                        self.stats['batched_builds'] += len(builds)
                        self.stats['molecule_builds'] += len(reqs)
                        self.stats['batch_sizes'].extend(len(b[1]) for b in builds)
                        for r in finished:
                            lane.reqs.remove(r)
                            self.active.remove(r)
                            self.in_use -= r.footprint
                            self.completed.append(r)
                            self._done.append(r)
                        self._cv.notify_all()
        except BaseException as exc:   # surfaced by step()
            with self._cv:
                self._error = exc
                self._cv.notify_all()

    # -- the caller's loop ---------------------------------------------------------
    def step(self, timeout=0.05):
        """Admit what fits, then wait until a request finishes (at most
        ``timeout`` seconds while requests are in flight).  Returns the
        requests finished since the previous call."""
        t0 = time.perf_counter()
        self._admit()
        self.stats['seconds']['admit'] += time.perf_counter() - t0
        with self._cv:
            if not self._done and self.active and self._error is None:
                self._cv.wait(timeout)
            if self._error is not None:
                raise RuntimeError('a serving lane failed') from self._error
            done, self._done = self._done, []
        return done

    def idle(self):
        with self._cv:
            return not self.queue and not self.active

    def drain(self):
        while not self.idle():
            self.step()
        return self.completed

    def close(self):
        """Finish the requests in flight and stop the lane threads."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for lane in self.lanes:
            if lane.thread is not None:
                lane.thread.join()
                lane.thread = None
        with self._cv:
            if self._error is not None:
                raise RuntimeError('a serving lane failed') from self._error


def poisson_workload(qps, duration, seed, lo=5, hi=120, basis='def2-svp'):
    """[(arrival offset s, nwater, mol seed)]: Poisson arrivals at rate qps
    over duration seconds, water counts log-uniform in [lo, hi]."""
    rng = numpy.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / qps)
        if t >= duration:
            return out
        n = int(round(numpy.exp(rng.uniform(numpy.log(lo), numpy.log(hi)))))
        out.append((t, n, int(rng.integers(1 << 30))))


def run_load(server, arrivals, make_mol, log=None):
    """Feed arrivals (offset, *args) into server in real time and run it until
    every request is done; returns (requests, wall seconds)."""
    t0 = time.perf_counter()
    pending = collections.deque(sorted(arrivals))
    reqs = []
    while pending or not server.idle():
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            off, *args = pending.popleft()
            reqs.append(server.submit(make_mol(*args), arrival=t0 + off))
            if log:
                log(f'arrival of request {reqs[-1].rid} at {off:.1f} s '
                    f'({reqs[-1].mol.nao_nr()} AOs), queue {len(server.queue)}')
        if server.idle():
            if pending:
                time.sleep(max(0.0, min(pending[0][0] - (time.perf_counter() - t0), 0.05)))
            continue
        server.step()
    return reqs, time.perf_counter() - t0
