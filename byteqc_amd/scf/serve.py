"""Request loop for J/K and SCF requests with continuous batching.

The reference has no serving layer: its callers run one PySCF SCF at a time
and each ``get_jk`` call launches per task from Python
(cuobc/scf/hf.py:53-227, 253-274).  Here many requests share the GPU:

* A request is either one J/K build for a given density ("jk") or a full
  RHF run to convergence ("scf": core-Hamiltonian guess, DIIS, mixed-
  precision J/K until |dE| < switch_de, then FP64).
* Admission by HBM footprint: a request leaves the FIFO queue only when its
  estimated device footprint (plan, pair data, SCF matrices, DIIS history,
  J/K buffers) fits in the budget beside the requests in flight; the oldest
  request goes first and is never overtaken, so large molecules do not starve.
* Continuous batching: every iteration of the loop builds J/K for *all*
  requests in flight in one batched launch set (``BQCbuild_jk_batch``: one
  launch per integral class and kernel kind covering every molecule), then
  advances each SCF by one step.  Requests that arrive while others run join
  at the next iteration; finished ones leave and free their memory at once.
* Per-request state (plan, pair lists, pair data, S, H, X, D, DIIS history)
  stays on the device between iterations.

Requests in one batched build share the number of Schwarz bins and the
precision; the loop groups requests by both, so a request's quartet set and
its result do not depend on which other requests are in flight.
"""
from __future__ import annotations

import collections
import itertools
import time

import numpy
import torch

from byteqc_amd.scf import hf
from byteqc_amd.scf.batch import BatchJK
from byteqc_amd.scf.rhf import RHF

# device matrices of one SCF request besides the J/K build buffers: S, H, X,
# D, F, MO coefficients, orbital-basis work matrices, DIIS history (F and
# error per slot)
_SCF_MATS = 10
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


class Server:
    """Continuous-batching request loop on one GPU."""

    def __init__(self, hbm_budget=None, max_active=64, precision='mixed_then_fp64',
                 conv_tol=1e-8, switch_de=1e-5, max_cycle=80, fp32_cut=1e-7, log=None):
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_active = max_active
        self.scf_defaults = dict(precision=precision, conv_tol=conv_tol, switch_de=switch_de,
                                 max_cycle=max_cycle, fp32_cut=fp32_cut)
        self.fp32_cut = fp32_cut
        self.queue = collections.deque()
        self.active = []
        self.completed = []
        self.in_use = 0
        self._ids = itertools.count()
        self.log = log
        self.stats = {'iterations': 0, 'batched_builds': 0, 'molecule_builds': 0,
                      'max_active': 0, 'max_in_use': 0, 'batch_sizes': []}

    # -- requests ----------------------------------------------------------------
    def submit(self, mol, kind='scf', dm=None, arrival=None, **scf_kwargs):
        r = Request(next(self._ids), mol, kind, dm, arrival, **scf_kwargs)
        r.footprint = estimate_footprint(mol)
        if r.footprint > self.hbm_budget:
            r.error = (f'request needs ~{r.footprint / 2**30:.2f} GiB of HBM, above the '
                       f'budget of {self.hbm_budget / 2**30:.2f} GiB')
            r.finished = time.perf_counter()
            self.completed.append(r)
            return r
        self.queue.append(r)
        return r

    def _admit(self):
        while self.queue and len(self.active) < self.max_active:
            r = self.queue[0]
            if self.active and self.in_use + r.footprint > self.hbm_budget:
                break
            self.queue.popleft()
            r.admitted = time.perf_counter()
            r.opt = hf._VHFOpt(r.mol).build()
            if r.kind == 'scf':
                kw = dict(self.scf_defaults)
                kw.update(r.scf_kwargs)
                r.scf = RHF(r.mol, vhfopt=r.opt, **kw).build()
            else:
                r.dm = torch.as_tensor(r.dm, dtype=torch.float64, device=hf._device())
            self.in_use += r.footprint
            self.active.append(r)
        self.stats['max_active'] = max(self.stats['max_active'], len(self.active))
        self.stats['max_in_use'] = max(self.stats['max_in_use'], self.in_use)

    def _retire(self, r):
        r.finished = time.perf_counter()
        self.in_use -= r.footprint
        if r.opt is not None:
            r.opt.clear()
        r.opt = None
        if r.scf is not None:
            s = r.scf
            r.result = {'e_tot': s.e_tot, 'converged': s.converged, 'cycles': s.cycles,
                        'jk_builds': dict(s.jk_builds), 'dm': s.dm.cpu().numpy(),
                        'mo_energy': s.mo_energy.cpu().numpy()}
            r.scf = None
        self.completed.append(r)
        if self.log:
            self.log(f'request {r.rid} ({r.kind}, {r.mol.nao_nr()} AOs) done in '
                     f'{r.latency:.2f} s, {r.jk_builds} J/K builds')

    # -- one iteration -------------------------------------------------------------
    def step(self):
        """Admit, build J/K for every request in flight, advance each one.
        Returns the requests finished in this iteration."""
        self._admit()
        if not self.active:
            return []
        groups = collections.defaultdict(list)
        for r in self.active:
            prec = r.scf.next_precision() if r.kind == 'scf' else 'fp64'
            groups[(prec, r.opt.default_nbins())].append(r)
        finished = []
        for (prec, nbins), rs in groups.items():
            dms = [r.scf.dm if r.kind == 'scf' else r.dm for r in rs]
            out = BatchJK(None, vhfopts=[r.opt for r in rs], precision=prec, nbins=nbins,
                          fp32_cut=self.fp32_cut).build(dms)
            self.stats['batched_builds'] += 1
            self.stats['molecule_builds'] += len(rs)
            self.stats['batch_sizes'].append(len(rs))
            for r, (j, k) in zip(rs, out):
                r.jk_builds += 1
                if r.kind == 'jk':
                    r.result = {'vj': j.cpu().numpy(), 'vk': k.cpu().numpy()}
                    finished.append(r)
                elif r.scf.absorb(j, k, prec) or r.scf.cycles >= r.scf.max_cycle:
                    finished.append(r)
        self.stats['iterations'] += 1
        for r in finished:
            self.active.remove(r)
            self._retire(r)
        return finished

    def idle(self):
        return not self.queue and not self.active

    def drain(self):
        while not self.idle():
            self.step()
        return self.completed


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
