"""J/K builds of many independent molecules in one set of kernel launches.

The reference builds J/K one molecule and one pair-type task at a time from
Python (cuobc/scf/hf.py:111-166, one ``GINTbuild_jk`` call per task, with a
device synchronisation per task in multigpu.py:277).  Here the task lists of
all molecules go to ``BQCbuild_jk_batch``, which emits one launch per integral
class and kernel kind across the whole batch; a fixed batch can be captured
once into a HIP graph (``BQCgraph_capture``) and replayed.

Per build the host work is: C D C^T per molecule (device GEMMs), the block
density bounds of all molecules in one kernel, the task lists in C++
(``BQCplan_tasks``), then the batched launch, symmetrisation and the
back-transform C^T V C per molecule.
"""
from __future__ import annotations

import ctypes

import torch

from byteqc_amd.lib import load_library
from byteqc_amd.scf import hf

_PREC = {'fp64': 0, 'mixed': 1}


class BatchJK:
    """J and K for a list of molecules (each with its own plan).

    mols: the molecules (plans are built for them when vhfopts is None), or
    None when vhfopts (built plans, one per molecule) are given."""

    def __init__(self, mols, vhfopts=None, precision='fp64', fp32_cut=1e-7, nbins=None,
                 with_j=True, with_k=True, stream=None):
        if vhfopts is None:
            self.mols = list(mols)
            self.opts = [hf._VHFOpt(m).build() for m in self.mols]
        else:
            self.opts = list(vhfopts)
            self.mols = list(mols) if mols is not None else [o.mol for o in self.opts]
        if not self.opts:
            raise ValueError('empty batch')
        self.precision = precision
        self.fp32_cut = fp32_cut
        self.with_j, self.with_k = with_j, with_k
        self.stream = stream
        self.nbins = nbins or max(o.default_nbins() for o in self.opts)
        dev = hf._device()
        self.nao_c = [int(o.coeff.shape[0]) for o in self.opts]
        self.dmc = [torch.zeros((n, n), dtype=torch.float64, device=dev) for n in self.nao_c]
        self.vj = [torch.zeros_like(d) for d in self.dmc] if with_j else None
        self.vk = [torch.zeros_like(d) for d in self.dmc] if with_k else None
        self.locs = [o.group_ao_loc() for o in self.opts]
        self._graph = None
        self._keep = None

    # -- host side of one build ------------------------------------------------
    def _load(self, dms):
        for o, buf, d in zip(self.opts, self.dmc, dms):
            c = o.coeff_dev
            d = torch.as_tensor(d, dtype=torch.float64, device=buf.device)
            buf.copy_(c @ d @ c.T)

    def _plan(self):
        conds = hf.dm_cond_batch(self.dmc, self.locs, self.stream)
        per = [o.tasks(cd, self.nbins) for o, cd in zip(self.opts, conds)]
        return per

    def _args(self, per):
        n = len(self.mols)
        keep = []

        def arr(xs, ctype=ctypes.c_void_p):
            a = (ctype * n)(*xs)
            keep.append(a)
            return a
        bps = arr([ctypes.cast(o.bpcache, ctypes.c_void_p) for o in self.opts])
        vj = arr([v.data_ptr() for v in self.vj]) if self.vj is not None else None
        vk = arr([v.data_ptr() for v in self.vk]) if self.vk is not None else None
        dm = arr([d.data_ptr() for d in self.dmc])
        nao = arr(self.nao_c, ctypes.c_int)
        ntask = arr([len(p[0]) for p in per], ctypes.c_int)
        tcp = arr([p[0].ctypes.data for p in per])
        tbins = arr([p[1].ctypes.data for p in per])
        bra = arr([o.bra_order.ctypes.data for o in self.opts])
        lq = arr([o.log_q_prec.ctypes.data for o in self.opts])
        tdm = arr([p[2].ctypes.data for p in per])
        keep.append(per)
        return keep, (n, bps, vj, vk, dm, nao, 1, ntask, tcp, tbins, self.nbins, bra,
                      _PREC[self.precision], lq, tdm, self.fp32_cut)

    def _zero(self):
        for vs in (self.vj, self.vk):
            if vs is not None:
                for v in vs:
                    v.zero_()

    def _finish(self):
        lib = load_library()
        out = []
        for i, o in enumerate(self.opts):
            c = o.coeff_dev
            vj = self.vj[i] if self.vj is not None else None
            vk = self.vk[i] if self.vk is not None else None
            lib.BQCsymmetrize_jk(hf._ptr(vj), hf._ptr(vk), self.nao_c[i], 1,
                                 hf._stream_ptr(self.stream))
            out.append((c.T @ vj @ c if vj is not None else None,
                        c.T @ vk @ c if vk is not None else None))
        return out

    # -- public ------------------------------------------------------------------
    def build(self, dms):
        """One batched J/K build; returns [(J, K)] device tensors per molecule."""
        with hf.on_stream(self.stream):
            return self._build(dms)

    def _build(self, dms):
        self._load(dms)
        self._zero()
        keep, args = self._args(self._plan())
        err = load_library().BQCbuild_jk_batch(*args, hf._stream_ptr(self.stream))
        if err:
            raise RuntimeError(f'batched J/K build failed ({err})')
        del keep
        return self._finish()

    def capture(self, dms):
        """Record the kernels of the build for dms into a HIP graph."""
        with hf.on_stream(self.stream):
            return self._capture(dms)

    def _capture(self, dms):
        self.release()
        self._load(dms)
        keep, args = self._args(self._plan())
        torch.cuda.synchronize()
        g = load_library().BQCgraph_capture(*args, hf._stream_ptr(self.stream))
        if not g:
            raise RuntimeError('graph capture failed')
        self._graph = g
        self._keep = keep
        return self

    def replay(self):
        """Zero J/K, replay the captured kernels and return [(J, K)]."""
        with hf.on_stream(self.stream):
            self.replay_kernels()
            return self._finish()

    def replay_kernels(self):
        """Zero J/K and replay the captured kernels only (no transforms)."""
        if not self._graph:
            raise RuntimeError('no captured graph')
        with hf.on_stream(self.stream):
            self._zero()
            if load_library().BQCgraph_launch(self._graph, hf._stream_ptr(self.stream)):
                raise RuntimeError('graph launch failed')

    def release(self):
        if self._graph:
            load_library().BQCgraph_destroy(self._graph)
            self._graph = None
            self._keep = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def get_jk_batch(mols, dms, precision='fp64', **kwargs):
    """[(J, K)] (device tensors) for molecules mols with densities dms."""
    return BatchJK(mols, precision=precision, **kwargs).build(dms)
