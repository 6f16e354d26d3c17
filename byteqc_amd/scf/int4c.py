"""Four-index ERIs on the GPU in the plan's decontracted Cartesian basis.

Mirrors cuobc/lib/int4c.py: ``get_int4c_slice`` fills the block of two pair
types through ``GINTfill_int2e`` (the same entry point and arguments as
cuobc/lib/int4c.py:115-127), ``get_int4c`` assembles the full tensor.  The
reference's ``get_int4c`` passes a keyword its plan builder does not accept
(int4c.py:49 vs hf.py:288) and overwrites blocks of diagonal pair types with
their transposes; here the plan keeps both orders of the shell pairs of a
group (``diag_block_with_triu=True``, no diagonal factor) and every block is
written once per index order, so the tensor is complete.
"""
from __future__ import annotations

import numpy
import torch

from byteqc_amd.lib import load_library
from byteqc_amd.scf import hf


def build_plan(mol, cutoff=1e-13):
    """Plan for ERI fills: full pair lists of diagonal groups, factor 1."""
    return hf._VHFOpt(mol, 'int2e').build(cutoff=cutoff, diag_block_with_triu=True)


def get_int4c_slice(cp_ij_id, cp_kl_id, vhfopt, nbins=1, out=None, stream=None):
    """(si, sj, sk, sl, block): ERIs of the pairs of pair types cp_ij_id and
    cp_kl_id (device tensor of shape (ni, nj, nk, nl) over the AO ranges of
    their shell groups), screened by the Schwarz bins as in the reference."""
    gi, gj = (int(x) for x in vhfopt.cp_grp[cp_ij_id])
    gk, gl = (int(x) for x in vhfopt.cp_grp[cp_kl_id])
    loc = vhfopt.group_ao_loc()
    i0, i1, j0, j1 = loc[gi], loc[gi + 1], loc[gj], loc[gj + 1]
    k0, k1, l0, l1 = loc[gk], loc[gk + 1], loc[gl], loc[gl + 1]
    ni, nj, nk, nl = i1 - i0, j1 - j0, k1 - k0, l1 - l0
    if out is None:
        out = torch.zeros((ni, nj, nk, nl), dtype=torch.float64, device=hf._device())
    ao_offsets = numpy.array([i0, j0, k0, l0], dtype=numpy.int32)
    strides = numpy.array([nj * nk * nl, nk * nl, nl, 1], dtype=numpy.int32)
    bins_ij = hf._make_s_index_offsets(vhfopt.log_qs[cp_ij_id], nbins, vhfopt.direct_scf_tol)
    bins_kl = hf._make_s_index_offsets(vhfopt.log_qs[cp_kl_id], nbins, vhfopt.direct_scf_tol)
    nao = int(vhfopt.mol.ao_loc_nr()[-1])
    with hf.on_stream(stream):
        err = load_library().GINTfill_int2e(vhfopt.bpcache, hf._ptr(out), nao, hf._ptr(strides),
                                            hf._ptr(ao_offsets), hf._ptr(bins_ij),
                                            hf._ptr(bins_kl), nbins, cp_ij_id, cp_kl_id,
                                            hf._stream_ptr())
    if err:
        raise RuntimeError(f'GINTfill_int2e failed ({err}) for pair types {cp_ij_id}, {cp_kl_id}')
    return slice(i0, i1), slice(j0, j1), slice(k0, k1), slice(l0, l1), out


def get_int4c(mol, vhfopt=None, nbins=1, cutoff=1e-13):
    """Full ERI tensor (nao_c^4, device) in the decontracted Cartesian basis
    of the plan (``vhfopt.mol``); returns (eri, vhfopt)."""
    if vhfopt is None:
        vhfopt = build_plan(mol, cutoff)
    n = int(vhfopt.mol.ao_loc_nr()[-1])
    eri = torch.zeros((n, n, n, n), dtype=torch.float64, device=hf._device())
    ncp = len(vhfopt.log_qs)
    for a in range(ncp):
        gi, gj = vhfopt.cp_grp[a]
        for b in range(ncp):
            gk, gl = vhfopt.cp_grp[b]
            si, sj, sk, sl, blk = get_int4c_slice(a, b, vhfopt, nbins)
            eri[si, sj, sk, sl] = blk
            if gi != gj:
                eri[sj, si, sk, sl] = blk.permute(1, 0, 2, 3)
            if gk != gl:
                eri[si, sj, sl, sk] = blk.permute(0, 1, 3, 2)
            if gi != gj and gk != gl:
                eri[sj, si, sl, sk] = blk.permute(1, 0, 3, 2)
    return eri, vhfopt
