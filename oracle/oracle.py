"""CPU FP64 oracle for ``get_jk`` -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as a checker.  The product
package ``byteqc_amd`` never imports it.

This is a restatement of the reference's host-side J/K algorithm in numpy,
driving the C restatement of its kernels in ``jk_oracle.c``:

* ``basis_seg_contraction``  -> cuobc/scf/hf.py:481-564
* ``VHFOpt.build``           -> cuobc/scf/hf.py:288-408 (shell sort by
  (l, nprim), Schwarz pair lists per group pair sorted by q descending,
  ``log q`` clipped at 0, diagonal blocks keep ish >= jsh)
* ``make_s_index_offsets``   -> cuobc/scf/hf.py:567-578
* ``get_jk``                 -> cuobc/scf/hf.py:53-227 (D to the decontracted
  Cartesian basis, density prescreen per task, bins, V -> J = 2(V+V^T),
  K = V+V^T, back-transform)

Parity status of each function is listed in the test suite; every function
here is pinned by at least one check in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = range(7)


def build():
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, '_build', 'liboracle.so')
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _LIB.oracle_rys_init()
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Boys / Rys
# ---------------------------------------------------------------------------
def rys(n, x):
    """Rys nodes (u = t^2/(1-t^2)) and weights from the oracle tables."""
    u = numpy.zeros(n)
    w = numpy.zeros(n)
    lib().oracle_rys(ctypes.c_int(n), ctypes.c_double(x), _ptr(u), _ptr(w))
    return u, w


def rys_direct(n, x):
    """Nodes s = t^2 and weights from the long-double Stieltjes procedure."""
    s = numpy.zeros(n)
    w = numpy.zeros(n)
    lib().oracle_rys_direct(ctypes.c_int(n), ctypes.c_double(x), _ptr(s), _ptr(w))
    return s, w


def boys(mmax, x):
    f = numpy.zeros(mmax + 1)
    lib().oracle_boys(ctypes.c_int(mmax), ctypes.c_double(x), _ptr(f))
    return f


# ---------------------------------------------------------------------------
# basis handling (cuobc/scf/hf.py:481-564)
# ---------------------------------------------------------------------------
def _gto_norm(l, es):
    from math import gamma
    es = numpy.asarray(es, dtype=float)
    n1 = l + 1.5
    return 1. / numpy.sqrt(gamma(n1) / (2. * (2 * es) ** n1))


def seg_contract(mol):
    """Decontract shells with nctr > 1 into primitive shells.

    Returns (bas, env, coeff) of the Cartesian primitive-shell basis and the
    (nao_c, nao0) map C with AO_sph = AO_c C, following hf.py:481-564
    (templates per atom keyed by its exponent pointers).
    """
    env = mol._env.copy()
    bas_out = []
    blocks = []
    templates = {}
    aoslices = mol.aoslice_by_atom()
    for ia, (ib0, ib1) in enumerate(aoslices[:, :2]):
        key = tuple(mol._bas[ib0:ib1, PTR_EXP])
        if key in templates:
            tb, tc = templates[key]
            tb = tb.copy()
            tb[:, ATOM_OF] = ia
        else:
            tb, tc = [], []
            for shell in mol._bas[ib0:ib1]:
                l = int(shell[ANG_OF])
                nf = (l + 1) * (l + 2) // 2
                nctr = int(shell[NCTR_OF])
                if nctr == 1:
                    tb.append(shell[None])
                    tc.append(numpy.eye(nf))
                    continue
                nprim = int(shell[NPRIM_OF])
                pe, pc = int(shell[PTR_EXP]), int(shell[PTR_COEFF])
                nrm = _gto_norm(l, env[pe:pe + nprim])
                c = env[pc:pc + nprim * nctr].reshape(nctr, nprim) / nrm
                # rows: primitive-shell Cartesian AOs (prim-major); cols:
                # contracted Cartesian AOs (ctr-major)
                blk = numpy.einsum('ip,ef->pfie', c, numpy.eye(nf)).reshape(
                    nprim * nf, nctr * nf)
                tc.append(blk)
                env[pc:pc + nprim] = nrm
                bs = numpy.repeat(shell[None], nprim, axis=0)
                bs[:, NPRIM_OF] = 1
                bs[:, NCTR_OF] = 1
                bs[:, PTR_EXP] = numpy.arange(pe, pe + nprim)
                bs[:, PTR_COEFF] = numpy.arange(pc, pc + nprim)
                tb.append(bs)
            tb = numpy.vstack(tb)
            templates[key] = (tb, tc)
        bas_out.append(tb)
        blocks.extend(tc)
    bas = numpy.asarray(numpy.vstack(bas_out), dtype=numpy.int32)
    rows = sum(b.shape[0] for b in blocks)
    cols = sum(b.shape[1] for b in blocks)
    cc = numpy.zeros((rows, cols))
    r = c = 0
    for b in blocks:
        cc[r:r + b.shape[0], c:c + b.shape[1]] = b
        r += b.shape[0]
        c += b.shape[1]
    if not mol.cart:
        cc = cc @ mol.cart2sph_coeff()
    return bas, env, cc


def cart_ao_loc(bas):
    l = bas[:, ANG_OF]
    return numpy.append(0, numpy.cumsum((l + 1) * (l + 2) // 2)).astype(numpy.int32)


def q_cond(atm, bas, env):
    nbas = len(bas)
    q = numpy.zeros((nbas, nbas))
    lib().oracle_q_cond(_ptr(q), _ptr(atm), ctypes.c_int(len(atm)), _ptr(bas),
                        ctypes.c_int(nbas), _ptr(env))
    return q


def eri_shells(ish, jsh, ksh, lsh, atm, bas, env):
    nf = [(int(bas[s, ANG_OF]) + 1) * (int(bas[s, ANG_OF]) + 2) // 2
          for s in (ish, jsh, ksh, lsh)]
    out = numpy.zeros(nf)
    lib().oracle_eri_shells(ctypes.c_int(ish), ctypes.c_int(jsh), ctypes.c_int(ksh),
                            ctypes.c_int(lsh), _ptr(atm), _ptr(bas), _ptr(env),
                            _ptr(out))
    return out


def eri_tensor(atm, bas, env):
    """Full Cartesian ERI tensor (nao_c^4) of a small primitive-shell basis."""
    ao_loc = cart_ao_loc(bas)
    n = ao_loc[-1]
    out = numpy.zeros((n, n, n, n))
    nb = len(bas)
    for i in range(nb):
        for j in range(nb):
            for k in range(nb):
                for l in range(nb):
                    out[ao_loc[i]:ao_loc[i + 1], ao_loc[j]:ao_loc[j + 1],
                        ao_loc[k]:ao_loc[k + 1], ao_loc[l]:ao_loc[l + 1]] = \
                        eri_shells(i, j, k, l, atm, bas, env)
    return out


# ---------------------------------------------------------------------------
# plan (cuobc/scf/hf.py:288-408, 567-578)
# ---------------------------------------------------------------------------
class Plan:
    pass


def build_plan(mol, cutoff=1e-13):
    bas, env, coeff = seg_contract(mol)
    atm = numpy.asarray(mol._atm, dtype=numpy.int32)
    key = bas[:, [ANG_OF, NPRIM_OF]]
    uniq, inv, counts = numpy.unique(key, axis=0, return_inverse=True,
                                     return_counts=True)
    inv = inv.ravel()
    order = numpy.argsort(inv, kind='stable')
    ao_loc = cart_ao_loc(bas)
    ao_idx = numpy.hstack([numpy.arange(ao_loc[i], ao_loc[i + 1]) for i in order])
    coeff = coeff[ao_idx]
    bas = numpy.ascontiguousarray(bas[order])
    q = q_cond(atm, bas, env)
    offs = numpy.append(0, numpy.cumsum(counts))
    log_qs, bra, ket, cp_l, cp_np = [], [], [], [], []
    for i in range(len(uniq)):
        p0, p1 = offs[i], offs[i + 1]
        for j in range(i + 1):
            q0, q1 = offs[j], offs[j + 1]
            sub = q[p0:p1, q0:q1].ravel()
            idx = sub.argsort()[::-1]
            qs = sub[idx]
            ii, jj = numpy.unravel_index(idx, (p1 - p0, q1 - q0))
            mask = qs > cutoff
            if i == j:
                mask &= ii >= jj
            bra.append(ii[mask] + p0)
            ket.append(jj[mask] + q0)
            lq = numpy.log(qs[mask])
            lq[lq > 0] = 0
            log_qs.append(lq)
            cp_l.append((int(uniq[i, 0]), int(uniq[j, 0])))
            cp_np.append(int(uniq[i, 1] * uniq[j, 1]))
    p = Plan()
    p.atm, p.bas, p.env, p.coeff = atm, bas, env, coeff
    p.ao_loc = cart_ao_loc(bas)
    p.uniq_l_ctr, p.l_ctr_offsets = uniq, offs
    p.log_qs = log_qs
    p.pair2bra = numpy.hstack(bra).astype(numpy.int32)
    p.pair2ket = numpy.hstack(ket).astype(numpy.int32)
    p.bas_pairs_locs = numpy.append(0, numpy.cumsum([len(x) for x in bra])).astype(
        numpy.int32)
    p.cp_l = cp_l
    p.cp_nprim = numpy.asarray(cp_np, dtype=numpy.int32)
    npp = p.cp_nprim * numpy.diff(p.bas_pairs_locs)
    p.prim_pairs_locs = numpy.append(0, numpy.cumsum(npp)).astype(numpy.int32)
    nprim_pairs = int(p.prim_pairs_locs[-1])
    p.aexyz = numpy.zeros(5 * max(nprim_pairs, 1))
    lib().oracle_pair_data(_ptr(p.aexyz), ctypes.c_int(nprim_pairs), _ptr(p.pair2bra),
                           _ptr(p.pair2ket), ctypes.c_int(len(p.pair2bra)),
                           ctypes.c_double(0.5), _ptr(atm), _ptr(bas), _ptr(env))
    p.nprim_pairs = nprim_pairs
    p.cutoff = cutoff
    p.q = q
    return p


def make_s_index_offsets(log_q, nbins=10, cutoff=1e-12):
    scale = nbins / numpy.log(min(cutoff, .1))
    s_index = numpy.floor(scale * log_q).astype(numpy.int32)
    bins = numpy.bincount(s_index, minlength=nbins)[:nbins]
    return numpy.append(0, numpy.cumsum(bins)).astype(numpy.int32)


def _condense_absmax(d, loc):
    n = len(loc) - 1
    out = numpy.zeros((n, n))
    for a in range(n):
        for b in range(n):
            blk = d[loc[a]:loc[a + 1], loc[b]:loc[b + 1]]
            out[a, b] = numpy.abs(blk).max() if blk.size else 0.
    return out


def task_list(plan, cond, nbins, prescreen=True):
    """Tasks (pair type a >= b), their bin offsets and density bounds, as the
    reference's loop in cuobc/scf/hf.py:111-145 (bounds: hf.py:119-126, bins:
    _make_s_index_offsets, hf.py:567-578)."""
    ncp = len(plan.log_qs)
    cp_i, cp_j = numpy.tril_indices(len(plan.uniq_l_ctr))
    tasks, bins, subs = [], [], []
    for t in range(ncp * (ncp + 1) // 2):
        a = int(numpy.floor(numpy.sqrt(2 * t)))
        if a * (a + 1) // 2 > t:
            a -= 1
        b = t - a * (a + 1) // 2
        lqa, lqb = plan.log_qs[a], plan.log_qs[b]
        if lqa.size == 0 or lqb.size == 0:
            continue
        i, j, k, l = cp_i[a], cp_j[a], cp_i[b], cp_j[b]
        sub = max(cond[i, j], cond[k, l], cond[i, k], cond[j, k], cond[i, l], cond[j, l])
        if prescreen and sub < plan.cutoff:
            continue
        with numpy.errstate(divide='ignore'):
            cut = plan.cutoff / sub
        tasks.append((a, b))
        subs.append(sub)
        bins.append(numpy.concatenate([make_s_index_offsets(lqa, nbins, cut),
                                       make_s_index_offsets(lqb, nbins, cut)]))
    task_cp = numpy.asarray(tasks, dtype=numpy.int32).reshape(-1, 2)
    task_bins = numpy.asarray(bins, dtype=numpy.int32).reshape(len(tasks), 2 * (nbins + 1))
    return task_cp, task_bins, numpy.asarray(subs)


def get_jk(mol, dm, plan=None, with_j=True, with_k=True, prescreen=True, nbins=None,
           return_count=False, kl_stride=1):
    """Oracle J and K for hermitian densities dm of shape (nao0,nao0) or (n,nao0,nao0)."""
    if plan is None:
        plan = build_plan(mol)
    coeff = plan.coeff
    nao, nao0 = coeff.shape
    dm0 = numpy.asarray(dm)
    dms = dm0.reshape(-1, nao0, nao0)
    n_dm = len(dms)
    dmc = numpy.ascontiguousarray(coeff @ dms @ coeff.T)
    if nbins is None:
        pairs_max = int(numpy.diff(plan.bas_pairs_locs).max())
        nbins = max(10, pairs_max // 100000)
    loc = plan.ao_loc[plan.l_ctr_offsets]
    cond = numpy.max([_condense_absmax(x, loc) for x in dmc], axis=0)
    task_cp, task_bins, _ = task_list(plan, cond, nbins, prescreen)
    tasks = task_cp
    vj = numpy.zeros((n_dm, nao, nao))
    vk = numpy.zeros((n_dm, nao, nao))
    nq = ctypes.c_longlong(0)
    err = lib().oracle_jk(
        ctypes.c_int(len(tasks)), _ptr(task_cp), _ptr(task_bins), ctypes.c_int(nbins),
        _ptr(plan.bas_pairs_locs), _ptr(plan.prim_pairs_locs), _ptr(plan.cp_nprim),
        _ptr(plan.pair2bra), _ptr(plan.pair2ket), ctypes.c_int(len(plan.pair2bra)),
        _ptr(plan.aexyz), ctypes.c_int(plan.nprim_pairs), _ptr(plan.atm), _ptr(plan.bas),
        _ptr(plan.env), _ptr(plan.ao_loc), ctypes.c_int(nao), ctypes.c_int(n_dm),
        _ptr(dmc), _ptr(vj) if with_j else None, _ptr(vk) if with_k else None,
        ctypes.byref(nq), ctypes.c_int(kl_stride))
    if err:
        raise RuntimeError('oracle_jk failed')
    out = []
    if with_j:
        vj = 2 * (vj + vj.transpose(0, 2, 1))
        out.append(coeff.T @ vj @ coeff)
    else:
        out.append(None)
    if with_k:
        vk = vk + vk.transpose(0, 2, 1)
        out.append(coeff.T @ vk @ coeff)
    else:
        out.append(None)
    if dm0.ndim == 2:
        out = [x[0] if x is not None else None for x in out]
    if return_count:
        return out[0], out[1], nq.value
    return out[0], out[1]


def jk_from_tensor(eri, dm):
    """Brute-force J_ij = sum_kl (ij|kl) D_kl, K_il = sum_jk (ij|kl) D_jk."""
    return (numpy.einsum('ijkl,kl->ij', eri, dm), numpy.einsum('ijkl,jk->il', eri, dm))
