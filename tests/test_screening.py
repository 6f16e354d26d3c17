"""Pins of the oracle's screening: Schwarz bounds, the task list with its
density prescreen, and the Schwarz-bin pairing of each task.

The oracle restates cuobc/scf/hf.py:111-145 (tasks and density bounds),
hf.py:567-578 (bins) and cuobc/lib/gvhf/nr_jk_driver.cu:252-275 (kl bin b
meets ij bins [0, nbins - b)).  These tests check it against what the
mathematics fixes, not against a second copy of the same formulas:

* q_cond(I, J) = sqrt(max |(ij|ij)|) over the components of the shell pair,
  taken from the Obara-Saika-pinned integrals and from an independent
  mpmath Obara-Saika evaluation;
* the Cauchy-Schwarz inequality |(ij|kl)| <= q_IJ q_KL over a full tensor
  (fails for a bound without the square root);
* exact safety of the kept quartet set, counted by brute force over every
  unique pair quartet: every quartet whose bound q'_P q'_Q exceeds the
  task's cutoff is kept, and no kept quartet lies below the cutoff raised
  to (nbins+1)/nbins, where q' = min(q, 1) (the reference clips log q at 0,
  hf.py:383-385);
* the kept set's size equals the number of quartets the oracle evaluates;
* screened J/K converge to the unscreened J/K (every unique quartet: pair
  cutoff 1e-300, one bin, no density prescreen) as the cutoff goes to 0.
"""
import math

import numpy
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from byteqc_amd.gto import M, gto_norm
from byteqc_amd.systems import water_mole, random_density
from oracle import oracle as orc
from tests import os_eri


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def _fac_sp(l):
    return {0: 0.282094791773878143, 1: 0.488602511902919921}.get(l, 1.)


def test_q_cond_is_sqrt_of_largest_diagonal_integral():
    mol = M(atom='O 0 0 0.58589194; H 0 0.75668992 0; H 0 -0.75668992 0', basis='cc-pvdz')
    bas, env, _ = orc.seg_contract(mol)
    atm = numpy.asarray(mol._atm, dtype=numpy.int32)
    q = orc.q_cond(atm, bas, env)
    nb = len(bas)
    assert numpy.allclose(q, q.T, rtol=0, atol=0)
    for i in range(nb):
        for j in range(nb):
            blk = orc.eri_shells(i, j, i, j, atm, bas, env)
            ni, nj = blk.shape[0], blk.shape[1]
            diag = numpy.array([[blk[a, b, a, b] for b in range(nj)] for a in range(ni)])
            ref = math.sqrt(numpy.abs(diag).max())
            assert q[i, j] == pytest.approx(ref, rel=1e-14), (i, j)


@pytest.mark.parametrize('la,lb', [(0, 0), (1, 0), (2, 1), (3, 2)])
def test_q_cond_matches_obara_saika(la, lb):
    """Primitive shell pairs: q from the oracle equals sqrt(max (ab|ab))
    evaluated by the independent mpmath recursion."""
    rng = numpy.random.default_rng(10 * la + lb)
    a, b = rng.uniform(0.3, 2.5, 2)
    ra, rb = rng.uniform(-1.0, 1.0, (2, 3))
    atoms = [('H', tuple(ra)), ('H', tuple(rb))]
    mol = M(atom=atoms, basis='cc-pvdz', unit='Bohr')
    env = list(mol._env[:28])
    bas = []
    for ia, (l, e) in enumerate(((la, a), (lb, b))):
        pe = len(env)
        env += [e, float(gto_norm(l, e))]
        bas.append([ia, l, 1, 1, 0, pe, pe + 1, 0])
    bas = numpy.asarray(bas, dtype=numpy.int32)
    env = numpy.asarray(env)
    q = orc.q_cond(numpy.asarray(mol._atm, dtype=numpy.int32), bas, env)
    ref = os_eri.primitive_eri(la, lb, la, lb, a, b, a, b, ra, rb, ra, rb)
    scale = (_fac_sp(la) * float(gto_norm(la, a)) * _fac_sp(lb) * float(gto_norm(lb, b))) ** 2
    ni, nj = _ncart(la), _ncart(lb)
    best = max(abs(float(ref[x][y][x][y])) for x in range(ni) for y in range(nj))
    assert q[0, 1] == pytest.approx(math.sqrt(best * scale), rel=1e-13)


def test_cauchy_schwarz_bound_over_full_tensor():
    mol = water_mole(1, 5.0, 3, 'def2-svp')
    bas, env, _ = orc.seg_contract(mol)
    atm = numpy.asarray(mol._atm, dtype=numpy.int32)
    q = orc.q_cond(atm, bas, env)
    eri = orc.eri_tensor(atm, bas, env)
    loc = orc.cart_ao_loc(bas)
    shell_of = numpy.repeat(numpy.arange(len(bas)), numpy.diff(loc))
    qa = q[shell_of[:, None], shell_of[None, :]]
    bound = qa[:, :, None, None] * qa[None, None, :, :]
    assert numpy.all(numpy.abs(eri) <= bound * (1 + 1e-12) + 1e-300)
    # the bound is attained on the diagonal of some pair (it is sharp)
    diag = numpy.einsum('ijij->ij', eri)
    assert numpy.isclose(numpy.sqrt(numpy.abs(diag).max()), q.max(), rtol=1e-14)


def _pair_bounds(plan):
    """Clipped bound q' = min(q, 1) of every plan pair, per pair type."""
    return [numpy.exp(lq) for lq in plan.log_qs]


def _kept_matrix(bins_ij, bins_kl, nij, nkl, nbins, diag):
    """Boolean (ij, kl) matrix of the quartets a task keeps: kl of bin b meets
    ij of bins [0, nbins - b); in a diagonal task only ij >= kl."""
    kept = numpy.zeros((nij, nkl), dtype=bool)
    for b in range(nbins):
        k0, k1 = bins_kl[b], bins_kl[b + 1]
        kept[:bins_ij[nbins - b], k0:k1] = True
    if diag:
        kept &= numpy.tri(nij, nkl, dtype=bool)
    return kept


def _check_task_set(mol, dm, tol, nbins, prescreen=True):
    plan = orc.build_plan(mol, cutoff=tol)
    dmc = plan.coeff @ dm @ plan.coeff.T
    loc = plan.ao_loc[plan.l_ctr_offsets]
    cond = orc._condense_absmax(dmc, loc)
    task_cp, task_bins, subs = orc.task_list(plan, cond, nbins, prescreen)
    qp = _pair_bounds(plan)
    ncp = len(plan.log_qs)
    cp_i, cp_j = numpy.tril_indices(len(plan.uniq_l_ctr))
    tasks = {(int(a), int(b)): t for t, (a, b) in enumerate(task_cp)}
    kept_total = 0
    for a in range(ncp):
        for b in range(a + 1):
            na, nb = len(qp[a]), len(qp[b])
            if na == 0 or nb == 0:
                assert (a, b) not in tasks
                continue
            i, j, k, l = cp_i[a], cp_j[a], cp_i[b], cp_j[b]
            sub = max(cond[i, j], cond[k, l], cond[i, k], cond[j, k], cond[i, l], cond[j, l])
            bound = numpy.outer(qp[a], qp[b]) * sub
            if a == b:
                bound = numpy.where(numpy.tri(na, nb, dtype=bool), bound, 0.)
            if (a, b) not in tasks:
                # a task is skipped only by the density prescreen
                assert prescreen and sub < tol
                assert bound.max() < tol
                continue
            t = tasks[(a, b)]
            assert subs[t] == sub
            bi = task_bins[t, :nbins + 1]
            bk = task_bins[t, nbins + 1:]
            kept = _kept_matrix(bi, bk, na, nb, nbins, a == b)
            cut = min(tol / sub, .1) if sub > 0 else .1
            # exact safety: nothing above the task's cutoff is dropped
            assert numpy.all(kept[bound / sub > cut]), (a, b)
            # and the bins keep nothing far below it
            if kept.any():
                assert (bound / sub)[kept].min() > cut ** ((nbins + 1) / nbins) * (1 - 1e-12)
            kept_total += int(kept.sum())
    _, _, nq = orc.get_jk(mol, dm, plan=plan, nbins=nbins, prescreen=prescreen,
                          return_count=True)
    assert nq == kept_total
    return kept_total


@settings(max_examples=12, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(nwater=st.integers(1, 2), seed=st.integers(0, 10 ** 6),
       log_scale=st.floats(-9.0, 0.0), tol_exp=st.sampled_from([-13, -11, -9, -7, -5]),
       nbins=st.integers(1, 12), basis=st.sampled_from(['cc-pvdz', 'def2-svp']))
def test_kept_quartets_are_exactly_safe(nwater, seed, log_scale, tol_exp, nbins, basis):
    mol = water_mole(nwater, 3.5, seed, basis)
    dm = random_density(mol.nao, 5 * nwater, seed) * 10. ** log_scale
    _check_task_set(mol, dm, 10. ** tol_exp, nbins)


def test_prescreen_off_keeps_every_task():
    mol = water_mole(2, 4.0, 17, 'cc-pvdz')
    dm = random_density(mol.nao, 10, 17)
    for scale in (1e-12, 1e-13, 1e-16):
        n_on = _check_task_set(mol, dm * scale, 1e-13, 10, prescreen=True)
        n_off = _check_task_set(mol, dm * scale, 1e-13, 10, prescreen=False)
        assert n_on <= n_off
    assert n_on == 0 and n_off > 0


def test_screened_jk_converges_to_unscreened():
    mol = water_mole(3, 5.0, 23, 'cc-pvdz')
    dm = random_density(mol.nao, 15, 23)
    full = orc.build_plan(mol, cutoff=1e-300)
    j0, k0, n0 = orc.get_jk(mol, dm, plan=full, nbins=1, prescreen=False, return_count=True)
    errs, counts = [], []
    for tol in (1e-5, 1e-7, 1e-9, 1e-11, 1e-13):
        plan = orc.build_plan(mol, cutoff=tol)
        j, k, n = orc.get_jk(mol, dm, plan=plan, return_count=True)
        errs.append(max(numpy.abs(j - j0).max(), numpy.abs(k - k0).max()))
        counts.append(n)
    print('screening errors', errs, 'quartets', counts, 'of', n0)
    assert all(c <= n0 for c in counts) and counts == sorted(counts)
    assert errs[0] > 0
    for e1, e2 in zip(errs, errs[1:]):
        assert e2 <= e1
    assert errs[-1] < 1e-10
