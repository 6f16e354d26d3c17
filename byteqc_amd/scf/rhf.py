"""Restricted Hartree-Fock on one GPU around the J/K engine.

The reference runs PySCF's SCF loop with ``get_jk`` and ``_eigh`` replaced by
its GPU versions (cuobc/scf/hf.py:253-274).  This driver keeps every SCF
quantity on the device (S, H, X, D, F, DIIS history) and advances one
iteration per ``step()`` call, so a request loop can interleave many SCF runs.

Iteration: D = 2 C_occ C_occ^T; J, K = get_jk(D); F = H + J - K/2;
E = 1/2 Tr[D (H + F)] + E_nuc; Pulay DIIS on e = X^T (F D S - S D F) X;
F' = X^T F X diagonalised (X = S^-1/2, canonical, small eigenvalues dropped).
Converged when |dE| < conv_tol and max|e| < conv_tol_grad.  ``level_shift``
(Eh, default 0) raises the virtual orbitals of the Fock matrix that is
diagonalised, as PySCF's ``level_shift`` does.

``precision='mixed_then_fp64'`` builds J/K in mixed precision while
|dE| >= switch_de and in FP64 afterwards (and for the final energy).

Incremental builds (``incremental=True``, the default): J and K of the
density change dD = D - D_last are added to those of D_last, as PySCF's
direct-SCF ``get_veff`` does (vhf = vhf_last + J(dD) - K(dD)/2, pyscf
scf/hf.py) -- the SCF loop the reference runs, since its RHF
(cuobc/scf/hf.py:272-274) only replaces ``get_jk`` and leaves ``_eri``
unset.  The density prescreen then drops the quartets whose Schwarz bound
times max|dD| of their blocks is below the cutoff, so late iterations
evaluate a small part of the tensor.  The first build, and the first
build after a change of precision, use the whole density.
"""
from __future__ import annotations

import ctypes
import time

import numpy
import torch

from byteqc_amd.lib import load_library
from byteqc_amd.scf import hf


def int1e(opt, stream=None):
    """(S, T, V) in the spherical contracted AO basis of the plan's molecule,
    as device tensors."""
    lib = load_library()
    pm = opt.mol
    dev = hf._device()
    nao_c = int(pm.ao_loc_nr()[-1])
    atm = torch.as_tensor(pm._atm, device=dev)
    bas = torch.as_tensor(pm._bas, device=dev)
    env = torch.as_tensor(pm._env, device=dev)
    aol = opt._dev['ao_loc']
    mats = [torch.zeros((nao_c, nao_c), dtype=torch.float64, device=dev) for _ in range(3)]
    grp_l = numpy.ascontiguousarray(opt.uniq_l_ctr[:, 0], dtype=numpy.int32)
    err = lib.BQCint1e(*(hf._ptr(m) for m in mats), hf._ptr(atm), pm.natm, hf._ptr(bas),
                       hf._ptr(env), hf._ptr(aol), nao_c, hf._ptr(opt.l_ctr_offsets),
                       len(grp_l), hf._ptr(grp_l), hf._stream_ptr(stream))
    if err:
        raise RuntimeError('one-electron integral kernel failed')
    c = opt.coeff_dev
    return tuple(c.T @ m @ c for m in mats)


def _frac_occ(e, nelec, degen=1e-6):
    """Aufbau occupations (spin-summed, at most 2 per orbital) for orbital
    energies e; a degenerate set at the Fermi level shares its electrons
    equally, which keeps an atom's density spherical."""
    order = numpy.argsort(e)
    occ = numpy.zeros(len(e))
    left, i = float(nelec), 0
    while left > 1e-12 and i < len(e):
        j = i + 1
        while j < len(e) and e[order[j]] - e[order[i]] < degen:
            j += 1
        put = min(2.0 * (j - i), left)
        occ[order[i:j]] = put / (j - i)
        left -= put
        i = j
    return occ


_ATOM_DM = {}


def atomic_density(symbol, basis, max_cycle=60, conv_tol=1e-9):
    """Spin-restricted, spherically averaged density of a neutral atom in
    the given basis (spherical contracted AOs, host array): fractional-
    occupation Hartree-Fock with DIIS, J/K from the GPU engine.  Cached per
    (element, basis)."""
    from byteqc_amd.gto import M
    key = (symbol, basis if isinstance(basis, str) else repr(basis))
    if key in _ATOM_DM:
        return _ATOM_DM[key]
    atom = M(atom=f'{symbol} 0 0 0', basis=basis)
    opt = hf._VHFOpt(atom).build()
    s, t, v = (m.cpu().numpy() for m in int1e(opt))
    h = t + v
    e, u = numpy.linalg.eigh(s)
    x = u[:, e > 1e-9] / numpy.sqrt(e[e > 1e-9])
    nelec = atom.nelectron

    def density(f):
        eps, c = numpy.linalg.eigh(x.T @ f @ x)
        c = x @ c
        return (c * _frac_occ(eps, nelec)) @ c.T

    dm = density(h)
    fs, errs, e_prev = [], [], None
    for _ in range(max_cycle):
        j, k = hf.get_jk(atom, dm, vhfopt=opt)
        f = h + j - 0.5 * k
        e_tot = 0.5 * float(numpy.sum(dm * (h + f)))
        sdf = s @ dm @ f
        err = x.T @ (sdf.T - sdf) @ x
        if e_prev is not None and abs(e_tot - e_prev) < conv_tol and abs(err).max() < 1e-5:
            break
        e_prev = e_tot
        fs.append(f)
        errs.append(err.ravel())
        fs, errs = fs[-8:], errs[-8:]
        n = len(fs)
        if n > 1:
            b = -numpy.ones((n + 1, n + 1))
            b[n, n] = 0
            b[:n, :n] = numpy.array(errs) @ numpy.array(errs).T
            rhs = numpy.zeros(n + 1)
            rhs[n] = -1
            coef = numpy.linalg.lstsq(b, rhs, rcond=None)[0][:n]
            f = numpy.einsum('i,ijk->jk', coef, numpy.array(fs))
        dm = density(f)
    opt.clear()
    _ATOM_DM[key] = dm
    return dm


def sad_density(mol):
    """Superposition of atomic densities: block-diagonal over the atoms'
    AO ranges (host array)."""
    nao = mol.nao_nr()
    dm = numpy.zeros((nao, nao))
    for ia, (_, _, a0, a1) in enumerate(mol.aoslice_by_atom()):
        sym = mol.atom_symbol(ia)
        basis = mol.basis[sym] if isinstance(mol.basis, dict) else mol.basis
        dm[a0:a1, a0:a1] = atomic_density(sym, basis)
    return dm


class RHF:
    """init_guess: 'atom' (superposition of spherically averaged atomic
    densities, as PySCF's init_guess='atom') or 'hcore' (core Hamiltonian)."""

    def __init__(self, mol, conv_tol=1e-8, conv_tol_grad=None, max_cycle=80, diis_space=8,
                 precision='fp64', switch_de=1e-5, fp32_cut=1e-7, lindep=1e-9, vhfopt=None,
                 stream=None, init_guess='atom', incremental=True, level_shift=0.0):
        if init_guess not in ('atom', 'hcore'):
            raise ValueError(f'unknown initial guess {init_guess!r}')
        self.init_guess = init_guess
        self.incremental = incremental
        self.level_shift = level_shift
        self.mol = mol
        self.conv_tol = conv_tol
        self.conv_tol_grad = conv_tol_grad if conv_tol_grad is not None else conv_tol ** 0.5
        self.max_cycle = max_cycle
        self.diis_space = diis_space
        self.precision = precision
        self.switch_de = switch_de
        self.fp32_cut = fp32_cut
        self.lindep = lindep
        self.stream = stream
        self._opt = vhfopt
        self.converged = False
        self.e_tot = None
        self.cycles = 0
        self.history = []
        self.build_flops = []   # algorithmic FLOPs of each standalone step's build
        self.jk_builds = {'fp64': 0, 'mixed': 0}
        # J, K of the density of the last build, its precision and density;
        # whether the build in flight is of dD (set by next_density)
        self._jk_last = None
        self._jk_last_prec = None
        self._dm_last = None
        self._pending = None

    # -- set-up --------------------------------------------------------------
    def build(self):
        with hf.on_stream(self.stream):
            return self._build()

    def _build(self):
        t0 = time.perf_counter()
        if self._opt is None:
            self._opt = hf._VHFOpt(self.mol).build()
        s, t, v = int1e(self._opt, self.stream)
        self.S, self.H = s, t + v
        e, u = torch.linalg.eigh(s)
        keep = e > self.lindep
        self.X = u[:, keep] / torch.sqrt(e[keep])
        self.e_nuc = float(self.mol.energy_nuc())
        self.nocc = self.mol.nelectron // 2
        self._f_hist, self._e_hist = [], []
        self.mo_coeff = self._diag(self.H)
        if self.init_guess == 'atom':
            self.dm = torch.as_tensor(sad_density(self.mol), device=s.device)
        else:
            self.dm = self._make_dm(self.mo_coeff)
        self.e_prev = None
        self.build_time = time.perf_counter() - t0
        return self

    def _diag(self, f):
        fo = self.X.T @ f @ self.X
        self.mo_energy, c = torch.linalg.eigh(fo)
        return self.X @ c

    def _make_dm(self, c):
        co = c[:, :self.nocc]
        return 2.0 * co @ co.T

    def _jk_precision(self):
        if self.precision == 'fp64':
            return 'fp64'
        if self.precision == 'mixed':
            return 'mixed'
        # mixed_then_fp64
        if self.e_prev is not None and self.history and abs(self.history[-1]['dE']) < self.switch_de:
            return 'fp64'
        return 'mixed'

    # -- one iteration ---------------------------------------------------------
    def step(self):
        """One SCF iteration; returns True once converged."""
        with hf.on_stream(self.stream):
            return self._step()

    def _step(self):
        prec = self.next_precision()
        j, k = hf.get_jk(self.mol, self.next_density(prec), vhfopt=self._opt,
                         return_device=True, precision=prec, fp32_cut=self.fp32_cut,
                         stream=self.stream)
        self.build_flops.append(sum(hf.last_build_flops()))
        return self.absorb(j, k, prec)

    def next_precision(self):
        """Precision of the next J/K build ('fp64' or 'mixed')."""
        return self._jk_precision()

    def next_density(self, prec):
        """The density the next J/K build (at precision prec) must use: the
        change since the last build when building incrementally, else the
        current density.  absorb() must follow with that build's J, K."""
        incr = (self.incremental and self._jk_last is not None
                and prec == self._jk_last_prec)
        self._pending = incr
        return self.dm - self._dm_last if incr else self.dm

    def absorb(self, j, k, prec):
        """Advance one iteration with J, K built (at precision prec, by any
        caller, e.g. a batched build) from the density next_density()
        returned -- or from the current density if it was not called;
        returns True once converged."""
        self.jk_builds[prec] += 1
        if self._pending:
            j = self._jk_last[0] + j
            k = self._jk_last[1] + k
        self._pending = None
        self._jk_last = (j, k)
        self._jk_last_prec = prec
        self._dm_last = self.dm
        f = self.H + j - 0.5 * k
        e_elec = 0.5 * float(torch.sum(self.dm * (self.H + f)))
        e_tot = e_elec + self.e_nuc
        sdf = self.S @ self.dm @ f
        err = self.X.T @ (sdf.T - sdf) @ self.X
        err_max = float(err.abs().max())
        de = e_tot - self.e_prev if self.e_prev is not None else e_tot
        self.history.append({'e': e_tot, 'dE': de, 'err': err_max, 'precision': prec})
        self.cycles += 1
        self.e_tot = e_tot
        conv = (self.e_prev is not None and abs(de) < self.conv_tol
                and err_max < self.conv_tol_grad and prec == 'fp64') or \
               (self.precision == 'mixed' and self.e_prev is not None
                and abs(de) < self.conv_tol and err_max < self.conv_tol_grad)
        self.e_prev = e_tot
        if conv:
            self.converged = True
            return True
        f = self._diis(f, err)
        if self.level_shift:
            # PySCF's level_shift (scf/hf.py): raise the virtual space by the
            # shift, f + shift (S - S (D/2) S); F D S - S D F and so the
            # converged solution do not change
            sd = self.S @ self.dm
            f = f + self.level_shift * (self.S - 0.5 * sd @ self.S)
        self.mo_coeff = self._diag(f)
        self.dm = self._make_dm(self.mo_coeff)
        return False

    def _diis(self, f, err):
        self._f_hist.append(f)
        self._e_hist.append(err)
        if len(self._f_hist) > self.diis_space:
            self._f_hist.pop(0)
            self._e_hist.pop(0)
        n = len(self._f_hist)
        if n < 2:
            return f
        b = torch.empty((n + 1, n + 1), dtype=torch.float64, device=f.device)
        e = torch.stack(self._e_hist).reshape(n, -1)
        b[:n, :n] = e @ e.T
        b[n, :n] = b[:n, n] = -1.0
        b[n, n] = 0.0
        rhs = torch.zeros(n + 1, dtype=torch.float64, device=f.device)
        rhs[n] = -1.0
        c = torch.linalg.lstsq(b.cpu(), rhs.cpu().unsqueeze(1)).solution.squeeze(1).to(f.device)
        return torch.einsum('i,ijk->jk', c[:n], torch.stack(self._f_hist))

    def kernel(self):
        if not hasattr(self, 'S'):
            self.build()
        for _ in range(self.max_cycle):
            if self.step():
                break
        return self.e_tot


def scf_energy(mol, **kwargs):
    mf = RHF(mol, **kwargs)
    return mf.kernel(), mf
