"""Molecule and basis container with PySCF's ``_atm/_bas/_env`` layout.

``get_jk`` in the reference reads a PySCF ``gto.Mole`` (cuobc/scf/hf.py:53,
280-324, 481-564): the integer tables ``_atm`` and ``_bas`` and the float
table ``_env``, with the slot meanings of cuobc/lib/gint/gint.h:28-46, plus
``ao_loc_nr``, ``cart2sph_coeff`` and ``aoslice_by_atom``.  PySCF is a
third-party dependency (pyscf 2.5.0) that this package does not ship, so this
module provides a Mole with the same attributes and the same numerical
conventions:

* primitive radial normalisation ``gto_norm(l, a) = 1/sqrt(int r^(2l+2)
  exp(-2 a r^2) dr)`` and renormalisation of every contracted radial function
  to unit self-overlap;
* general contractions kept as one shell with ``nctr`` columns, coefficients
  stored column by column (``env[ptr_coeff + ictr*nprim + iprim]``);
* Cartesian component order xx, xy, xz, yy, yz, zz (``lx`` from l down to 0,
  then ``ly`` from l-lx down to 0);
* real spherical functions in the order m = -l..l, except p which stays
  (x, y, z); s and p carry no angular factor in the Cartesian-to-spherical
  map because the integral code applies 1/(2 sqrt(pi)) and sqrt(3/(4 pi)) to
  them (cuobc/lib/gint/pair_data.c:27-28).

A real ``pyscf.gto.Mole`` exposes the same attributes and can be passed to
``get_jk`` instead of this class.
"""
from __future__ import annotations

import math
import os
from fractions import Fraction
from functools import lru_cache

import numpy

BOHR = 0.52917721092  # Angstrom per Bohr, the constant PySCF uses

# slots of _atm / _bas (cuobc/lib/gint/gint.h:28-46)
CHARGE_OF, PTR_COORD, NUC_MOD_OF, PTR_ZETA, PTR_FRAC_CHARGE, RESERVE_ATMSLOT = range(6)
ATM_SLOTS = 6
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF, RESERVE_BASLOT = range(8)
BAS_SLOTS = 8
PTR_ENV_START = 20

ANGULAR = 'spdfghik'
CHARGES = {'H': 1, 'He': 2, 'Li': 3, 'Be': 4, 'B': 5, 'C': 6, 'N': 7, 'O': 8,
           'F': 9, 'Ne': 10}

_BASIS_DIR = os.path.join(os.path.dirname(__file__), 'basis')
_BASIS_ALIAS = {
    'ccpvdz': 'cc-pvdz', 'cc-pvdz': 'cc-pvdz',
    'ccpvtz': 'cc-pvtz', 'cc-pvtz': 'cc-pvtz',
    'def2svp': 'def2-svp', 'def2-svp': 'def2-svp',
}


# ---------------------------------------------------------------------------
# radial normalisation
# ---------------------------------------------------------------------------
def gaussian_int(n, alpha):
    """int_0^inf x^n exp(-alpha x^2) dx."""
    n1 = (n + 1) * .5
    return math.gamma(n1) / (2. * alpha ** n1)


def gto_norm(l, expnt):
    """Normalisation of the radial part r^l exp(-a r^2) (int r^2 R^2 dr = 1)."""
    expnt = numpy.asarray(expnt, dtype=float)
    g = numpy.vectorize(lambda a: gaussian_int(2 * l + 2, 2 * a))(expnt)
    return 1. / numpy.sqrt(g)


def _normalize_contracted(l, es, cs):
    """Scale every column of cs so the contracted radial function has unit norm."""
    ee = es[:, None] + es[None, :]
    s = numpy.vectorize(lambda a: gaussian_int(2 * l + 2, a))(ee)
    nrm = numpy.einsum('pi,pq,qi->i', cs, s, cs)
    return cs / numpy.sqrt(nrm)


# ---------------------------------------------------------------------------
# basis-set files (NWChem format)
# ---------------------------------------------------------------------------
def parse_nwchem(text, symbol):
    """Return the shells of ``symbol`` as ``[[l, [e, c1, c2, ...], ...], ...]``.

    Blocks of one angular momentum that share an identical exponent list are
    merged into one general contraction, primitives whose coefficients are
    all zero are dropped and shells are ordered by l (the layout PySCF's
    basis parser produces).
    """
    blocks = []
    cur = None
    for raw in text.splitlines():
        line = raw.split('#')[0].strip()
        if not line or line.upper() == 'END':
            continue
        tok = line.split()
        if tok[0].isalpha():
            if tok[0] == symbol:
                cur = [ANGULAR.index(tok[1].lower()), []]
                blocks.append(cur)
            else:
                cur = None
            continue
        if cur is not None:
            cur[1].append([float(x.replace('D', 'E')) for x in tok])
    if not blocks:
        raise KeyError(f'no basis data for {symbol}')
    merged = {}
    for l, rows in blocks:
        ec = numpy.array(rows)
        es, cs = ec[:, 0], ec[:, 1:]
        lst = merged.setdefault(l, [])
        for item in lst:
            if item[0].shape == es.shape and numpy.array_equal(item[0], es):
                item[1] = numpy.hstack([item[1], cs])
                break
        else:
            lst.append([es, cs])
    shells = []
    for l in sorted(merged):
        for es, cs in merged[l]:
            keep = numpy.any(cs != 0, axis=1)
            es, cs = es[keep], cs[keep]
            order = numpy.argsort(-es, kind='stable')
            shells.append([l] + numpy.hstack([es[order, None], cs[order]]).tolist())
    return shells


@lru_cache(maxsize=None)
def _read_basis_file(name):
    with open(os.path.join(_BASIS_DIR, name + '.dat')) as f:
        return f.read()


def load_basis(name, symbol):
    key = name.lower().replace('_', '-')
    key = _BASIS_ALIAS.get(key, _BASIS_ALIAS.get(key.replace('-', ''), key))
    return parse_nwchem(_read_basis_file(key), symbol)


# ---------------------------------------------------------------------------
# Cartesian -> real spherical harmonics
# ---------------------------------------------------------------------------
def cart_components(l):
    """Cartesian exponents (lx, ly, lz) of a shell in storage order."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1)
            for ly in range(l - lx, -1, -1)]


@lru_cache(maxsize=None)
def cart2sph(l):
    """(ncart, 2l+1) map from Cartesian to real spherical functions of one shell.

    Column m+l holds the coefficients of the unit-normalised real solid
    harmonic r^l Y_lm (no Condon-Shortley phase) in the monomials
    x^a y^b z^c.  s and p are the identity (their angular factor lives in the
    integral prefactor, as in libcint).
    """
    if l < 2:
        return numpy.eye(2 * l + 1)
    comps = cart_components(l)
    index = {c: i for i, c in enumerate(comps)}
    out = numpy.zeros((len(comps), 2 * l + 1))
    for m in range(-l, l + 1):
        am = abs(m)
        vm = Fraction(0) if m >= 0 else Fraction(1, 2)
        nlm = math.sqrt(2 * math.factorial(l + am) * math.factorial(l - am)
                        / (2 if m == 0 else 1)) / (2 ** am * math.factorial(l))
        nlm *= math.sqrt((2 * l + 1) / (4 * math.pi))
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                v = vm
                vmax = int(Fraction(am, 2) - vm) + vm
                while v <= vmax:
                    two_v = int(2 * v)
                    sign = -1 if int(t + v - vm) % 2 else 1
                    c = (Fraction(sign) * Fraction(1, 4) ** t * math.comb(l, t)
                         * math.comb(l - t, am + t) * math.comb(t, u)
                         * math.comb(am, two_v))
                    a = 2 * t + am - int(2 * (u + v))
                    b = int(2 * (u + v))
                    cz = l - 2 * t - am
                    out[index[(a, b, cz)], m + l] += float(c) * nlm
                    v += 1
    return out


# ---------------------------------------------------------------------------
# Mole
# ---------------------------------------------------------------------------
def _parse_atoms(atom, unit):
    if isinstance(atom, str):
        items = []
        for line in atom.replace(';', '\n').splitlines():
            tok = line.split()
            if tok:
                items.append((tok[0], [float(x) for x in tok[1:4]]))
    else:
        items = [(a[0], list(a[1])) for a in atom]
    scale = 1. / BOHR if unit.lower().startswith('a') else 1.
    return [(sym, numpy.asarray(xyz, dtype=float) * scale) for sym, xyz in items]


class Mole:
    """Molecule with PySCF's integral tables (``_atm``, ``_bas``, ``_env``)."""

    def __init__(self, atom=None, basis='cc-pvdz', unit='Angstrom', cart=False,
                 verbose=0, charge=0, spin=0):
        self.atom = atom
        self.basis = basis
        self.unit = unit
        self.cart = cart
        self.verbose = verbose
        self.charge = charge
        self.spin = spin
        self._atm = self._bas = self._env = None

    def build(self):
        atoms = _parse_atoms(self.atom, self.unit)
        self._atom = atoms
        env = [0.] * PTR_ENV_START
        atm = []
        for sym, xyz in atoms:
            ptr = len(env)
            env.extend(xyz.tolist())
            env.append(0.)  # zeta of a point nucleus
            atm.append([CHARGES[sym], ptr, 1, ptr + 3, 0, 0])
        bas = []
        templates = {}
        for ia, (sym, _) in enumerate(atoms):
            bname = self.basis[sym] if isinstance(self.basis, dict) else self.basis
            key = (bname if isinstance(bname, str) else repr(bname), sym)
            if key not in templates:
                shells = []
                # a basis name, or PySCF's explicit form [[l, [e, c1, ...], ...], ...]
                src = load_basis(bname, sym) if isinstance(bname, str) else bname
                for shell in src:
                    l = shell[0]
                    ec = numpy.array(shell[1:])
                    es, cs = ec[:, 0], ec[:, 1:]
                    cs = cs * gto_norm(l, es)[:, None]
                    cs = _normalize_contracted(l, es, cs)
                    pe = len(env)
                    env.extend(es.tolist())
                    pc = len(env)
                    env.extend(cs.T.ravel().tolist())
                    shells.append([l, len(es), cs.shape[1], 0, pe, pc])
                templates[key] = shells
            for l, npr, nctr, kappa, pe, pc in templates[key]:
                bas.append([ia, l, npr, nctr, kappa, pe, pc, 0])
        self._atm = numpy.asarray(atm, dtype=numpy.int32).reshape(-1, ATM_SLOTS)
        self._bas = numpy.asarray(bas, dtype=numpy.int32).reshape(-1, BAS_SLOTS)
        self._env = numpy.asarray(env, dtype=numpy.float64)
        return self

    # -- sizes ---------------------------------------------------------------
    @property
    def natm(self):
        return len(self._atm)

    @property
    def nbas(self):
        return len(self._bas)

    @property
    def nelectron(self):
        return int(self._atm[:, CHARGE_OF].sum()) - self.charge

    @property
    def nao(self):
        return self.nao_nr()

    def nao_nr(self, cart=None):
        return int(self.ao_loc_nr(cart)[-1])

    def ao_loc_nr(self, cart=None):
        if cart is None:
            cart = self.cart
        l = self._bas[:, ANG_OF]
        nctr = self._bas[:, NCTR_OF]
        dims = ((l + 1) * (l + 2) // 2 if cart else 2 * l + 1) * nctr
        return numpy.append(0, numpy.cumsum(dims)).astype(numpy.int32)

    @property
    def ao_loc(self):
        return self.ao_loc_nr()

    # -- shell accessors -------------------------------------------------------
    def bas_angular(self, i):
        return int(self._bas[i, ANG_OF])

    def bas_nprim(self, i):
        return int(self._bas[i, NPRIM_OF])

    def bas_nctr(self, i):
        return int(self._bas[i, NCTR_OF])

    def bas_atom(self, i):
        return int(self._bas[i, ATOM_OF])

    def bas_exp(self, i):
        p = self._bas[i, PTR_EXP]
        return self._env[p:p + self._bas[i, NPRIM_OF]].copy()

    def bas_ctr_coeff(self, i):
        """(nprim, nctr) coefficients including the primitive normalisation."""
        b = self._bas[i]
        p, npr, nc = b[PTR_COEFF], b[NPRIM_OF], b[NCTR_OF]
        return self._env[p:p + npr * nc].reshape(nc, npr).T.copy()

    def atom_coords(self):
        ptr = self._atm[:, PTR_COORD]
        return numpy.stack([self._env[p:p + 3] for p in ptr]).reshape(-1, 3)

    def atom_charges(self):
        return self._atm[:, CHARGE_OF].copy()

    def atom_symbol(self, i):
        return self._atom[i][0]

    def aoslice_by_atom(self, ao_loc=None):
        """(natm, 4) table of [shell0, shell1, ao0, ao1] per atom."""
        if ao_loc is None:
            ao_loc = self.ao_loc_nr()
        out = numpy.zeros((self.natm, 4), dtype=int)
        atom_of = self._bas[:, ATOM_OF]
        for ia in range(self.natm):
            idx = numpy.where(atom_of == ia)[0]
            if idx.size:
                out[ia] = [idx[0], idx[-1] + 1, ao_loc[idx[0]], ao_loc[idx[-1] + 1]]
            elif ia > 0:
                out[ia] = [out[ia - 1, 1]] * 2 + [out[ia - 1, 3]] * 2
        return out

    def cart2sph_coeff(self):
        """(nao_cart, nao_sph) block-diagonal Cartesian-to-spherical map."""
        blocks = []
        for b in self._bas:
            c = cart2sph(int(b[ANG_OF]))
            blocks.extend([c] * int(b[NCTR_OF]))
        nc = sum(x.shape[0] for x in blocks)
        ns = sum(x.shape[1] for x in blocks)
        out = numpy.zeros((nc, ns))
        r = s = 0
        for x in blocks:
            out[r:r + x.shape[0], s:s + x.shape[1]] = x
            r += x.shape[0]
            s += x.shape[1]
        return out

    def energy_nuc(self):
        z = self.atom_charges().astype(float)
        r = self.atom_coords()
        e = 0.
        for i in range(self.natm):
            for j in range(i):
                e += z[i] * z[j] / numpy.linalg.norm(r[i] - r[j])
        return e

    def copy(self):
        import copy as _copy
        new = _copy.copy(self)
        for k in ('_atm', '_bas', '_env'):
            v = getattr(self, k)
            setattr(new, k, None if v is None else v.copy())
        return new


def M(**kwargs):
    """Build a Mole in one call, like ``pyscf.gto.M``."""
    return Mole(**kwargs).build()
