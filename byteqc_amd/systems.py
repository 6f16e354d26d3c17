"""Synthetic water clusters and densities used by tests and benchmarks.

Water monomer: r(OH) = 0.9572 Angstrom, angle(HOH) = 104.52 degrees.  Each
monomer gets a seeded uniform random rotation and a uniform oxygen position in
a cube of side ``box``; a placement is rejected if any O-O distance would fall
below ``min_oo`` (2.7 Angstrom).  Densities are D = 2 C C^T with C the Q factor
of a seeded N(0,1) matrix of shape (nao, nocc), five occupied orbitals per
water.
"""
from __future__ import annotations

import math

import numpy

R_OH = 0.9572
THETA_HOH = 104.52
WATER_MASS_G = 18.01528 / 6.02214076e23   # grams per molecule
MIN_OO = 2.7


def _monomer():
    half = math.radians(THETA_HOH) / 2
    return numpy.array([[0., 0., 0.],
                        [R_OH * math.sin(half), 0., R_OH * math.cos(half)],
                        [-R_OH * math.sin(half), 0., R_OH * math.cos(half)]])


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= numpy.linalg.norm(q)
    w, x, y, z = q
    return numpy.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(nwater, box, seed, min_oo=MIN_OO, max_tries=2000000):
    """Atom list [(symbol, (x, y, z)), ...] in Angstrom for ``nwater`` waters."""
    rng = numpy.random.default_rng(seed)
    mono = _monomer()
    oxy = numpy.empty((nwater, 3))
    atoms = []
    n = 0
    tries = 0
    while n < nwater:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f'cannot place {nwater} waters in a {box} A cube')
        rot = _random_rotation(rng)
        pos = rng.uniform(0., box, 3)
        if n and numpy.min(numpy.linalg.norm(oxy[:n] - pos, axis=1)) < min_oo:
            continue
        oxy[n] = pos
        xyz = mono @ rot.T + pos
        atoms += [('O', tuple(xyz[0])), ('H', tuple(xyz[1])), ('H', tuple(xyz[2]))]
        n += 1
    return atoms


def box_for_density(nwater, grams_per_cm3=1.0):
    """Cube side (Angstrom) holding ``nwater`` waters at the given density."""
    vol_cm3 = nwater * WATER_MASS_G / grams_per_cm3
    return (vol_cm3 * 1e24) ** (1. / 3)


def random_density(nao, nocc, seed):
    """D = 2 C C^T, C = Q of a seeded (nao, nocc) standard-normal matrix."""
    rng = numpy.random.default_rng(seed)
    c = numpy.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    return 2. * c @ c.T


def water_mole(nwater, box, seed, basis, **kwargs):
    from byteqc_amd.gto import M
    return M(atom=water_cluster(nwater, box, seed), basis=basis, **kwargs)


# Named workloads (sizes fixed; seeds recorded here and in every result).
WORKLOADS = {
    'c1': dict(nwater=8, box=9.0, basis='cc-pvdz', seed=1, dm_seed=101),
    'c2': dict(nwater=64, box=18.0, basis='cc-pvdz', seed=2, dm_seed=102),
    'c4': dict(nwater=128, box=23.0, basis='cc-pvtz', seed=4, dm_seed=104),
}


def workload(name, **override):
    """(mol, dm, spec) for one of the named single-molecule workloads."""
    spec = dict(WORKLOADS[name])
    spec.update(override)
    mol = water_mole(spec['nwater'], spec['box'], spec['seed'], spec['basis'])
    dm = random_density(mol.nao_nr(), 5 * spec['nwater'], spec['dm_seed'])
    return mol, dm, spec


def batch_specs(nmol=256, seed=3, lo=4, hi=40):
    """Per-molecule (nwater, box, seed) for the multi-molecule batch."""
    rng = numpy.random.default_rng(seed)
    counts = rng.integers(lo, hi + 1, size=nmol)
    return [dict(nwater=int(n), box=box_for_density(int(n)), seed=seed * 100003 + i,
                 dm_seed=seed * 200003 + i, basis='def2-svp')
            for i, n in enumerate(counts)]
