"""Write oracle J/K fixtures for the committed workloads.

jk_c1.npz: 8 waters (9 A cube, seed 1), cc-pVDZ (192 AOs), density
D = 2 C C^T from a QR-orthonormalised seeded (192 x 40) Gaussian (seed 101);
J and K from the CPU oracle (oracle/oracle.py) with Schwarz cutoff 1e-13.

Run: python tests/golden/make_jk_fixtures.py
"""
import os
import sys

import numpy

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from byteqc_amd.systems import workload  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def main():
    mol, dm, spec = workload('c1')
    j, k, nq = orc.get_jk(mol, dm, return_count=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'jk_c1.npz')
    numpy.savez_compressed(path, vj=j, vk=k, nquartets=nq,
                           spec=numpy.array(repr(sorted(spec.items()))))
    print('wrote', path, 'quartets', nq)


if __name__ == '__main__':
    main()
