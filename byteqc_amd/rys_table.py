"""Host-side reader of the Rys interpolation table compiled into the library.

Evaluates the same interpolants the device code evaluates (Clenshaw on the
unit interval containing x; large-x Gauss-Hermite limit above XBIG(n)), so
the table data can be checked without a GPU.
"""
import os
import re

import numpy

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'csrc', 'rys_table.inc')
_CACHE = {}


def _load():
    if not _CACHE:
        text = open(_PATH).read()
        _CACHE['nmax'] = int(re.search(r'BQC_RYS_NMAX (\d+)', text).group(1))
        _CACHE['deg'] = int(re.search(r'BQC_RYS_DEG (\d+)', text).group(1))

        def arr(name, typ):
            body = re.search(name + r'\[\d+\] = \{(.*?)\};', text, re.S).group(1)
            return numpy.array([typ(v) for v in body.replace('\n', ' ').split(',')
                                if v.strip()])
        _CACHE['off'] = arr('BQC_RYS_OFFSET', int)
        _CACHE['coef'] = arr('BQC_RYS_COEF', float)
        _CACHE['hs'] = arr('BQC_RYS_HERM_S', float)
        _CACHE['hw'] = arr('BQC_RYS_HERM_W', float)
    return _CACHE


def xbig(n):
    return 36 + 8 * n


def rys(n, x):
    """(u, w) for n roots at x, as the device evaluates them."""
    t = _load()
    nmax, deg = t['nmax'], t['deg']
    if x >= xbig(n):
        s = t['hs'][(n - 1) * nmax:(n - 1) * nmax + n]
        w = t['hw'][(n - 1) * nmax:(n - 1) * nmax + n]
        return s / (x - s), w / numpy.sqrt(x)
    it = int(x)
    y = 2 * (x - it) - 1
    base = t['off'][n] + it * 2 * n * (deg + 1)
    c = t['coef'][base:base + 2 * n * (deg + 1)].reshape(2 * n, deg + 1)
    b1 = numpy.zeros(2 * n)
    b2 = numpy.zeros(2 * n)
    for j in range(deg, 0, -1):
        b1, b2 = 2 * y * b1 - b2 + c[:, j], b1
    v = y * b1 - b2 + c[:, 0]
    return v[:n], v[n:]
