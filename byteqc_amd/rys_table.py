"""Host-side reader of the Rys interpolation tables compiled into the library.

Evaluates the same interpolants the device code evaluates (Clenshaw on the
unit interval containing x; large-x Gauss-Hermite limit above XBIG(n)), so
the table data can be checked without a GPU.  rys_table.inc holds n = 1..7
roots, rys_table_hi.inc n = 8..11.
"""
import os
import re

import numpy

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'csrc')
_PARTS = (('rys_table.inc', ''), ('rys_table_hi.inc', '_HI'))
_CACHE = {}


def _load(fname, suffix):
    if fname not in _CACHE:
        text = open(os.path.join(_DIR, fname)).read()
        t = {'nmin': int(re.search(r'BQC_RYS%s_NMIN (\d+)' % suffix, text).group(1)),
             'nmax': int(re.search(r'BQC_RYS%s_NMAX (\d+)' % suffix, text).group(1)),
             'deg': int(re.search(r'BQC_RYS_DEG (\d+)', text).group(1))}

        def arr(name, typ):
            body = re.search(name + r'\[\d+\] = \{(.*?)\};', text, re.S).group(1)
            return numpy.array([typ(v) for v in body.replace('\n', ' ').split(',')
                                if v.strip()])
        t['off'] = arr('BQC_RYS%s_OFFSET' % suffix, int)
        t['coef'] = arr('BQC_RYS%s_COEF' % suffix, float)
        t['hs'] = arr('BQC_RYS%s_HERM_S' % suffix, float)
        t['hw'] = arr('BQC_RYS%s_HERM_W' % suffix, float)
        _CACHE[fname] = t
    return _CACHE[fname]


def _table(n):
    for fname, suffix in _PARTS:
        t = _load(fname, suffix)
        if t['nmin'] <= n <= t['nmax']:
            return t
    raise ValueError(f'no Rys table for {n} roots')


def xbig(n):
    return 36 + 8 * n


def rys(n, x):
    """(u, w) for n roots at x, as the device evaluates them."""
    t = _table(n)
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
