"""Write high-precision Boys-function and Rys root/weight vectors.

Everything here is computed from the mathematics with mpmath at 80 digits:

* Boys function F_m(x) = int_0^1 t^(2m) exp(-x t^2) dt
  = gamma_lower(m + 1/2, x) / (2 x^(m + 1/2)),  F_m(0) = 1/(2m+1).
* Rys quadrature of order n: nodes s_i = t_i^2 in (0,1) and weights w_i with
  sum_i w_i s_i^k = F_k(x) for k = 0..2n-1.  Obtained by Golub-Welsch: the
  Cholesky factor of the Hankel moment matrix H_ij = F_(i+j)(x) gives the
  three-term recurrence of the orthogonal polynomials in s, whose Jacobi
  matrix eigenvalues are the nodes and whose first eigenvector components
  give the weights.

Output: rys_golden.npz with arrays x (nx,), boys (nx, 20) and, for every
n in 1..11, roots_n / weights_n (nx, n) as float64 (nodes as u = s/(1-s), the
form the integral recurrences consume, and as s).

Run:  python tests/golden/make_rys_golden.py
"""
import os

import mpmath
import numpy

mpmath.mp.dps = 80
NMAX = 11
MMAX = 20


def boys(m, x):
    x = mpmath.mpf(x)
    if x == 0:
        return mpmath.mpf(1) / (2 * m + 1)
    a = m + mpmath.mpf(1) / 2
    return mpmath.gammainc(a, 0, x) / (2 * x ** a)


def rys(n, x):
    mom = [boys(k, x) for k in range(2 * n + 1)]
    h = mpmath.matrix(n + 1, n + 1)
    for i in range(n + 1):
        for j in range(n + 1):
            h[i, j] = mom[i + j]
    r = mpmath.cholesky(h).T  # upper triangular, H = R^T R
    alpha = []
    beta = []
    for j in range(n):
        a = r[j, j + 1] / r[j, j]
        if j > 0:
            a -= r[j - 1, j] / r[j - 1, j - 1]
        alpha.append(a)
        if j + 1 < n:
            beta.append(r[j + 1, j + 1] / r[j, j])
    jac = mpmath.matrix(n, n)
    for j in range(n):
        jac[j, j] = alpha[j]
        if j + 1 < n:
            jac[j, j + 1] = jac[j + 1, j] = beta[j]
    ev, vec = mpmath.eigsy(jac)
    nodes = [ev[i] for i in range(n)]
    weights = [mom[0] * vec[0, i] ** 2 for i in range(n)]
    order = sorted(range(n), key=lambda i: nodes[i])
    return [nodes[i] for i in order], [weights[i] for i in order]


def x_grid():
    fixed = [0., 1e-14, 1e-10, 1e-7, 1e-5, 1e-3, 0.01, 0.05, 0.1, 0.3, 0.5, 0.9,
             1., 1.5, 2., 2.5, 3., 3.7, 4., 5., 6., 7., 8., 9., 10., 11.5, 12.,
             14., 15., 17., 18., 20., 22., 24., 25., 27., 30., 33., 35., 38., 40.,
             45., 50., 55., 60., 70., 80., 100., 150., 200., 400., 1000., 1e4]
    rng = numpy.random.default_rng(20241202)
    rand = sorted(rng.uniform(0., 80., 40).tolist())
    return numpy.array(sorted(set(fixed + rand)))


def main():
    xs = x_grid()
    out = {'x': xs}
    out['boys'] = numpy.array([[float(boys(m, x)) for m in range(MMAX)] for x in xs])
    for n in range(1, NMAX + 1):
        s_all, u_all, w_all = [], [], []
        for x in xs:
            s, w = rys(n, x)
            s_all.append([float(v) for v in s])
            u_all.append([float(v / (1 - v)) for v in s])
            w_all.append([float(v) for v in w])
        out[f's_{n}'] = numpy.array(s_all)
        out[f'u_{n}'] = numpy.array(u_all)
        out[f'w_{n}'] = numpy.array(w_all)
        print('n =', n, 'done', flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'rys_golden.npz')
    numpy.savez_compressed(path, **out)
    print('wrote', path)


if __name__ == '__main__':
    main()
