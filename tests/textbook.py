"""Closed-form change-of-basis matrices that hi2lo must reproduce (small n).

hi2lo takes a column of order m to order m%2 (sph, disk) or 0 (tri): it is
the orthonormal change of basis between two families of orthonormal functions
(docs/docs.doc:166-175 sph, :215-218 disk, :328-332 tri).  Entry (k, i) is the
inner product of input basis function i with output basis function k,
evaluated by Gauss quadrature that is exact for the polynomial integrands.

* sph:  P~_{m+i}^m(x) in the basis P~_{p+k}^p(x), p = m % 2, on [-1, 1]
        (normalised associated Legendre functions, positive phase).
* tri:  (1-t)^m P~_i^{(2m+b, a)}(t) in the basis P~_k^{(b, a)}(t),
        b = beta+gamma+1, a = alpha, weight (1-t)^b (1+t)^a.
* disk: (1+u)^{floor(m/2)} P~_i^{(beta, m+alpha)}(u) in the basis
        P~_k^{(beta, p+alpha)}(u), p = m % 2, weight (1-u)^beta (1+u)^{p+alpha}.

P~^{(a,b)} is the Jacobi polynomial orthonormal for (1-t)^a (1+t)^b.
"""
from math import exp, lgamma, log

import numpy as np
from scipy.special import eval_jacobi, gammaln, lpmv, roots_jacobi, roots_legendre


def jacobi_orthonormal(k, a, b, t):
    lh = (a + b + 1) * log(2) - log(2 * k + a + b + 1) + lgamma(k + a + 1) + lgamma(k + b + 1) \
        - lgamma(k + a + b + 1) - lgamma(k + 1)
    return eval_jacobi(k, a, b, t) / exp(0.5 * lh)


def legendre_orthonormal(l, m, x):
    return (-1) ** m * np.sqrt((2 * l + 1) / 2 * np.exp(gammaln(l - m + 1) - gammaln(l + m + 1))) * lpmv(m, l, x)


def sph_matrix(n, m):
    x, w = roots_legendre(2 * n + 4)
    p = m % 2
    C = np.zeros((n, n - m))
    for i in range(n - m):
        f = legendre_orthonormal(m + i, m, x)
        for k in range(n - p):
            C[k, i] = np.sum(w * f * legendre_orthonormal(p + k, p, x))
    return C


def tri_matrix(n, m, alpha, beta, gamma):
    b = beta + gamma + 1
    t, w = roots_jacobi(2 * n + 4, b, alpha)
    C = np.zeros((n, n - m))
    for i in range(n - m):
        f = (1 - t) ** m * jacobi_orthonormal(i, 2 * m + b, alpha, t)
        for k in range(n):
            C[k, i] = np.sum(w * f * jacobi_orthonormal(k, b, alpha, t))
    return C


def disk_matrix(n, m, alpha, beta):
    p = m % 2
    rows = n - (m + 1) // 2
    u, w = roots_jacobi(2 * n + 4 + m, beta, p + alpha)
    C = np.zeros((n, rows))
    for i in range(rows):
        f = (1 + u) ** (m // 2) * jacobi_orthonormal(i, beta, m + alpha, u)
        for k in range(n):
            C[k, i] = np.sum(w * f * jacobi_orthonormal(k, beta, p + alpha, u))
    return C


def matrix(kind, n, m, params):
    if kind == "sph":
        return sph_matrix(n, m)
    if kind == "tri":
        return tri_matrix(n, m, *params)
    return disk_matrix(n, m, *params)


def input_rows(kind, n, m):
    """Rows a column of order m holds before hi2lo."""
    if kind == "sph":
        return n - m
    if kind == "tri":
        return n - m
    return n - (m + 1) // 2


def orders(kind, n):
    return range(2, 2 * n - 1) if kind == "disk" else (range(1, n) if kind == "tri" else range(2, n))


def low_order(kind, m):
    return 0 if kind == "tri" else m % 2
