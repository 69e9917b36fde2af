"""The annulus rotation plan (ft_plan_rotannulus, rotations.c:194-260) on the CPU.

The oracle's plan (oracle/ft_oracle.c) is pinned by

* an independent dense computation: numpy's Householder QR (diagonal of R
  made positive, which fixes the factorization) of the shifted Jacobi matrix,
  the dense product R Q for the next order, and numpy's Cholesky factor for
  the odd orders.  The Jacobi matrix there comes from the textbook recurrence
  of normalized Jacobi polynomials and is itself checked against scipy's
  Gauss-Jacobi nodes (its eigenvalues).  The product of the oracle's Givens
  rotations for every order must equal numpy's Q;
* the reference's own known answer (test_drivers.c:996-1006): at rho = 0 the
  annulus rotations are the disk's, within its bounds 6 sqrt(N) eps (rel-2)
  and 5 N eps (rel-inf);
* s^2 + c^2 = 1 and the round trip through the disk executors.

The library's plan must equal the oracle's bit for bit.
"""
import numpy as np
import pytest
from scipy.special import roots_jacobi

import oracle as O

EPS = 2.0 ** -52


def _jacobi_matrix(N, a, b):
    """Symmetric Jacobi matrix of the orthonormal P^{(a,b)} (weight (1-x)^a (1+x)^b)."""
    j = np.arange(N, dtype=float)
    s = 2 * j + a + b
    with np.errstate(divide="ignore", invalid="ignore"):
        diag = np.where(j == 0, (b - a) / (a + b + 2), (b * b - a * a) / (s * (s + 2)))
    k = np.arange(1, N, dtype=float)
    t = 2 * k + a + b
    with np.errstate(divide="ignore", invalid="ignore"):
        off = 2 / t * np.sqrt(k * (k + a) * (k + b) * (k + a + b) / ((t - 1) * (t + 1)))
    if N > 1:  # k = 1: t - 1 = k + a + b cancels (0/0 when a + b = -1)
        off[0] = 2 / (a + b + 2) * np.sqrt((a + 1) * (b + 1) / (a + b + 3))
    return np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)


def _dense_annulus(n, alpha, beta, rho):
    """{order m: the orthogonal factor Q of that order's QR}, computed densely."""
    N = n + 1
    t = 1 + 2 * rho * rho / (1 - rho * rho)
    J = _jacobi_matrix(N, beta, alpha)  # P^{(beta, alpha)}, as rotations.c:202 asks
    nodes = roots_jacobi(N, beta, alpha)[0]
    assert np.allclose(np.sort(np.linalg.eigvalsh(J)), np.sort(nodes), atol=1e-13)
    out = {}
    for odd in (0, 1):
        K = J + t * np.eye(N)
        if odd:
            U = np.linalg.cholesky(K).T  # K = U^T U
            K = (U @ U.T)[:-1, :-1]      # U K U^{-1}, leading block
        for m in range(odd, 2 * n - 1, 2):
            Q, R = np.linalg.qr(K)
            D = np.sign(np.diag(R))
            Q, R = Q * D, D[:, None] * R
            out[m] = Q
            K = (R @ Q)[:-1, :-1]
            K = np.tril(K) + np.tril(K, -1).T
    return out


def _givens_product(P, m, size):
    """Q = G_0 G_1 ... G_{size-2}, G_l = [[c, s], [-s, c]] on rows (l, l+1)."""
    G = np.eye(size)
    for l in range(size - 1):
        i = P.index(l, m)
        c, s = P.c[i], P.s[i]
        Gl = np.eye(size)
        Gl[l, l], Gl[l, l + 1], Gl[l + 1, l], Gl[l + 1, l + 1] = c, s, -s, c
        G = G @ Gl
    return G


@pytest.mark.parametrize("n,alpha,beta,rho", [(1, 0.0, 0.0, 0.3), (3, 0.0, 0.0, 0.3), (12, 0.5, -0.25, 0.6),
                                              (30, 0.0, 0.0, np.sqrt(1 / 3)), (40, 1.5, 2.0, 0.2),
                                              (40, -0.5, -0.5, 0.9)])
def test_annulus_plan_against_dense_qr(n, alpha, beta, rho):
    P = O.Plan("annulus", n, alpha, beta, rho=rho)
    for m, Q in _dense_annulus(n, alpha, beta, rho).items():
        assert Q.shape[0] - 1 == n - (m + 1) // 2  # one rotation per stored entry
        assert np.max(np.abs(_givens_product(P, m, Q.shape[0]) - Q)) < 1e-13, m


@pytest.mark.parametrize("alpha,beta", [(0.0, 0.0), (-0.5, -0.5), (1.0, 0.5)])
def test_annulus_at_rho_zero_is_the_disk(alpha, beta):
    """test_drivers.c:996-1006 with its bounds."""
    for n in (64, 65, 128, 257):
        A = O.Plan("annulus", n, alpha, beta, rho=0.0)
        D = O.Plan("disk", n, alpha, beta)
        e2 = np.hypot(np.linalg.norm(A.s - D.s), np.linalg.norm(A.c - D.c)) / \
            np.hypot(np.linalg.norm(D.s), np.linalg.norm(D.c))
        einf = max(np.max(np.abs(A.s - D.s)), np.max(np.abs(A.c - D.c))) / \
            max(np.max(np.abs(D.s)), np.max(np.abs(D.c)))
        assert e2 < 6 * np.sqrt(n) * EPS, e2
        assert einf < 5 * n * EPS, einf


def test_annulus_unit_norm_and_roundtrip():
    for n, rho in ((64, 0.5), (129, np.sqrt(1 / 3))):
        P = O.Plan("annulus", n, 0.0, 0.0, rho=rho)
        assert np.max(np.abs(P.s ** 2 + P.c ** 2 - 1)) <= 4 * EPS
        M = 4 * n - 3
        A = O.fill("disk", n, M, 13)
        X = O.execute(P, O.LO2HI, O.execute(P, O.HI2LO, A.copy(order="F")))
        assert O.rel2(X, A) < 4 * np.sqrt(n) * EPS
        assert O.relinf(X, A) < 2 * n * EPS


@pytest.mark.parametrize("n,alpha,beta,rho", [(1, 0.0, 0.0, 0.0), (2, 0.0, 0.0, 0.3), (17, 0.5, -0.25, 0.6),
                                              (256, 0.0, 0.0, np.sqrt(1 / 3)), (300, 1.5, 2.0, 0.2)])
def test_product_annulus_plan_bitwise_equal_oracle(ft, n, alpha, beta, rho):
    P = ft.plan_rotannulus(n, alpha, beta, 0.0, rho)
    Q = O.Plan("annulus", n, alpha, beta, rho=rho)
    assert P.kind == ft.DISK and P.family == "annulus" and P.size == n * n
    assert np.array_equal(P.s, Q.s) and np.array_equal(P.c, Q.c)


def test_product_annulus_rejects_unimplemented(ft):
    """gamma != 0 is not implemented in the reference (rotations.c:252-254);
    rho must lie in [0, 1)."""
    for args in ((8, 0.0, 0.0, 0.5, 0.3), (8, 0.0, 0.0, 0.0, 1.0), (8, 0.0, 0.0, 0.0, -0.1), (0, 0.0, 0.0, 0.0, 0.3)):
        with pytest.raises(ft.GPUError):
            ft.plan_rotannulus(*args)
    assert ft.last_error() == 0
