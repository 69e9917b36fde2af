"""The connection matrices of the harmonic transforms (transforms.c:50-93) and
the harmonic plan, on the CPU.

The oracle's matrices (oracle/ft_oracle.c) are pinned against exact rational
conversions that do not use the reference's method: the three-term
recurrences of the Legendre and Gegenbauer polynomials carried out in the
Chebyshev (T) or Gegenbauer C^mu basis with Python fractions, then scaled by
the textbook normalizations (orthonormal Legendre sqrt(l + 1/2); orthonormal
Chebyshev 1/sqrt(pi), sqrt(2/pi); orthonormal Gegenbauer 1/sqrt(h_l),
h_l = pi 2^(1-2 lambda) Gamma(l + 2 lambda) / (l! (l + lambda) Gamma(lambda)^2)).
Inverse pairs multiply to the identity.  The library's plan matrices equal
the oracle's bit for bit, and the composite oracle inverts itself.  The
Jacobi connections of tri2cheb and disk2cxf (the library's host code) are
pinned against scipy's Jacobi polynomials evaluated at Chebyshev points.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

EPS = 2.0 ** -52


def _legendre_in_chebyshev(n):
    """C[k, l]: P_l = sum_k C[k, l] T_k, exactly (x T_0 = T_1, x T_k = (T_{k-1} + T_{k+1}) / 2)."""
    def times_x(v):
        out = [Fraction(0)] * (len(v) + 1)
        for k, a in enumerate(v):
            if a == 0:
                continue
            if k == 0:
                out[1] += a
            else:
                out[k - 1] += a / 2
                out[k + 1] += a / 2
        return out
    P = [[Fraction(1)], [Fraction(0), Fraction(1)]]
    for l in range(1, n - 1):
        xp = times_x(P[l])
        prev = P[l - 1] + [Fraction(0)] * (len(xp) - len(P[l - 1]))
        P.append([((2 * l + 1) * a - l * b) / (l + 1) for a, b in zip(xp, prev)])
    C = np.zeros((n, n))
    for l in range(n):
        for k, a in enumerate(P[l][:n]):
            C[k, l] = float(a)
    return C


def _gegenbauer_in_gegenbauer(n, lam, mu):
    """C[k, l]: C^lam_l = sum_k C[k, l] C^mu_k, exactly (lam, mu rational > 0):
    (l+1) C_{l+1} = 2 (l + lam) x C_l - (l + 2 lam - 1) C_{l-1}, and
    x C^mu_k = ((k + 1) C^mu_{k+1} + (k + 2 mu - 1) C^mu_{k-1}) / (2 (k + mu))."""
    lam, mu = Fraction(lam), Fraction(mu)

    def times_x(v):
        out = [Fraction(0)] * (len(v) + 1)
        for k, a in enumerate(v):
            if a == 0:
                continue
            out[k + 1] += a * (k + 1) / (2 * (k + mu))
            if k > 0:
                out[k - 1] += a * (k + 2 * mu - 1) / (2 * (k + mu))
        return out
    G = [[Fraction(1)], times_x([2 * lam])]
    G[1] = [a for a in G[1]]
    for l in range(1, n - 1):
        xg = times_x(G[l])
        prev = G[l - 1] + [Fraction(0)] * (len(xg) - len(G[l - 1]))
        G.append([(2 * (l + lam) * a - (l + 2 * lam - 1) * b) / (l + 1) for a, b in zip(xg, prev)])
    C = np.zeros((n, n))
    for l in range(n):
        for k, a in enumerate(G[l][:n]):
            C[k, l] = float(a)
    return C


def _gegenbauer_norm(l, lam):
    """sqrt(h_l) of C^lam_l under the weight (1 - x^2)^(lam - 1/2)."""
    lg = math.lgamma(l + 2 * lam) - math.lgamma(l + 1) - 2 * math.lgamma(lam)
    return math.sqrt(math.pi * 2.0 ** (1 - 2 * lam) / (l + lam) * math.exp(lg))


def _close(V, W):
    return np.max(np.abs(V - W)) <= 8 * EPS * max(1.0, np.max(np.abs(W)))


@pytest.mark.parametrize("normleg,normcheb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_legendre_to_chebyshev_exact(normleg, normcheb):
    n = 40
    C = _legendre_in_chebyshev(n)
    dl = np.array([math.sqrt(l + 0.5) if normleg else 1.0 for l in range(n)])
    dc = np.array([(1 / math.sqrt(math.pi) if k == 0 else math.sqrt(2 / math.pi)) if normcheb else 1.0
                   for k in range(n)])
    want = C * dl[None, :] / dc[:, None]
    V = O.connection("leg2cheb", normleg, normcheb, n)
    assert _close(V, want)
    W = O.connection("cheb2leg", normcheb, normleg, n)
    assert np.max(np.abs(W @ V - np.eye(n))) < 64 * EPS


@pytest.mark.parametrize("lam,mu", [(1.5, 1.0), (1.0, 1.5), (0.5, 1.0), (2.0, 0.5), (2.5, 1.5)])
@pytest.mark.parametrize("norm1,norm2", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_ultraspherical_exact(lam, mu, norm1, norm2):
    n = 30
    C = _gegenbauer_in_gegenbauer(n, lam, mu)
    d1 = np.array([1 / _gegenbauer_norm(l, lam) if norm1 else 1.0 for l in range(n)])
    d2 = np.array([1 / _gegenbauer_norm(k, mu) if norm2 else 1.0 for k in range(n)])
    want = C * d1[None, :] / d2[:, None]
    V = O.connection("ultra2ultra", norm1, norm2, n, lam, mu)
    assert np.max(np.abs(V - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))
    W = O.connection("ultra2ultra", norm2, norm1, n, mu, lam)
    assert np.max(np.abs(W @ V - np.eye(n))) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 17, 256])
def test_product_connection_matrices_bitwise_equal_oracle(ft, n):
    H = ft.plan_sph2fourier(n)
    assert np.array_equal(H.P(0), O.connection("leg2cheb", 1, 0, n))
    assert np.array_equal(H.P(1), O.connection("ultra2ultra", 1, 0, n, 1.5, 1.0))
    assert np.array_equal(H.Pinv(0), O.connection("cheb2leg", 0, 1, n))
    assert np.array_equal(H.Pinv(1), O.connection("ultra2ultra", 0, 1, n, 1.0, 1.5))
    head = H._head()
    assert head.NRP == 1 and head.NP == 2 and head.NMP == 0 and not head.B


def _sph_plan(n):
    return O.Plan("sph", n), O.connection("leg2cheb", 1, 0, n), O.connection("ultra2ultra", 1, 0, n, 1.5, 1.0), \
        O.connection("cheb2leg", 0, 1, n), O.connection("ultra2ultra", 0, 1, n, 1.0, 1.5)


@pytest.mark.parametrize("vec", [False, True])
def test_harmonic_oracle_round_trip_and_transpose(vec):
    """fourier2sph inverts sph2fourier (README.md:55-56 reports ~1e-15 for the
    reference); 'T' is the transpose of 'N': <y, F x> = <F^T y, x>."""
    n = 64
    Q, P0, P1, I0, I1 = _sph_plan(n)
    M = 2 * n - 1
    fwd, inv = ("sphv2fourier", "fourier2sphv") if vec else ("sph2fourier", "fourier2sph")
    A = O.fill("sph", n, M, 5)
    X = O.harmonic(inv, "N", Q, (P0, P1), (I0, I1), O.harmonic(fwd, "N", Q, (P0, P1), (I0, I1), A.copy(order="F")))
    assert O.rel2(X, A) < 4 * np.sqrt(n) * EPS
    rng = np.random.default_rng(1)
    x = np.asfortranarray(rng.standard_normal((n, M)))
    y = np.asfortranarray(rng.standard_normal((n, M)))
    Fx = O.harmonic(fwd, "N", Q, (P0, P1), (I0, I1), x.copy(order="F"))
    Fty = O.harmonic(fwd, "T", Q, (P0, P1), (I0, I1), y.copy(order="F"))
    assert abs(np.sum(y * Fx) - np.sum(Fty * x)) < 1e-12 * np.sqrt(np.sum(y * y) * np.sum(x * x)) * n


def test_harmonic_plan_rejects(ft):
    with pytest.raises(ft.GPUError):
        ft.plan_sph2fourier(0)
    H = ft.plan_sph2fourier(8)
    A = np.zeros((8, 15), order="F")
    with pytest.raises(ft.GPUError):
        ft.execute_sph2fourier("X", H, A)        # TRANS
    with pytest.raises(ft.GPUError):
        ft.execute_sph2fourier("N", H, np.zeros((8, 16), order="F"))  # even M
    with pytest.raises(ft.GPUError):
        ft.execute_sph2fourier("N", H, np.zeros((9, 15), order="F"), N=9)  # N != plan's n
    assert ft.last_error() == 0


# ----------------------------------------------------------------- Jacobi ----

def _jacobi_norm(l, a, b):
    """sqrt(h_l) of P_l^{(a,b)} under (1-x)^a (1+x)^b."""
    if l == 0:
        lg = (a + b + 1) * math.log(2) + math.lgamma(a + 1) + math.lgamma(b + 1) - math.lgamma(a + b + 2)
    else:
        lg = (a + b + 1) * math.log(2) + math.lgamma(l + a + 1) + math.lgamma(l + b + 1) - \
            math.log(2 * l + a + b + 1) - math.lgamma(l + a + b + 1) - math.lgamma(l + 1)
    return math.exp(0.5 * lg)


def _jacobi_connection_by_evaluation(n, a, b, c, d, norm1, norm2):
    """V with P_l^{(a,b)} = sum_k V[k, l] P_k^{(c,d)} (normalized where asked),
    from scipy's Jacobi polynomials at n Chebyshev points (a square solve)."""
    from scipy.special import eval_jacobi
    x = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    src = np.array([eval_jacobi(l, a, b, x) / (_jacobi_norm(l, a, b) if norm1 else 1) for l in range(n)]).T
    dst = np.array([eval_jacobi(k, c, d, x) / (_jacobi_norm(k, c, d) if norm2 else 1) for k in range(n)]).T
    return np.linalg.solve(dst, src)


def _lib_matrix(ft, name, n, *args):
    fn = getattr(ft._lib, name)
    fn.restype = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_int] * (len(args) - sum(isinstance(v, float) for v in args)) + \
        [ctypes.c_double] * sum(isinstance(v, float) for v in args)
    p = fn(*args)
    V = np.ctypeslib.as_array(p, shape=(n * n,)).copy().reshape(n, n, order="F")
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(p, ctypes.c_void_p))
    return V


@pytest.mark.parametrize("a,b,c,d", [(0.0, 0.0, -0.5, -0.5), (1.5, 0.0, -0.5, -0.5), (0.5, -0.25, -0.5, 0.5),
                                     (-0.5, -0.5, 1.0, 0.0), (-0.5, 0.5, 0.0, 1.0), (2.0, 1.0, 0.5, 0.25)])
@pytest.mark.parametrize("norm1,norm2", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_jacobi_connection_against_evaluation(ft, a, b, c, d, norm1, norm2):
    """The library's Jacobi -> Jacobi matrix (transforms.c:95-108) against
    scipy's Jacobi polynomials at n Chebyshev points; it is upper triangular."""
    n = 24
    V = _lib_matrix(ft, "ftamd_plan_jac2jac", n, norm1, norm2, n, a, b, c, d)
    W = _jacobi_connection_by_evaluation(n, a, b, c, d, norm1, norm2)
    assert np.array_equal(V, np.triu(V))
    assert np.max(np.abs(V - W)) <= 1e-11 * max(1.0, np.max(np.abs(W)))
    Vi = _lib_matrix(ft, "ftamd_plan_jac2jac", n, norm2, norm1, n, c, d, a, b)
    assert np.max(np.abs(Vi @ V - np.eye(n))) < 1e-12


def test_jacobi_chebyshev_scalings(ft):
    """transforms.c:157-183: the unnormalized Chebyshev side scales rows (or
    columns) of the normalized connection by sqrt(1/pi), sqrt(2/pi)."""
    n = 20
    V = _lib_matrix(ft, "ftamd_plan_jac2jac", n, 1, 1, n, 0.5, -0.25, -0.5, -0.5)
    C = _lib_matrix(ft, "ftamd_plan_jac2cheb", n, 1, 0, n, 0.5, -0.25)
    row = np.array([math.sqrt(1 / math.pi)] + [math.sqrt(2 / math.pi)] * (n - 1))
    assert np.max(np.abs(C - V * row[:, None])) <= 2 * EPS * np.max(np.abs(C))  # the constants' last bit
    Vi = _lib_matrix(ft, "ftamd_plan_jac2jac", n, 1, 1, n, -0.5, -0.5, 0.5, -0.25)
    Ci = _lib_matrix(ft, "ftamd_plan_cheb2jac", n, 0, 1, n, 0.5, -0.25)
    col = np.array([math.sqrt(math.pi)] + [math.sqrt(math.pi / 2)] * (n - 1))
    assert np.max(np.abs(Ci - Vi * col[None, :])) <= 2 * EPS * np.max(np.abs(Ci))
    assert np.max(np.abs(Ci @ C - np.eye(n))) < 1e-12


@pytest.mark.parametrize("which,params", [("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)), ("tri", (0.5, 1.0, -0.25)),
                                          ("disk", (0.0, 0.0)), ("disk", (0.5, 1.5))])
def test_tri_disk_plans_and_oracle_round_trip(ft, which, params):
    """The plans hold the reference's connection matrices (drivers.c:901-985);
    the composite oracle inverts itself and 'T' is the transpose of 'N'."""
    n = 40
    if which == "tri":
        H = ft.plan_tri2cheb(n, *params)
        a, b, g = params
        assert np.array_equal(H.P(0), _lib_matrix(ft, "ftamd_plan_jac2jac", n, 1, 1, n, b + g + 1.0, a, -0.5, -0.5))
        assert np.array_equal(H.Pinv(1), _lib_matrix(ft, "ftamd_plan_jac2jac", n, 1, 1, n, -0.5, -0.5, g, b))
        Q, M, fwd, inv = O.Plan("tri", n, *params), n, "tri2cheb", "cheb2tri"
    else:
        H = ft.plan_disk2cxf(n, *params)
        a, b = params
        cst0 = 2.0 ** (0.5 * (a + b + 2.0))
        assert np.array_equal(H.P(0), np.triu(_lib_matrix(ft, "ftamd_plan_jac2cheb", n, 1, 0, n, b, a)) * cst0)
        Q, M, fwd, inv = O.Plan("disk", n, *params), 4 * n - 3, "disk2cxf", "cxf2disk"
    P, Pinv = (H.P(0), H.P(1)), (H.Pinv(0), H.Pinv(1))
    A = O.fill(which, n, M, 3)
    X = O.harmonic(inv, "N", Q, P, Pinv, O.harmonic(fwd, "N", Q, P, Pinv, A.copy(order="F"), params=params),
                   params=params)
    assert O.rel2(X, A) < 1e-13
    rng = np.random.default_rng(2)
    x = np.asfortranarray(rng.standard_normal((n, M)))
    y = np.asfortranarray(rng.standard_normal((n, M)))
    Fx = O.harmonic(fwd, "N", Q, P, Pinv, x.copy(order="F"), params=params)
    Fty = O.harmonic(fwd, "T", Q, P, Pinv, y.copy(order="F"), params=params)
    assert abs(np.sum(y * Fx) - np.sum(Fty * x)) < 1e-12 * np.sqrt(np.sum(y * y) * np.sum(x * x)) * n



@pytest.mark.parametrize("beta", [0.0, 0.5, -0.25])
def test_rectdisk2cheb_plan_and_oracle_round_trip(ft, beta):
    """ft_plan_rectdisk2cheb's matrices (drivers.c:1123-1141): P[0] takes the
    normalized ultraspherical(beta+1) polynomials (Jacobi(beta+1/2, beta+1/2))
    to Chebyshev T (checked against scipy's polynomials at Chebyshev points);
    inverse pairs multiply to the identity; the composite oracle inverts
    itself."""
    n = 30
    H = ft.plan_rectdisk2cheb(n, beta)
    head = H._head()
    assert head.NP == 3
    from scipy.special import eval_chebyt, eval_jacobi
    x = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    lam = beta + 1.0
    src = np.array([eval_jacobi(l, lam - 0.5, lam - 0.5, x) / _jacobi_norm(l, lam - 0.5, lam - 0.5)
                    for l in range(n)]).T
    dst = np.array([eval_chebyt(k, x) for k in range(n)]).T
    W = np.linalg.solve(dst, src)  # normalized ultraspherical(lam) in the Chebyshev T basis
    assert np.max(np.abs(H.P(0) - W)) <= 1e-11 * np.max(np.abs(W))
    Pm = [H.P(k) for k in range(3)]
    Pi = [np.ctypeslib.as_array(head.Pinv[k], shape=(n * n,)).copy().reshape(n, n, order="F") for k in range(3)]
    for k in range(3):
        assert np.max(np.abs(Pi[k] @ Pm[k] - np.eye(n))) < 1e-12
    Q = O.Plan("rectdisk", n, beta)
    A = O.fill("tri", n, n, 9)
    X = O.harmonic("cheb2rectdisk", "N", Q, Pm, Pi, O.harmonic("rectdisk2cheb", "N", Q, Pm, Pi, A.copy(order="F")))
    assert O.rel2(X, A) < 1e-13


@pytest.mark.parametrize("alpha,beta,rho", [(0.0, 0.0, math.sqrt(1.0 / 3.0)), (0.5, -0.25, 0.2), (-0.5, -0.5, 0.9),
                                            (1.5, 0.5, 0.0)])
def test_annulus_factor_is_the_gram_cholesky(ft, alpha, beta, rho):
    """ft_plan_ann2cxf's MP[1] (drivers.c:1036-1046): R^T R is the Gram matrix
    of the orthonormal Jacobi(beta, alpha) polynomials under the extra weight
    t + x, t = (1 + rho^2) / (1 - rho^2) (the annulus's r^2 in x), computed by
    Gauss-Jacobi quadrature with scipy's polynomials; the oracle's operator
    has the Gauss-Jacobi nodes as eigenvalues and its Cholesky factor agrees
    with the library's to rounding."""
    from scipy.special import eval_jacobi, roots_jacobi
    n = 40
    H = ft.plan_ann2cxf(n, alpha, beta, 0.0, rho)
    R = H.R()
    x, w = roots_jacobi(n + 1, beta, alpha)
    P = np.array([eval_jacobi(k, beta, alpha, x) / _jacobi_norm(k, beta, alpha) for k in range(n)])
    t = (1 + rho * rho) / (1 - rho * rho)
    G = (P * (w * (t + x))) @ P.T
    assert np.max(np.abs(R.T @ R - G)) <= 1e-12 * np.max(np.abs(G))
    d, e = O.jacobi_operator(n, beta, alpha)
    nodes = np.sort(np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1)))
    assert np.max(np.abs(nodes - np.sort(roots_jacobi(n, beta, alpha)[0]))) < 1e-14
    Ro = O.annulus_factor(n, alpha, beta, rho)
    assert np.max(np.abs(R - Ro)) <= 8 * EPS * np.max(np.abs(Ro))


@pytest.mark.parametrize("alpha,beta,rho", [(0.0, 0.0, math.sqrt(1.0 / 3.0)), (0.5, 1.5, 0.3)])
def test_ann2cxf_plan_and_oracle_round_trip(ft, alpha, beta, rho):
    """The plan holds Jacobi(beta, alpha) -> Chebyshev scaled by sqrt(2) (2 /
    (1 - rho^2))^((alpha + beta + 1) / 2) (drivers.c:1054-1063), and its
    inverse by the reciprocal; the composite oracle (drivers.c:1073-1121)
    inverts itself and its 'T' is the transpose of its 'N'."""
    n = 40
    H = ft.plan_ann2cxf(n, alpha, beta, 0.0, rho)
    head = H._head()
    assert head.NP == 1 and head.NMP == 2 and head.rho == rho
    cst = math.sqrt(2.0) * (2.0 / (1.0 - rho * rho)) ** (0.5 * (alpha + beta + 1.0))
    C = np.triu(_lib_matrix(ft, "ftamd_plan_jac2cheb", n, 1, 0, n, beta, alpha))
    assert np.max(np.abs(H.P(0) - C * cst)) <= 2 * EPS * np.max(np.abs(C * cst))
    assert np.max(np.abs(H.Pinv(0) @ H.P(0) - np.eye(n))) < 1e-12
    Q, M = O.Plan("annulus", n, alpha, beta, rho=rho), 4 * n - 3
    P, Pinv, R = (H.P(0),), (H.Pinv(0),), H.R()
    A = O.fill("disk", n, M, 5)
    X = O.harmonic("cxf2ann", "N", Q, P, Pinv, O.harmonic("ann2cxf", "N", Q, P, Pinv, A.copy(order="F"), R=R), R=R)
    assert O.rel2(X, A) < 1e-13
    rng = np.random.default_rng(4)
    x = np.asfortranarray(rng.standard_normal((n, M)))
    y = np.asfortranarray(rng.standard_normal((n, M)))
    Fx = O.harmonic("ann2cxf", "N", Q, P, Pinv, x.copy(order="F"), R=R)
    Fty = O.harmonic("ann2cxf", "T", Q, P, Pinv, y.copy(order="F"), R=R)
    assert abs(np.sum(y * Fx) - np.sum(Fty * x)) < 1e-12 * np.sqrt(np.sum(y * y) * np.sum(x * x)) * n


@pytest.mark.parametrize("args", [(10, 0.0, 0.0, 0.5, 0.5), (10, 0.0, 0.0, 0.0, 1.0), (10, 0.0, 0.0, 0.0, -0.1),
                                  (0, 0.0, 0.0, 0.0, 0.5)])
def test_ann2cxf_rejects_what_the_reference_cannot_plan(ft, args):
    """gamma != 0 (drivers.c:1048-1051 only warns), rho outside [0, 1) and
    n < 1 are errors."""
    with pytest.raises(ft.GPUError):
        ft.plan_ann2cxf(*args)


def test_annulus_plan_mismatch_is_rejected(ft):
    """An annulus plan does not run the disk's transforms, nor a disk plan the
    annulus's (the reference would read members the plan lacks)."""
    A = np.asfortranarray(np.zeros((8, 29)))
    H = ft.plan_ann2cxf(8, 0.0, 0.0, 0.0, 0.5)
    D = ft.plan_disk2cxf(8, 0.0, 0.0)
    with pytest.raises(ft.GPUError):
        ft.execute_disk2cxf("N", H, A)
    with pytest.raises(ft.GPUError):
        ft.execute_ann2cxf("N", D, A)
