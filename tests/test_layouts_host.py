"""The vector-field (sphv) and rectangle-disk (rectdisk) executors on the CPU.

The oracle side is pinned first:

* the rectangle-disk plan's closed form (rotations.c:270-287) entry by entry
  against exact rationals, s^2 + c^2 = 1 as an exact identity of its terms,
  and its beta = -1/2 case, where the closed form reduces to the sphere's;
* its rotations are orthogonal per order and lo2hi undoes hi2lo;
* both executors as compositions of the (separately pinned) single-column
  kernels on the columns the reference's drivers name (drivers.c:204-220,
  496-508), with the columns outside the path untouched;
* bit for bit against the reference's own compiled executors (oracle/_ref).

Then the library's host side: its rectangle-disk plan equals the oracle's
bit for bit, and the band-systolic chains it builds for both column layouts,
run through the step-by-step emulation of the kernel, reproduce the oracle.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import sys_emulator as se

EPS = 2.0 ** -52
BETAS = [0.0, 0.5, -0.25, 1.0, 3.5]


def _rectdisk_exact(beta, l, m):
    """(nums, numc, den) of rotations.c:277-279 as exact rationals."""
    b = Fraction(beta)
    return (Fraction((l + 1) * (l + 2)), (2 * m + 2 * b + 3) * (2 * l + 2 * m + 2 * b + 6),
            (l + 2 * m + 2 * b + 4) * (l + 2 * m + 2 * b + 5))


@pytest.mark.parametrize("beta", BETAS)
def test_rectdisk_plan_exact_rounding(beta):
    """Every entry is fl(sqrt(fl(num/den))) of the exact closed form (the terms
    are exact in FP64 for these beta), within an ulp of the true value, and the
    terms satisfy nums + numc = den exactly (so s^2 + c^2 = 1)."""
    n = 30
    P = O.Plan("rectdisk", n, beta)
    for m in range(n):
        for l in range(n - m):
            nums, numc, den = _rectdisk_exact(beta, l, m)
            assert nums + numc == den
            i = P.index(l, m)
            assert P.s[i] == float(np.sqrt(float(nums / den))), (l, m)
            assert P.c[i] == float(np.sqrt(float(numc / den))), (l, m)
    assert np.max(np.abs(P.s ** 2 + P.c ** 2 - 1.0)) <= 4 * EPS


def test_rectdisk_half_is_the_sphere():
    """At beta = -1/2 the rectangle-disk terms are the sphere's (rotations.c:122-124
    with 2m+2beta+3 = 2m+2): the tables agree bit for bit."""
    for n in (1, 5, 64, 200):
        R = O.Plan("rectdisk", n, -0.5)
        S = O.Plan("sph", n)
        assert np.array_equal(R.s, S.s) and np.array_equal(R.c, S.c)


@pytest.mark.parametrize("beta", [0.0, 1.5])
def test_rectdisk_rotations_orthogonal(beta):
    """Per order m, hi2lo is orthogonal and lo2hi is its transpose."""
    n = 16
    P = O.Plan("rectdisk", n, beta)
    for m in range(2, n):
        H = np.zeros((n, n))
        T = np.zeros((n, n))
        for i in range(n):
            e = np.zeros(n)
            e[i] = 1.0
            H[:, i] = O.kernel(P, O.HI2LO, m % 2, m, e.copy())
            T[:, i] = O.kernel(P, O.LO2HI, m % 2, m, e.copy())
        assert np.max(np.abs(H.T @ H - np.eye(n))) < 50 * EPS
        assert np.max(np.abs(T - H.T)) < 50 * EPS


def _compose(P, d, A, pairs):
    """The executor as single-column kernels: (column, order) pairs."""
    X = A.copy(order="F")
    for j, m in pairs:
        col = np.ascontiguousarray(X[:, j])
        O.kernel(P, d, m % 2, m, col)
        X[:, j] = col
    return X


def _sphv_pairs(M):
    return [(2 * m + t, m) for m in range(2, M // 2) for t in (1, 2)]


def _rectdisk_pairs(M):
    return [(m, m) for m in range(2, M)]


@pytest.mark.parametrize("n,M", [(3, 5), (9, 17), (40, 79), (40, 81), (40, 33), (41, 7)])
def test_sphv_oracle_is_the_sphere_kernel_on_its_columns(n, M):
    """drivers.c:204-220: order m on columns 2m+1 and 2m+2, m = 2 .. M/2-1;
    columns 0 .. 4 untouched."""
    P = O.Plan("sph", n)
    A = np.asfortranarray(np.random.default_rng(n * M).standard_normal((n, M)))
    for d in (O.HI2LO, O.LO2HI):
        X = O.execute_sphv(P, d, A.copy(order="F"), M)
        assert np.array_equal(X, _compose(P, d, A, _sphv_pairs(M)))
        assert np.array_equal(X[:, :5], A[:, :5])


@pytest.mark.parametrize("n,M,beta", [(3, 3, 0.0), (12, 12, 0.5), (40, 40, -0.25), (40, 17, 2.0), (64, 1, 0.0)])
def test_rectdisk_oracle_is_its_kernel_on_each_column(n, M, beta):
    """drivers.c:496-508: order m on column m, m = 2 .. M-1; columns 0, 1 untouched."""
    P = O.Plan("rectdisk", n, beta)
    A = np.asfortranarray(np.random.default_rng(n + M).standard_normal((n, M)))
    for d in (O.HI2LO, O.LO2HI):
        X = O.execute(P, d, A.copy(order="F"), M)
        assert np.array_equal(X, _compose(P, d, A, _rectdisk_pairs(M)))
        assert np.array_equal(X[:, :2], A[:, :2])


@pytest.mark.parametrize("layout", ["sphv", "rectdisk"])
def test_roundtrip_bounds(layout):
    """test_drivers.c:52-57 bounds on hi2lo then lo2hi."""
    for n in (65, 128):
        if layout == "sphv":
            P, M = O.Plan("sph", n), 2 * n - 1
            run = O.execute_sphv
        else:
            P, M = O.Plan("rectdisk", n, 0.5), n
            run = O.execute
        A = O.fill("sph" if layout == "sphv" else "tri", n, M, 21)
        X = run(P, O.LO2HI, run(P, O.HI2LO, A.copy(order="F"), M), M)
        assert O.rel2(X, A) < 4 * np.sqrt(n) * EPS
        assert O.relinf(X, A) < 2 * n * EPS


needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (needs the reference tree)")


@needs_ref
@pytest.mark.parametrize("n", [7, 40, 131])
def test_oracle_equals_reference_bitwise(n):
    """The unfused oracle against the reference's compiled default executors
    (execute_sphv_*_default, execute_rectdisk_*_default and the reference's
    own ft_plan_rotrectdisk through the same closed form)."""
    try:
        O.set_fused(False)
        rng = np.random.default_rng(n)
        P = O.Plan("sph", n)
        M = 2 * n - 1
        A = np.asfortranarray(rng.standard_normal((n, M)))
        for d in (O.HI2LO, O.LO2HI):
            assert np.array_equal(O.execute_sphv(P, d, A.copy(order="F")), O.ref_execute_sphv(P, d, A.copy(order="F")))
        for beta in (0.0, 0.5, -0.25):
            R = O.Plan("rectdisk", n, beta)
            B = np.asfortranarray(rng.standard_normal((n, n)))
            for d in (O.HI2LO, O.LO2HI):
                assert np.array_equal(O.execute(R, d, B.copy(order="F")), O.ref_execute_rectdisk(R, d, B.copy(order="F")))
    finally:
        O.set_fused(True)


# ------------------------------------------------------------ the library ----

@pytest.mark.parametrize("beta", BETAS)
@pytest.mark.parametrize("n", [1, 2, 3, 17, 256])
def test_product_rectdisk_plan_bitwise_equal_oracle(ft, n, beta):
    P = ft.plan_rotrectdisk(n, beta)
    Q = O.Plan("rectdisk", n, beta)
    assert P.kind == ft.RECTDISK and P.n == n
    assert np.array_equal(P.s, Q.s) and np.array_equal(P.c, Q.c)


@pytest.mark.parametrize("layout", ["sphv", "rectdisk"])
@pytest.mark.parametrize("n,M,S,maxpart", [(3, None, 2, 16), (20, None, 2, 16), (100, None, 2, 16), (100, None, 4, 16),
                                           (150, None, 2, 3), (77, 31, 2, 16), (300, None, 2, 4)])
@pytest.mark.parametrize("d", [0, 1])
def test_host_chains_for_layouts(ft, layout, n, M, S, maxpart, d):
    """Chains of the sphv and rectdisk column layouts through the kernel's
    step-by-step emulation against the oracle (relative 2-norm 1e-14: the
    emulation rounds the rotation in two steps where the kernel fuses it)."""
    if layout == "sphv":
        P, Q = ft.plan_rotsphere(n), O.Plan("sph", n)
        M = 2 * n - 1 if M is None else M
    else:
        P, Q = ft.plan_rotrectdisk(n, 0.5), O.Plan("rectdisk", n, 0.5)
        M = n if M is None else min(M, n)
    A = np.asfortranarray(np.random.default_rng(n + 7 * S + d).standard_normal((n, M)))
    bands, tab, chains = se.host_build(ft._lib, P.ptr, d, S, M, maxpart, layout)
    got = se.emulate(bands, tab, chains, S, A.copy(order="F"))
    ref = (O.execute_sphv if layout == "sphv" else O.execute)(Q, d, A.copy(order="F"), M)
    assert O.rel2(got, ref) <= 1e-14
    # rectangle-disk chain rows hold one order per column: 8 orders of one parity
    if layout == "rectdisk":
        for c in chains:
            cols = [x for x in c.col if x >= 0]
            assert len(set(x % 2 for x in cols)) <= 1 and len(set(cols)) == len(cols)


def test_invalid_widths_and_plans_rejected(ft):
    P = ft.plan_rotsphere(8)
    with pytest.raises(ft.GPUError):
        ft.execute_sphv_hi2lo(P, np.zeros((8, 16), order="F"))  # even M
    with pytest.raises(ft.GPUError):
        ft.execute_sphv_lo2hi(P, np.zeros((8, 19), order="F"))  # M > 2n+1
    R = ft.plan_rotrectdisk(8, 0.0)
    with pytest.raises(ft.GPUError):
        ft.execute_rectdisk_hi2lo(R, np.zeros((8, 9), order="F"))  # M > n
    T = ft.plan_rottriangle(8, 0.0, 0.0, 0.0)
    with pytest.raises(ValueError):
        ft.execute_rectdisk_hi2lo(T, np.zeros((8, 8), order="F"))  # triangle tables
    import ctypes
    A = np.zeros((8, 8), order="F")
    assert ft._lib.ft_gpu_execute(T.ptr, ft.RECTDISK, 0, 0, A.ctypes.data_as(ctypes.c_void_p), 8, None) == \
        ft.FT_GPU_EINVAL
    assert ft._lib.ft_gpu_execute(R.ptr, 5, 0, 0, A.ctypes.data_as(ctypes.c_void_p), 8, None) == ft.FT_GPU_EINVAL
    assert ft.last_error() == 0


def test_rotation_counts(ft):
    """utilities.rotations for the new kinds against a direct count of the
    reference's loop nests (drivers.c:204-220, 496-508; rotations_default.c:31-36)."""
    from fasttransforms_amd.utilities import rotations

    def count(n, pairs):
        return sum(max(n - 2 - j, 0) for _, m in pairs for j in range(m % 2, m - 1, 2))

    for n in (3, 10, 64, 101):
        M = 2 * n - 1
        assert rotations("sphv", n, M) == count(n, _sphv_pairs(M))
        assert rotations("sphv", n, 37 if n > 20 else 5) == count(n, _sphv_pairs(37 if n > 20 else 5))
        assert rotations("rectdisk", n, n) == count(n, _rectdisk_pairs(n))
        assert rotations("annulus", n) == rotations("disk", n)
    assert rotations("rectdisk", 1024) * 2 == rotations("sph", 1024)
