"""Pins the CPU oracle (oracle/ft_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle (a wrong term
in a closed form, a wrong row or column index, a transposed rotation, a wrong
loop bound) fails it:

* the change-of-basis identities of tests/textbook.py (independent of the
  reference's code: Legendre and Jacobi quadrature),
* bit-for-bit agreement with the reference's own compiled kernels
  (oracle/_ref) and with the committed golden vectors they produced,
* s^2 + c^2 = 1 (test_rotations.c:21-23), orthogonality of the dense operator,
  the round trip, agreement with the __float128 evaluation, and an
  exact-rational check of every table entry's rounding.
"""
import hashlib
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import textbook

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = 2.0 ** -52
PLANS = [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)), ("tri", (0.0, 0.0, -0.5)),
         ("disk", (0.0, 0.0)), ("disk", (0.5, 1.0))]


def full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]


# ---------------------------------------------------------------- plans ----

@pytest.mark.parametrize("kind,params", PLANS)
def test_plan_unit_norm(kind, params):
    # test_rotations.c:21-23: sqrt(sum (hypot(s,c)-1)^2) / sqrt(#entries) < eps
    for n in (64, 128, 256):
        P = O.Plan(kind, n, *params)
        err = np.sqrt(np.sum((np.hypot(P.s, P.c) - 1.0) ** 2)) / np.sqrt(P.s.size)
        assert err < EPS


def _exact_entry(kind, params, l, m):
    """(nums, numc, den) as exact rationals of the closed forms (rotations.c:113-190)."""
    if kind == "sph":
        return Fraction((l + 1) * (l + 2)), Fraction((2 * m + 2) * (2 * l + 2 * m + 5)), \
            Fraction((l + 2 * m + 3) * (l + 2 * m + 4))
    if kind == "tri":
        a, b, g = (Fraction(x) for x in params)
        return (l + 1) * (l + a + 1), (2 * m + b + g + 2) * (2 * l + 2 * m + a + b + g + 4), \
            (l + 2 * m + b + g + 3) * (l + 2 * m + a + b + g + 3)
    a, b = (Fraction(x) for x in params)
    return (l + 1) * (l + b + 1), (m + a + 1) * (2 * l + m + a + b + 3), (l + m + a + 2) * (l + m + a + b + 2)


def _round_sqrt_of_rounded_ratio(num, den):
    """fl(sqrt(fl(num/den))) computed from exact rationals (num, den exact in FP64)."""
    q = float(num / den)  # Fraction -> float is correctly rounded
    return float(np.sqrt(q))  # IEEE sqrt is correctly rounded


@pytest.mark.parametrize("kind,params", PLANS)
def test_plan_entries_exact_rounding(kind, params):
    """Every entry is fl(sqrt(fl(num/den))) of the exact closed form (the sums and
    products of these parameters are exact in FP64), and lies within one ulp of
    the true value sqrt(num/den)."""
    n = 24
    P = O.Plan(kind, n, *params)
    orders = range(2 * n - 1) if kind == "disk" else range(n)
    for m in orders:
        rows = n - (m + 1) // 2 if kind == "disk" else n - m
        for l in range(rows):
            nums, numc, den = _exact_entry(kind, params, l, m)
            want_s = _round_sqrt_of_rounded_ratio(nums, den)
            want_c = _round_sqrt_of_rounded_ratio(numc, den)
            if kind == "disk":
                want_s = -want_s
            i = P.index(l, m)
            assert P.s[i] == want_s and P.c[i] == want_c, (m, l)
            true_s = float(np.sqrt(float(nums / den)))
            assert abs(abs(P.s[i]) - true_s) <= 2 * np.spacing(true_s)


# -------------------------------------------------------------- kernels ----

@pytest.mark.parametrize("kind,params", [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)),
                                         ("disk", (0.0, 0.0))])
@pytest.mark.parametrize("n", [10, 16])
def test_textbook_change_of_basis(kind, params, n):
    """hi2lo on unit vectors reproduces the quadrature change-of-basis matrix."""
    P = O.Plan(kind, n, *params)
    worst = 0.0
    for m in textbook.orders(kind, n):
        rows = textbook.input_rows(kind, n, m)
        C = textbook.matrix(kind, n, m, params)
        for i in range(rows):
            e = np.zeros(n)
            e[i] = 1.0
            O.kernel(P, O.HI2LO, textbook.low_order(kind, m), m, e)
            worst = max(worst, np.max(np.abs(e - C[:, i])))
    assert worst < 5e-12, worst


@pytest.mark.parametrize("kind,params", PLANS)
def test_dense_operator_orthogonal(kind, params):
    """The matrix of hi2lo for every order is orthogonal and lo2hi is its transpose."""
    n = 14
    P = O.Plan(kind, n, *params)
    for m in textbook.orders(kind, n):
        m1 = textbook.low_order(kind, m)
        H = np.zeros((n, n))
        T = np.zeros((n, n))
        for i in range(n):
            e = np.zeros(n)
            e[i] = 1.0
            H[:, i] = O.kernel(P, O.HI2LO, m1, m, e.copy())
            T[:, i] = O.kernel(P, O.LO2HI, m1, m, e.copy())
        assert np.max(np.abs(H.T @ H - np.eye(n))) < 50 * EPS
        assert np.max(np.abs(T - H.T)) < 50 * EPS


@pytest.mark.parametrize("kind,params", PLANS)
def test_roundtrip_bounds(kind, params):
    """test_drivers.c:52-57 bounds: rel-2 < 4 sqrt(N) eps (8 for tri), rel-inf < 2N eps."""
    for n in (65, 128):
        M = full_M(kind, n)
        P = O.Plan(kind, n, *params)
        A = O.fill(kind, n, M, 7)
        X = O.execute(P, O.LO2HI, O.execute(P, O.HI2LO, A.copy(order="F")))
        c2 = 8 if kind == "tri" else 4
        assert O.rel2(X, A) < c2 * np.sqrt(n) * EPS
        assert O.relinf(X, A) < 2 * n * EPS


@pytest.mark.parametrize("kind,params", PLANS)
def test_against_float128(kind, params):
    """FP64 oracle vs the same loop nest in __float128 (s, c promoted), and the
    __float128 evaluation with s, c recomputed in quadruple precision."""
    n = 96
    M = full_M(kind, n)
    P = O.Plan(kind, n, *params)
    A = O.fill(kind, n, M, 11)
    for d in (O.HI2LO, O.LO2HI):
        X = O.execute(P, d, A.copy(order="F"))
        Q = O.execute_q(P, d, A.copy(order="F"))
        H = O.execute_q(P, d, A.copy(order="F"), hp_plan=True)
        assert O.rel2(X, Q) < 4 * np.sqrt(n) * EPS
        assert O.rel2(X, H) < 8 * np.sqrt(n) * EPS


def test_columns_outside_path_untouched():
    """Orders 0, 1 (sph/disk) and 0 (tri) are never touched (drivers.c:94-110)."""
    for kind, params in (("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("disk", (0.0, 0.0))):
        n = 20
        M = full_M(kind, n)
        A = O.fill(kind, n, M, 3)
        X = O.execute(O.Plan(kind, n, *params), O.HI2LO, A.copy(order="F"))
        keep = 1 if kind == "tri" else 3
        assert np.array_equal(X[:, :keep], A[:, :keep])


# ----------------------------------------------- the reference's own code ---

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (needs the reference tree)")


@needs_ref
@pytest.mark.parametrize("kind,params", PLANS)
@pytest.mark.parametrize("n", [7, 37, 130])
def test_unfused_oracle_equals_reference_default_bitwise(kind, params, n):
    """With the arithmetic of the reference's scalar build, the oracle's loop
    nests reproduce the reference's compiled default executor bit for bit."""
    M = full_M(kind, n)
    P = O.Plan(kind, n, *params)
    A = O.fill(kind, n, M, 5)
    try:
        O.set_fused(False)
        for d in (O.HI2LO, O.LO2HI):
            X = O.execute(P, d, A.copy(order="F"))
            R = O.ref_execute(P, d, A.copy(order="F"), isa=0)
            assert np.array_equal(X, R)
    finally:
        O.set_fused(True)


@needs_ref
@pytest.mark.parametrize("kind,params", PLANS)
def test_fused_oracle_vs_reference_dispatch(kind, params):
    """The fused oracle vs the reference's dispatched SIMD executor (which mixes
    fused SIMD tiles with unfused scalar heads): rounding-level differences."""
    for n in (64, 257):
        M = full_M(kind, n)
        P = O.Plan(kind, n, *params)
        A = O.fill(kind, n, M, 9)
        for d in (O.HI2LO, O.LO2HI):
            X = O.execute(P, d, A.copy(order="F"))
            R = O.ref_execute(P, d, A.copy(order="F"))
            assert O.rel2(X, R) < 2 * np.sqrt(n) * EPS


@needs_ref
def test_single_column_kernels_equal_reference():
    """ft_kernel_* (rotations.c:135-141, 162-168, 262-268) incl. strides and
    m1 of either parity, against the reference's compiled default kernels."""
    try:
        O.set_fused(False)
        for kind, params in (("sph", ()), ("tri", (0.0, 0.0, -0.5)), ("disk", (0.0, 0.0))):
            n = 40
            P = O.Plan(kind, n, *params)
            for S in (1, 3):
                for m1, m2 in ((0, 8), (1, 9), (2, 17), (1, 2), (3, 30), (0, 0)):
                    base = np.arange((n - 1) * S + 1, dtype=np.float64) / 7.0 - 1.3
                    for d in (O.HI2LO, O.LO2HI):
                        X = O.kernel(P, d, m1, m2, base.copy(), S)
                        R = O.ref_kernel(P, d, m1, m2, base.copy(), S)
                        assert np.array_equal(X, R), (kind, S, m1, m2, d)
    finally:
        O.set_fused(True)


# -------------------------------------------------------- golden vectors ----

def _golden():
    with open(os.path.join(HERE, "golden", "manifest.json")) as f:
        man = json.load(f)
    data = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))
    return man, data


def test_golden_vectors_match_oracle():
    """The committed vectors (the reference's compiled executors) vs the oracle:
    bitwise for the scalar executor (oracle in unfused mode), rounding-level for
    the SIMD executor (oracle in fused mode)."""
    man, data = _golden()
    for key, case in man["cases"].items():
        kind, params, n, M = case["kind"], tuple(case["params"]), case["n"], case["M"]
        A = O.fill(kind, n, M, case["seed"])
        assert hashlib.sha256(A.tobytes(order="F")).hexdigest() == case["input_sha256"], key
        P = O.Plan(kind, n, *params)
        for d, dname in ((O.HI2LO, "hi2lo"), (O.LO2HI, "lo2hi")):
            ref_default = data["%s_%s_default" % (key, dname)].T
            ref_simd = data["%s_%s_simd" % (key, dname)].T
            try:
                O.set_fused(False)
                X = O.execute(P, d, A.copy(order="F"))
            finally:
                O.set_fused(True)
            assert np.array_equal(X, ref_default), (key, dname)
            Y = O.execute(P, d, A.copy(order="F"))
            assert O.rel2(Y, ref_simd) < 2 * np.sqrt(n) * EPS, (key, dname)


# ---------------------------------------------------------------- fills -----

def test_fill_stream_matches_numpy_restatement(ft):
    """oracle.fill (C) and the package's seeded fill (numpy) are the same stream."""
    from fasttransforms_amd import utilities
    for kind, n in (("sph", 9), ("tri", 11), ("disk", 6)):
        M = full_M(kind, n)
        assert np.array_equal(O.fill(kind, n, M, 1234), utilities.fill(kind, n, M, 1234))


@needs_ref
@pytest.mark.parametrize("kind,params,n", [("sph", (), 1024), ("sph", (), 2048), ("tri", (0.0, 0.0, 0.0), 1024),
                                           ("tri", (-0.5, -0.5, -0.5), 1024), ("disk", (0.0, 0.0), 1024)])
def test_fused_oracle_vs_reference_dispatch_at_scale(kind, params, n):
    """The headline sizes: the fused oracle (which the GPU matches bit for
    bit) against the reference's dispatched executor, both directions, within
    the 1e-14 relative 2-norm bar (measured here ~1e-15)."""
    M = full_M(kind, n)
    P = O.Plan(kind, n, *params)
    A = O.fill(kind, n, M, 1000 + n)
    for d in (O.HI2LO, O.LO2HI):
        X = O.execute(P, d, A.copy(order="F"))
        R = O.ref_execute(P, d, A.copy(order="F"))
        err = O.rel2(X, R)
        print("%s%s n=%d dir=%d rel2 vs reference dispatch %.3e" % (kind, params, n, d, err))
        assert err <= 1e-14


# ------------------------------------------------------------------ FP32 ----

def _f32_exact(q):
    """Correct rounding (to nearest, ties to even) of a Fraction to binary32,
    returned as a Fraction (normal range only: every value here is)."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    elif Fraction(2) ** (e + 1) <= a:
        e += 1
    scale = Fraction(2) ** (23 - e)
    m = a * scale
    mi = m.numerator // m.denominator
    r = m - mi
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and mi % 2 == 1):
        mi += 1
    return sign * Fraction(mi) / scale


@pytest.mark.parametrize("kind,params", [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("disk", (0.0, 0.0))])
def test_fp32_rotation_arithmetic_exact(kind, params):
    """One FP32 rotation of the oracle's float loop nest against exact
    rational arithmetic with correct binary32 rounding: s and c rounded once
    from the FP64 plan, the product s*y rounded, then c*x + (s*y) rounded once
    (the fused form), and the transposed rotation for lo2hi.  An unfused
    evaluation, a product kept in double, or unrounded s/c would differ in
    some of these cases."""
    n = 3 if kind == "sph" else 2
    m2 = 2 if kind != "tri" else 1
    P = O.Plan(kind, 6, *params)  # a table with several distinct entries
    rng = np.random.default_rng(5)
    lo_row, hi_row = 0, (2 if kind == "sph" else 1)
    mism = 0
    for trial in range(400):
        # take a random table entry by building a plan view whose only rotation is (0, j)
        x0, y0 = (np.float32(v) for v in rng.uniform(-1, 1, 2))
        Q = O.Plan(kind, n, *params)
        k = int(rng.integers(0, len(P.s)))
        Q.s[:] = P.s[k]
        Q.c[:] = P.c[k]
        S, C = _f32_exact(Fraction(float(P.s[k]))), _f32_exact(Fraction(float(P.c[k])))
        x, y = Fraction(float(x0)), Fraction(float(y0))
        for d in (O.HI2LO, O.LO2HI):
            buf = np.zeros(n, dtype=np.float32)
            buf[lo_row], buf[hi_row] = x0, y0
            O.kernel_f32(Q, d, 0, m2, buf)
            if d == O.HI2LO:
                ex = _f32_exact(C * x + _f32_exact(S * y))
                ey = _f32_exact(C * y - _f32_exact(S * x))
            else:
                ex = _f32_exact(C * x - _f32_exact(S * y))
                ey = _f32_exact(C * y + _f32_exact(S * x))
            got = (Fraction(float(buf[lo_row])), Fraction(float(buf[hi_row])))
            mism += got != (ex, ey)
    assert mism == 0


@pytest.mark.parametrize("kind,params", PLANS)
def test_fp32_oracle_close_to_fp64(kind, params):
    """The FP32 loop nests are the FP64 oracle's (one macro text, two
    instantiations): the FP32 executor stays within an FP32 rounding bound of
    the FP64 one in both directions, really computes in float (it is not the
    FP64 result rounded), and its round trip returns the input within the
    FP32 analogue of the reference's bound."""
    for n in (33, 200):
        M = full_M(kind, n)
        P = O.Plan(kind, n, *params)
        A = O.fill(kind, n, M, 61)
        for d in (O.HI2LO, O.LO2HI):
            R = O.execute(P, d, A.copy(order="F"))
            F = O.execute_f32(P, d, A.astype(np.float32, order="F"))
            err = O.rel2(F.astype(np.float64), R)
            assert 2.0 ** -26 < err < 8 * np.sqrt(n) * 2.0 ** -23, (n, d, err)
            assert not np.array_equal(F, R.astype(np.float32))
        F = O.execute_f32(P, O.HI2LO, A.astype(np.float32, order="F"))
        F = O.execute_f32(P, O.LO2HI, F)
        assert O.rel2(F.astype(np.float64), A.astype(np.float32).astype(np.float64)) < 4 * np.sqrt(n) * 2.0 ** -23
