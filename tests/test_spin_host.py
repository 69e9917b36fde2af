"""The spin-weighted sphere's executors (ft_execute_spinsph_*, drivers.c:645-705)
on the CPU.

The oracle side is pinned first:

* the tables (rotations.c:443-475): s1, c1 are the sphere's table bit for
  bit; s2(., ., d), c2(., ., d) are the triangle table of parameters
  (alpha, beta, gamma) = (d, d-1, 0) bit for bit (the closed forms coincide
  term by term), every entry fl(sqrt(fl(num/den))) of exact integers with
  nums + numc = den (so s^2 + c^2 = 1);
* unfused, bit-identical to the reference's compiled scalar executor
  (oracle/_ref: kernel_spinsph_*_default under drivers.c:687-705's glue) for
  spins -4 .. 4, both directions;
* fused, bit-identical to the reference's dispatched executor
  (kernel_spinsph_*_AVX_FMA, its twisted fused rotations, on the permuted
  column pairs) on the columns of orders |m| >= 2, which that executor runs
  fused, and to the unfused oracle on orders 0, +-1, which it runs scalar;
* lo2hi undoes hi2lo within the reference's bounds (test_drivers.c:1492-1500:
  4 sqrt(N) eps rel-2, 2N eps rel-inf) on its spinsphones input.

Then the library's host side: its tables equal the oracle's bit for bit, and
the band-systolic chains of the second stage, run through the kernel's
emulation, are the sphere's sequences of order |m| - |s| on the real and
imaginary parts, with every address inside the array.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import sys_emulator as se

EPS = 2.0 ** -52
SPINS = list(range(-4, 5))


def _rand(n, M, seed):
    g = np.random.default_rng(seed)
    return np.asfortranarray(g.uniform(-1, 1, (n, M)) + 1j * g.uniform(-1, 1, (n, M)))


def _ones(n, M, s):
    """spinsphones (ftutilities.c:149-157)."""
    A = np.zeros((n, M), dtype=np.complex128, order="F")
    for i in range(n - abs(s)):
        A[i, : max(M - 2 * i, 0)] = 1 + 1j
    return A


@pytest.mark.parametrize("s", [0, 1, -2, 4])
def test_tables_exact_and_identities(s):
    n = 24
    P = O.SpinPlan(n, s)
    sph = O.Plan("sph", n)
    for m in range(n):
        for l in range(n - m):
            assert P.s1[l + m * n] == sph.s[sph.index(l, m)] and P.c1[l + m * n] == sph.c[sph.index(l, m)]
    a = abs(s)
    for d in range(n):
        T = O.Plan("tri", n, float(d), float(d - 1), 0.0) if a else None
        for j in range(a):
            for l in range(n - d - j):
                nums = (l + 1) * (l + d + 1)
                numc = (2 * j + d + 1) * (2 * l + 2 * j + 2 * d + 3)
                den = (l + 2 * j + d + 2) * (l + 2 * j + 2 * d + 2)
                assert nums + numc == den
                i = l + j * n + d * a * n
                assert P.s2[i] == float(np.sqrt(float(Fraction(nums, den))))
                assert P.c2[i] == float(np.sqrt(float(Fraction(numc, den))))
                assert P.s2[i] == T.s[T.index(l, j)] and P.c2[i] == T.c[T.index(l, j)]


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref is not built")
@pytest.mark.parametrize("s", SPINS)
@pytest.mark.parametrize("n,M", [(1, 1), (2, 3), (7, 13), (20, 39), (20, 7), (67, 133)])
@pytest.mark.parametrize("d", [0, 1])
def test_oracle_matches_reference(s, n, M, d):
    P = O.SpinPlan(n, s)
    A = _rand(n, M, 1000 * n + 10 * M + d + (s + 4) * 7)
    O.set_fused(False)
    try:
        unfused = O.execute_spinsph(P, d, A.copy(order="F"))
    finally:
        O.set_fused(True)
    ref = O.ref_execute_spinsph(P, d, A.copy(order="F"), isa="default")
    assert np.array_equal(unfused, ref)
    fused = O.execute_spinsph(P, d, A.copy(order="F"))
    disp = O.ref_execute_spinsph(P, d, A.copy(order="F"), isa="avx_fma")
    k = min(3, M)
    assert np.array_equal(disp[:, :k], unfused[:, :k])
    assert np.array_equal(disp[:, k:], fused[:, k:])


@pytest.mark.parametrize("s", SPINS)
def test_round_trip(s):
    n = 64
    M = 2 * n - 1
    P = O.SpinPlan(n, s)
    A = _ones(n, M, s)
    B = O.execute_spinsph(P, 0, A.copy(order="F"))
    assert abs(np.linalg.norm(B) - np.linalg.norm(A)) <= 4 * np.sqrt(n) * EPS * np.linalg.norm(A)
    X = O.execute_spinsph(P, 1, B)
    assert O.rel2(X, A) < 4 * np.sqrt(n) * EPS
    assert O.relinf(X, A) < 2 * n * EPS


def test_kernel_is_executor_column():
    """ft_kernel_spinsph on a strided vector equals the executor's column of
    that order (drivers.c:687-705 calls it on each column)."""
    n, s = 15, -3
    P = O.SpinPlan(n, s)
    A = _rand(n, 9, 5)
    full = O.execute_spinsph(P, 0, A.copy(order="F"))
    for c, m in [(0, 0), (3, -2), (4, 2), (7, -4), (8, 4)]:
        buf = np.zeros(3 * (n - 1) + 1, dtype=np.complex128)
        buf[::3] = A[:, c]
        O.kernel_spinsph(P, 0, m, buf, 3)
        assert np.array_equal(buf[::3], full[:, c]), (c, m)


def test_library_tables_equal_oracle(ft):
    for n, s in [(1, 0), (9, 2), (33, -4), (50, 7)]:
        P, Q = ft.plan_rotspinsphere(n, s), O.SpinPlan(n, s)
        a = abs(s)
        assert np.array_equal(P.s1, Q.s1) and np.array_equal(P.c1, Q.c1)
        assert np.array_equal(P.s2, Q.s2[: a * n * n]) and np.array_equal(P.c2, Q.c2[: a * n * n])


@pytest.mark.parametrize("s", [0, 1, -2, 3])
@pytest.mark.parametrize("n,M", [(3, 5), (20, 39), (40, 23), (90, 179)])
@pytest.mark.parametrize("S,maxpart", [(2, 16), (2, 2), (4, 16)])
@pytest.mark.parametrize("d", [0, 1])
def test_host_chains_of_second_stage(ft, s, n, M, S, maxpart, d):
    """The second stage's chains through the kernel's emulation equal the
    sphere's kernel of orders d%2 .. d (d = |am - as|) on the real and the
    imaginary part of each column (relative 1e-14: the emulation rounds the
    rotation in two steps)."""
    R = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = _rand(n, M, 17 * n + M + s)
    bands, tab, chains = se.host_build_spin(ft._lib, R.ptr, d, S, s, M, maxpart)
    size = 2 * n * M
    for ch in chains:
        top = max(bands[ch.band0 + w].R for w in range(ch.nb))
        for c in ch.col:
            if c >= 0:
                assert 0 <= ch.aoff + c + ch.rpar and ch.aoff + c + ch.rpar + ch.rstep * top < size
    flat = np.ascontiguousarray(A.reshape(-1, order="F")).view(np.float64).copy()
    se.emulate(bands, tab, chains, S, flat)
    got = flat.view(np.complex128).reshape((n, M), order="F")
    ref = A.copy(order="F")
    for c in range(M):
        m = 0 if c == 0 else (-(c + 1) // 2 if c % 2 else c // 2)
        dd = abs(abs(m) - abs(s))
        for part in (ref.real, ref.imag):
            v = np.ascontiguousarray(part[:, c])
            O.kernel(Q, d, dd % 2, dd, v)
            part[:, c] = v
    assert O.rel2(got, ref) <= 1e-14 or np.linalg.norm(ref) == 0


def test_invalid_calls_rejected(ft):
    P = ft.plan_rotspinsphere(8, 2)
    A = np.zeros((8, 16), dtype=np.complex128, order="F")
    for M in (2, 16, -1):
        with pytest.raises(ft.GPUError):
            ft.execute_spinsph_hi2lo(P, A, M)
    with pytest.raises(ft.GPUError):
        ft.kernel_spinsph_lo2hi(P, 1, A[:, 0].copy(), 0)
