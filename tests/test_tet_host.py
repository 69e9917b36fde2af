"""The tetrahedron's executors (ft_execute_tet_*, drivers.c:510-533) on the CPU.

The oracle side is pinned first:

* bit for bit against the reference's own compiled triangle kernels
  (oracle/_ref, unfused as the reference's scalar executor computes): the
  slab columns exactly as drivers.c runs them, the block rotations
  (rotations.c:297-321) as that kernel on each slab row at stride n with the
  table of the size-L plan -- a different loop nest over a different table;
* the block rotations at L = n are the oracle's separately pinned triangle
  kernel on every row of the slab at stride n;
* the triangle plan's entries do not depend on its size (the identity the
  size-L table relies on), in the oracle and in the library;
* lo2hi undoes hi2lo and hi2lo keeps the 2-norm, within the reference's
  test bounds (test_drivers.c:1296-1305).

Then the library's host side: the band-systolic chains of both stages, run
through the step-by-step emulation of the kernel, reproduce the oracle, and
every address they reach lies inside the array.
"""
import numpy as np
import pytest

import oracle as O
import sys_emulator as se

EPS = 2.0 ** -52
# the reference's test parameters (test_drivers.c:15-20 use alpha, beta,
# gamma, delta with these defaults and others)
PARAMS = [(0.0, 0.0, 0.0, 0.0), (0.1, 0.2, 0.3, 0.4), (-0.5, -0.5, -0.5, -0.5), (1.5, -0.25, 0.75, 2.0)]
SHAPES = [(1, 1, 1), (2, 2, 2), (5, 5, 5), (9, 9, 9), (9, 4, 9), (9, 9, 3), (12, 7, 15), (17, 17, 17), (20, 1, 4),
          (16, 13, 2)]


def _plans(N, p):
    a, b, g, d = p
    return O.Plan("tri", N, a, b, g + d + 1.0), O.Plan("tri", N, b, g, d)


def _array(N, L, M, seed):
    return np.asfortranarray(np.random.default_rng(seed).uniform(-1, 1, (N, L, M)))


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref is not built")
@pytest.mark.parametrize("p", PARAMS)
@pytest.mark.parametrize("N,L,M", SHAPES)
@pytest.mark.parametrize("d", [0, 1])
def test_oracle_matches_reference_kernels(p, N, L, M, d):
    P1, P2 = _plans(N, p)
    PL = O.Plan("tri", L, p[1], p[2], p[3])
    A = _array(N, L, M, N * 100 + L * 10 + M + d)
    O.set_fused(False)
    try:
        got = O.execute_tet(P1, P2, d, A.copy(order="F"), L, M)
    finally:
        O.set_fused(True)
    ref = O.ref_execute_tet(P1, PL, d, A.copy(order="F"), L, M)
    assert np.array_equal(got, ref), (p, N, L, M, d)
    # the fused form differs from the unfused by rounding only
    fused = O.execute_tet(P1, P2, d, A.copy(order="F"), L, M)
    assert O.rel2(fused, ref) <= 4 * EPS * np.sqrt(N * L * M) * N


@pytest.mark.parametrize("N,m", [(1, 0), (2, 1), (7, 3), (16, 15), (16, 20)])
@pytest.mark.parametrize("d", [0, 1])
def test_block_rotations_are_the_triangle_kernel_at_stride(N, m, d):
    """At L = n, rotations.c:297-321 on the slab is the triangle kernel of
    orders m-1 .. 0 on each row at stride n: rows are independent, so the
    order of its loops changes no bits."""
    P = O.Plan("tri", N, 0.3, -0.2, 0.6)
    A = _array(N, N, 1, 7 * N + m)
    got = O.kernel_tet(P, d, N, m, A.copy(order="F"))
    flat = A.copy(order="F").reshape(-1, order="F")
    for k in range(N):
        view = flat[k:]
        O.kernel(P, d, 0, min(m, N), view, S=N)
    assert np.array_equal(got.reshape(-1, order="F"), flat)


def test_triangle_entries_do_not_depend_on_size(ft):
    """s(l, j), c(l, j) of the triangle plan (rotations.c:143-160) read only
    l, j and the parameters: the size-L table is the size-n one cut to
    l + j <= L - 1, in the oracle and in the library."""
    for a, b, g in [(0.0, 0.0, 0.0), (0.2, 0.3, 1.4), (-0.5, 0.5, -0.25)]:
        big, bigp = O.Plan("tri", 40, a, b, g), ft.plan_rottriangle(40, a, b, g)
        for L in (1, 2, 13, 40):
            small, smallp = O.Plan("tri", L, a, b, g), ft.plan_rottriangle(L, a, b, g)
            for j in range(L):
                for l in range(L - j):
                    assert small.s[small.index(l, j)] == big.s[big.index(l, j)]
                    assert small.c[small.index(l, j)] == big.c[big.index(l, j)]
            assert np.array_equal(smallp.s, small.s) and np.array_equal(smallp.c, small.c)
        assert np.array_equal(bigp.s, big.s) and np.array_equal(bigp.c, big.c)


@pytest.mark.parametrize("N", [16, 33, 64])
@pytest.mark.parametrize("p", PARAMS[:2])
def test_round_trip_and_norm(N, p):
    """hi2lo then lo2hi within the reference's bounds (test_drivers.c:1299-1305:
    8 (NLM)^(1/3) eps in the 2-norm, 4 (NLM)^(2/3) eps in the max norm) on its
    tetones input, and hi2lo keeps the 2-norm."""
    L = M = N
    P1, P2 = _plans(N, p)
    A = np.zeros((N, L, M), order="F")
    for m in range(M):
        for l in range(L - m):
            A[: N - l - m, l, m] = 1.0
    B = O.execute_tet(P1, P2, 0, A.copy(order="F"), L, M)
    assert abs(np.linalg.norm(B) - np.linalg.norm(A)) <= 8 * (N * L * M) ** (1 / 3) * EPS * np.linalg.norm(A)
    X = O.execute_tet(P1, P2, 1, B, L, M)
    assert O.rel2(X, A) < 8 * (N * L * M) ** (1 / 3) * EPS
    assert O.relinf(X, A) < 4 * (N * L * M) ** (2 / 3) * EPS


def _addresses_in_bounds(bands, chains, size):
    for ch in chains:
        top = max(bands[ch.band0 + w].R for w in range(ch.nb))
        for c in ch.col:
            if c < 0:
                continue
            lo = ch.aoff + c * ch.lda + ch.rpar
            hi = lo + ch.rstep * top
            assert 0 <= lo and hi < size, (lo, hi, size)


@pytest.mark.parametrize("N,L,M", [(1, 1, 1), (2, 2, 3), (9, 9, 9), (9, 4, 9), (12, 7, 15), (17, 17, 17),
                                   (20, 3, 2), (40, 33, 12)])
@pytest.mark.parametrize("S,maxpart", [(2, 16), (4, 16), (2, 1)])
@pytest.mark.parametrize("d", [0, 1])
def test_host_chains_of_both_stages(ft, N, L, M, S, maxpart, d):
    """The chains of the slab columns (on the size-N plan 1) and of the slab
    rows (on the size-L plan with plan 2's parameters, as the library's
    executor takes it) in the executor's order through the kernel's
    emulation, against the oracle (relative 2-norm 1e-14: the emulation rounds
    the rotation in two steps where the kernel fuses it)."""
    p = PARAMS[1]
    a, b, g, dl = p
    P1 = ft.plan_rottriangle(N, a, b, g + dl + 1.0)
    PB = ft.plan_rottriangle(L, b, g, dl)
    Q1, Q2 = _plans(N, p)
    A = _array(N, L, M, 31 * N + L + M + d)
    got = A.copy(order="F")
    for part in ((0, 1) if d == 0 else (1, 0)):
        plan = P1 if part == 0 else PB
        bands, tab, chains = se.host_build_tet(ft._lib, plan.ptr, d, S, part, N, L, 0, M, maxpart)
        _addresses_in_bounds(bands, chains, N * L * M)
        se.emulate(bands, tab, chains, S, got)
    ref = O.execute_tet(Q1, Q2, d, A.copy(order="F"), L, M)
    assert O.rel2(got, ref) <= 1e-14 or np.linalg.norm(ref) == 0


@pytest.mark.parametrize("m0,count", [(0, 1), (3, 1), (5, 2)])
def test_host_chains_of_slab_kernels(ft, m0, count):
    """The block rotations of slabs from order m0 (the kernels' schedule)
    against rotations.c:297-321 restated, slab by slab."""
    N, L, S = 11, 8, 2
    P = ft.plan_rottriangle(L, 0.2, 0.3, 0.4)
    Q = O.Plan("tri", N, 0.2, 0.3, 0.4)
    for d in (0, 1):
        A = _array(N, L, count, 5 * m0 + d)
        got = A.copy(order="F")
        bands, tab, chains = se.host_build_tet(ft._lib, P.ptr, d, S, 1, N, L, m0, count)
        _addresses_in_bounds(bands, chains, N * L * count)
        se.emulate(bands, tab, chains, S, got)
        for s in range(count):
            ref = O.kernel_tet(Q, d, L, m0 + s, np.asfortranarray(A[:, :, s]))
            assert O.rel2(got[:, :, s], ref) <= 1e-14


def test_invalid_shapes_and_plans_rejected(ft):
    """The library checks what the reference leaves unchecked: triangle plans
    of one size, 0 <= L <= n, M >= 0 (a larger L would read the block
    rotations beyond plan 2's table)."""
    n = 6
    P1, P2 = ft.plan_rottriangle(n, 0, 0, 1), ft.plan_rottriangle(n, 0, 0, 0)
    A = np.zeros((n, n + 1, 3), order="F")
    with pytest.raises(ft.GPUError):
        ft.execute_tet_hi2lo(P1, P2, A, n + 1, 3)
    with pytest.raises(ft.GPUError):
        ft.execute_tet_lo2hi(P1, ft.plan_rottriangle(n + 1, 0, 0, 0), A, n, 3)
    with pytest.raises(ft.GPUError):
        ft.execute_tet_hi2lo(P1, P2, A, n, -1)
    with pytest.raises(ValueError):
        ft.execute_tet_hi2lo(ft.plan_rotsphere(n), P2, A, n, 3)
    with pytest.raises(ft.GPUError):
        ft.kernel_tet_hi2lo(P2, n + 1, 2, A)


def test_rotation_count():
    """utilities.tet_rotations (the benchmark's flop count) against the loop
    bounds of drivers.c:510-533 and rotations.c:297-321 counted one by one."""
    from fasttransforms_amd.utilities import tet_rotations
    for N, L, M in [(1, 1, 1), (5, 5, 5), (9, 4, 12), (12, 12, 3), (20, 17, 25)]:
        count = 0
        for m in range(M):
            for l in range(max(L - m, 0)):
                for j in range(l + m - 1, -1, -1):
                    count += max(N - 1 - j, 0)
            for j in range(m - 1, -1, -1):
                count += max(L - 1 - j, 0) * N
        assert tet_rotations(N, L, M) == count, (N, L, M)
