"""GPU parity of the spin-weighted sphere's executors (ft_execute_spinsph_*,
drivers.c:645-705) and kernels (ft_kernel_spinsph_*, rotations_default.c:
87-117) against the oracle: bit for bit (both stages apply every rotation in
the oracle's fused form and order on each entry).  At n = 1024, the
reference's accuracy test on the device: lo2hi undoes hi2lo within its
bounds.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _rand(n, M, seed):
    g = np.random.default_rng(seed)
    return np.asfortranarray(g.uniform(-1, 1, (n, M)) + 1j * g.uniform(-1, 1, (n, M)))


@pytest.mark.parametrize("s", list(range(-4, 5)) + [9, -70])
@pytest.mark.parametrize("n,M", [(1, 1), (2, 3), (5, 9), (20, 39), (20, 7), (67, 133), (130, 259), (256, 511)])
def test_executors_bitwise(ft, s, n, M):
    P, Q = ft.plan_rotspinsphere(n, s), O.SpinPlan(n, s)
    for d, run in ((0, ft.execute_spinsph_hi2lo), (1, ft.execute_spinsph_lo2hi)):
        A = _rand(n, M, 31 * n + M + d + 3 * (s + 80))
        _say("spinsph s=%d n=%d M=%d dir=%d" % (s, n, M, d))
        got = run(P, A.copy(order="F"), M)
        ref = O.execute_spinsph(Q, d, A.copy(order="F"), M)
        assert np.array_equal(got, ref), (s, n, M, d, O.rel2(got, ref))


@pytest.mark.parametrize("s", [-2, 0, 3])
def test_device_resident_bitwise(ft, s):
    import torch
    n, M = 300, 599
    P, Q = ft.plan_rotspinsphere(n, s), O.SpinPlan(n, s)
    A = _rand(n, M, 7 + s)
    for d in (0, 1, 0):
        _say("gpu_execute_spinsph s=%d n=%d dir=%d" % (s, n, d))
        T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        ft.gpu_execute_spinsph(P, d, T)
        ft.synchronize()
        got = T.cpu().numpy().T
        ref = O.execute_spinsph(Q, d, A.copy(order="F"))
        assert np.array_equal(got, ref), (s, d, O.rel2(got, ref))


@pytest.mark.parametrize("s", [-3, 0, 2, 9])
def test_kernels_bitwise(ft, s):
    n = 40
    P, Q = ft.plan_rotspinsphere(n, s), O.SpinPlan(n, s)
    for m in (0, 1, -1, 2, -5, 12, -39):
        for S in (1, 3):
            for d, run in ((0, ft.kernel_spinsph_hi2lo), (1, ft.kernel_spinsph_lo2hi)):
                buf = np.random.default_rng(abs(m) * 10 + S + d).uniform(-1, 1, (S * (n - 1) + 1, 2))
                v = np.ascontiguousarray(buf[:, 0] + 1j * buf[:, 1])
                _say("kernel_spinsph s=%d m=%d S=%d dir=%d" % (s, m, S, d))
                got = run(P, m, v.copy(), S)
                ref = O.kernel_spinsph(Q, d, m, v.copy(), S)
                assert np.array_equal(got, ref), (s, m, S, d)


@pytest.mark.parametrize("s", [-2, 1])
def test_round_trip_1024(ft, s):
    """test_drivers.c:1479-1500 at N = 1024 on the device: spinsphones,
    hi2lo then lo2hi, against 4 sqrt(N) eps (rel-2) and 2N eps (rel-inf)."""
    import torch
    n = 1024
    M = 2 * n - 1
    P = ft.plan_rotspinsphere(n, s)
    A = np.zeros((n, M), dtype=np.complex128, order="F")
    for i in range(n - abs(s)):
        A[i, : max(M - 2 * i, 0)] = 1 + 1j
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    _say("spinsph round trip s=%d n=1024" % s)
    ft.gpu_execute_spinsph(P, 0, T)
    ft.gpu_execute_spinsph(P, 1, T)
    ft.synchronize()
    X = T.cpu().numpy().T
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("spinsph s=%d n=1024 round trip rel2 %.2e relinf %.2e" % (s, e2, einf))
    assert e2 < 4 * np.sqrt(n) * EPS and einf < 2 * n * EPS
