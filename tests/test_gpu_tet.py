"""GPU parity of the tetrahedron's executors (ft_execute_tet_*, drivers.c:510-533)
and kernels (ft_kernel_tet_*, rotations.c:297-321) against the oracle:
bit for bit, since the two band-systolic launches apply every rotation in
the oracle's fused form and in its order on each entry.  At N = L = M = 256
(where the oracle would take minutes) through size-independent properties:
lo2hi undoes hi2lo within the reference's bounds and hi2lo keeps the norm.
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


def _plans(ft, N, p):
    a, b, g, d = p
    return (ft.plan_rottriangle(N, a, b, g + d + 1.0), ft.plan_rottriangle(N, b, g, d),
            O.Plan("tri", N, a, b, g + d + 1.0), O.Plan("tri", N, b, g, d))


SHAPES = [(1, 1, 1), (2, 2, 2), (5, 5, 5), (9, 4, 9), (12, 7, 15), (17, 17, 17), (20, 1, 4), (33, 33, 33),
          (64, 40, 70), (100, 100, 37), (128, 128, 128)]


@pytest.mark.parametrize("N,L,M", SHAPES)
@pytest.mark.parametrize("p", [(0.0, 0.0, 0.0, 0.0), (0.1, 0.2, 0.3, 0.4), (1.5, -0.25, 0.75, 2.0)])
def test_executors_bitwise(ft, N, L, M, p):
    P1, P2, Q1, Q2 = _plans(ft, N, p)
    for d, run in ((0, ft.execute_tet_hi2lo), (1, ft.execute_tet_lo2hi)):
        A = np.asfortranarray(np.random.default_rng(N + L + M + d).uniform(-1, 1, (N, L, M)))
        _say("tet", p, "N=%d L=%d M=%d dir=%d" % (N, L, M, d))
        got = run(P1, P2, A.copy(order="F"), L, M)
        ref = O.execute_tet(Q1, Q2, d, A.copy(order="F"), L, M)
        assert np.array_equal(got, ref), (N, L, M, d, O.rel2(got, ref))


@pytest.mark.parametrize("N,L,M", [(9, 4, 9), (64, 64, 64), (100, 77, 120)])
def test_device_resident_bitwise(ft, N, L, M):
    """ft_gpu_execute_tet on a device tensor (asynchronous, then synchronize),
    repeated so the cached schedules run again."""
    import torch
    p = (0.1, 0.2, 0.3, 0.4)
    P1, P2, Q1, Q2 = _plans(ft, N, p)
    A = np.asfortranarray(np.random.default_rng(3 * N + L).uniform(-1, 1, (N, L, M)))
    for d in (0, 1, 0):
        _say("gpu_execute_tet N=%d L=%d M=%d dir=%d" % (N, L, M, d))
        T = torch.from_numpy(A.reshape(-1, order="F").copy()).cuda()
        ft.gpu_execute_tet(P1, P2, d, T, L, M)
        ft.synchronize()
        got = T.cpu().numpy().reshape((N, L, M), order="F")
        ref = O.execute_tet(Q1, Q2, d, A.copy(order="F"), L, M)
        assert np.array_equal(got, ref), (d, O.rel2(got, ref))


@pytest.mark.parametrize("N,L", [(1, 1), (7, 7), (16, 9), (64, 64)])
def test_slab_kernels_bitwise(ft, N, L):
    P = ft.plan_rottriangle(N, 0.2, 0.3, 0.4)
    Q = O.Plan("tri", N, 0.2, 0.3, 0.4)
    for m in sorted({0, 1, L // 2, L - 1, L + 3}):
        for d, run in ((0, ft.kernel_tet_hi2lo), (1, ft.kernel_tet_lo2hi)):
            A = np.asfortranarray(np.random.default_rng(N * m + d).uniform(-1, 1, (N, L)))
            _say("kernel_tet N=%d L=%d m=%d dir=%d" % (N, L, m, d))
            got = run(P, L, m, A.copy(order="F"))
            ref = O.kernel_tet(Q, d, L, m, A.copy(order="F"))
            assert np.array_equal(got, ref), (N, L, m, d)


def test_round_trip_256(ft):
    """The reference's accuracy test (test_drivers.c:1283-1305) at N = L = M =
    256 on the device: tetones, hi2lo then lo2hi, against 8 (NLM)^(1/3) eps
    (2-norm) and 4 (NLM)^(2/3) eps (max norm); hi2lo keeps the 2-norm."""
    import torch
    N = L = M = 256
    P1, P2, _, _ = _plans(ft, N, (0.0, 0.0, 0.0, 0.0))
    A = np.zeros((N, L, M), order="F")
    for m in range(M):
        for l in range(L - m):
            A[: N - l - m, l, m] = 1.0
    T = torch.from_numpy(A.reshape(-1, order="F").copy()).cuda()
    _say("tet round trip N=L=M=256")
    ft.gpu_execute_tet(P1, P2, 0, T, L, M)
    ft.synchronize()
    nB = float(torch.linalg.norm(T))
    ft.gpu_execute_tet(P1, P2, 1, T, L, M)
    ft.synchronize()
    X = T.cpu().numpy().reshape((N, L, M), order="F")
    nA = np.linalg.norm(A)
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("tet 256 round trip rel2 %.2e relinf %.2e, norm change %.2e" % (e2, einf, abs(nB - nA) / nA))
    cube = (N * L * M) ** (1 / 3)
    assert e2 < 8 * cube * EPS and einf < 4 * cube ** 2 * EPS
    assert abs(nB - nA) <= 8 * cube * EPS * nA
