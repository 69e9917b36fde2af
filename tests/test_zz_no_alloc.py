"""ft_gpu_execute_noalloc (the allocation-free entry point) against
ft_gpu_execute, through the C ABI: bit-identical results on two calls in a
row, the second on a smaller input, in both directions and precisions; the
preconditions and limits are rejected with FT_GPU_EINVAL."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _noalloc(ft, P, direction, T, precision, stream, kind=None):
    return ft._lib.ft_gpu_execute_noalloc(P.ptr, P.kind if kind is None else kind, int(direction), int(precision),
                                          ctypes.c_void_p(T.data_ptr()), int(T.shape[0]),
                                          ctypes.c_void_p(stream.cuda_stream))


def _case(ft, n, direction, precision, seed):
    """The same input through ft_gpu_execute and ft_gpu_execute_noalloc."""
    import torch
    P = ft.plan_rotsphere(n)
    dt = torch.float64 if precision == ft.FP64 else torch.float32
    A = O.fill("sph", n, 2 * n - 1, seed)
    X = torch.from_numpy(np.ascontiguousarray(A.T)).to(dt).cuda()  # (M, n): row j = column j
    Y = X.clone()
    stream = torch.cuda.current_stream()
    # the allocating entry point first: it also leaves the plan's tables resident
    ft.set_engine(ft.ENGINE_AUTO)
    ft.gpu_execute(P, direction, X, precision=precision, stream=stream)
    ft.synchronize(stream)
    assert ft.last_engine() == ft.ENGINE_SYSTOLIC
    _say("noalloc n=%d dir=%d precision=%d" % (n, direction, precision))
    rc = _noalloc(ft, P, direction, Y, precision, stream)
    assert rc == 0, rc
    ft.synchronize(stream)
    assert torch.equal(X, Y), "allocation-free result differs from ft_gpu_execute"
    return P


def test_noalloc_matches_execute_two_calls_second_smaller(ft):
    _case(ft, 300, 0, ft.FP64, 71)
    _case(ft, 150, 0, ft.FP64, 72)


def test_noalloc_directions_and_precisions(ft):
    for n, d, p in ((257, 1, ft.FP64), (200, 0, ft.FP32), (120, 1, ft.FP32), (1024, 0, ft.FP64)):
        _case(ft, n, d, p, 80 + n)


def test_noalloc_rejects_what_it_cannot_run(ft):
    import torch
    stream = torch.cuda.current_stream()
    # tables not resident yet: rejected, nothing launched
    P = ft.plan_rotsphere(90)
    T = torch.zeros(2 * 90 - 1, 90, dtype=torch.float64, device="cuda")
    assert _noalloc(ft, P, 0, T, ft.FP64, stream) == ft.FT_GPU_EINVAL
    # a work list beyond the scratch (sph n = 4096: over 5000 chain parts)
    Q = ft.plan_rotsphere(4096)
    U = torch.zeros(2 * 4096 - 1, 4096, dtype=torch.float64, device="cuda")
    ft.gpu_execute(Q, 0, U)
    ft.synchronize(stream)
    assert _noalloc(ft, Q, 0, U, ft.FP64, stream) == ft.FT_GPU_EINVAL
    # wrong kind and bad M
    assert ft._lib.ft_gpu_execute_noalloc(P.ptr, ft.TRI, 0, 0, ctypes.c_void_p(T.data_ptr()), 179,
                                          ctypes.c_void_p(stream.cuda_stream)) == ft.FT_GPU_EINVAL
    assert ft._lib.ft_gpu_execute_noalloc(P.ptr, ft.SPH, 0, 0, ctypes.c_void_p(T.data_ptr()), 180,
                                          ctypes.c_void_p(stream.cuda_stream)) == ft.FT_GPU_EINVAL
    ft.synchronize(stream)


def test_noalloc_vector_and_rectangle_layouts(ft):
    """The sphv and rectdisk executor kinds through the allocation-free entry
    point: the same bits as ft_gpu_execute, two calls in a row, the second
    smaller."""
    import torch
    stream = torch.cuda.current_stream()
    for kind, n in ((ft.SPHV, 300), (ft.SPHV, 140), (ft.RECTDISK, 400), (ft.RECTDISK, 90)):
        P = ft.plan_rotsphere(n) if kind == ft.SPHV else ft.plan_rotrectdisk(n, 0.5)
        M = 2 * n - 1 if kind == ft.SPHV else n
        A = np.random.default_rng(n).uniform(-1.0, 1.0, (M, n))
        X = torch.from_numpy(A).cuda()
        Y = X.clone()
        ft.set_engine(ft.ENGINE_SYSTOLIC)
        try:
            ft.gpu_execute(P, 1, X, kind=kind, stream=stream)
            ft.synchronize(stream)
        finally:
            ft.set_engine(ft.ENGINE_AUTO)
        _say("noalloc kind=%d n=%d" % (kind, n))
        assert _noalloc(ft, P, 1, Y, ft.FP64, stream, kind=kind) == 0
        ft.synchronize(stream)
        assert torch.equal(X, Y), (kind, n)
