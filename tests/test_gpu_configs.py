"""GPU checks at the benchmark configurations and of the boundary contract.

* C1: the n = 256 FP64 round trip, each step bit-identical to the oracle,
  within the reference's bounds (test_drivers.c:52-57) and close to the
  reference's own dispatched executor (oracle/_ref).
* C2: n = 1024 and 2048 (and tri/disk at 1024) against the reference's own
  dispatched executor directly, within 1e-14 relative 2-norm.
* C3: the FP32 variant bit-identical to the oracle's FP32 loop nests at
  n = 512, 1024, 2048, both directions, with its error against the FP64 and
  the __float128 evaluations reported per degree.
* C4: the 512-transform batch (seed 4242, n log-uniform on [32, 1024]),
  eager and replayed from a captured HIP graph, every transform bit-identical.
* C5: batches of 64 tri (0,0,0), tri (-1/2,-1/2,-1/2) and disk (0,0)
  transforms with n uniform on [128, 1024].
* The contract: concurrent calls on one plan from two host threads, batch
  schedules across destroyed and re-created plans, cached single-vector
  schedules, and the asynchronous error check.
"""
import ctypes
import sys
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]


def _maker(ft, kind):
    return {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]


def _oracle_many(jobs):
    """Runs O.execute over (plan, dir, A) jobs on host threads (the oracle's C
    code releases the interpreter lock)."""
    with ThreadPoolExecutor(max_workers=4) as ex:
        return list(ex.map(lambda j: O.execute(j[0], j[1], j[2]), jobs))


def test_c1_sph_256_roundtrip(ft):
    n, M = 256, 511
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = O.fill("sph", n, M, 256)
    _say("C1 sph n=256 hi2lo")
    G1 = ft.execute_sph_hi2lo(P, A.copy(order="F"))
    R1 = O.execute(Q, 0, A.copy(order="F"))
    assert np.array_equal(G1, R1)
    _say("C1 sph n=256 lo2hi")
    G2 = ft.execute_sph_lo2hi(P, G1.copy(order="F"))
    R2 = O.execute(Q, 1, R1.copy(order="F"))
    assert np.array_equal(G2, R2)
    e2, einf = O.rel2(G2, A), O.relinf(G2, A)
    _say("C1 round trip rel2 %.3e relinf %.3e" % (e2, einf))
    assert e2 < 4 * np.sqrt(n) * EPS
    assert einf < 2 * n * EPS
    if O.have_ref():
        X = O.ref_execute(Q, 0, A.copy(order="F"))
        assert O.rel2(G1, X) <= 1e-14
        X = O.ref_execute(Q, 1, X)
        assert O.rel2(G2, X) <= 1e-14


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref (the reference's compiled kernels) not built")
@pytest.mark.parametrize("kind,params,n", [("sph", (), 1024), ("sph", (), 2048), ("tri", (0.0, 0.0, 0.0), 1024),
                                           ("tri", (-0.5, -0.5, -0.5), 1024), ("disk", (0.0, 0.0), 1024)])
def test_c2_against_reference_dispatch(ft, kind, params, n):
    """The GPU output against the reference's own dispatched executor on the
    same seeded input: relative 2-norm <= 1e-14."""
    import torch
    M = full_M(kind, n)
    P = _maker(ft, kind)(n, *params)
    Q = O.Plan(kind, n, *params)
    A = O.fill(kind, n, M, 2024)
    for d in (0, 1):
        T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        _say("C2", kind, params, n, d)
        ft.gpu_execute(P, d, T)
        ft.synchronize()
        G = T.cpu().numpy().T
        R = O.ref_execute(Q, d, A.copy(order="F"))
        err = O.rel2(G, R)
        _say("C2 %s%s n=%d dir=%d rel2 vs reference %.3e" % (kind, params, n, d, err))
        assert err <= 1e-14


@pytest.mark.parametrize("n", [512, 1024, 2048])
def test_c3_fp32_bitwise_and_errors(ft, n):
    """FP32 on the GPU equals the oracle's FP32 loop nests bit for bit; its
    error against FP64 (and, up to n = 1024, __float128) is reported."""
    import torch
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = O.fill("sph", n, 2 * n - 1, 77)
    for d in (0, 1):
        T = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()
        _say("C3 fp32 n=%d dir=%d" % (n, d))
        ft.gpu_execute(P, d, T, precision=ft.FP32)
        ft.synchronize()
        G = T.cpu().numpy().T
        F = O.execute_f32(Q, d, A.astype(np.float32, order="F"))
        assert np.array_equal(G, F), "FP32 not bitwise equal to the FP32 oracle"
        R = O.execute(Q, d, A.copy(order="F"))
        e2, einf = O.rel2(G.astype(np.float64), R), O.relinf(G.astype(np.float64), R)
        msg = "C3 fp32 n=%d dir=%d vs FP64: rel2 %.3e relinf %.3e" % (n, d, e2, einf)
        if n <= 1024:
            H = O.execute_q(Q, d, A.copy(order="F"), hp_plan=True)
            msg += "; vs float128: rel2 %.3e relinf %.3e" % (O.rel2(G.astype(np.float64), H),
                                                          O.relinf(G.astype(np.float64), H))
        _say(msg)
        assert e2 < 8 * np.sqrt(n) * 2.0 ** -23


def _c4_inputs(count=512, seed=4242):
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(32), np.log(1024), count))).astype(int)
    return ns, [O.fill("sph", int(n), 2 * int(n) - 1, seed + i) for i, n in enumerate(ns)]


def test_c4_batch_512(ft):
    """The benchmark's batch itself: 512 sph hi2lo transforms (seed 4242),
    one launch, then a batch plan replayed from a captured HIP graph; every
    transform bit-identical to the oracle."""
    import torch
    ns, As = _c4_inputs()
    plans = {int(n): ft.plan_rotsphere(int(n)) for n in np.unique(ns)}
    oplans = {int(n): O.Plan("sph", int(n)) for n in np.unique(ns)}
    pl = [plans[int(n)] for n in ns]
    _say("C4 oracle, %d transforms, n %d..%d" % (len(ns), ns.min(), ns.max()))
    R = _oracle_many([(oplans[int(n)], 0, A.copy(order="F")) for n, A in zip(ns, As)])
    T = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
    _say("C4 batched launch")
    ft.gpu_execute_batch(pl, 0, T)
    ft.synchronize()
    for i, (t, r) in enumerate(zip(T, R)):
        assert np.array_equal(t.cpu().numpy().T, r), (i, int(ns[i]))
    # the same batch from a captured graph, replayed twice with eager work between
    G = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
    bp = ft.BatchPlan(pl, "hi2lo", G)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        bp.execute(s)
    torch.cuda.synchronize()
    for rep in range(2):
        for t, A in zip(G, As):
            t.copy_(torch.from_numpy(np.ascontiguousarray(A.T)))
        torch.cuda.synchronize()
        _say("C4 graph replay", rep)
        g.replay()
        ft.synchronize(torch.cuda.current_stream())
        torch.cuda.synchronize()
        for i, (t, r) in enumerate(zip(G, R)):
            assert np.array_equal(t.cpu().numpy().T, r), (rep, i, int(ns[i]))
        _ = (G[0] * 2).sum().item()
    del g
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind,params", [("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)), ("disk", (0.0, 0.0))])
def test_c5_batch_64(ft, kind, params):
    """64 transforms with n uniform on [128, 1024] in one launch, each
    bit-identical to the oracle."""
    import torch
    rng = np.random.default_rng(5150)
    ns = rng.integers(128, 1025, 64)
    plans = {int(n): _maker(ft, kind)(int(n), *params) for n in np.unique(ns)}
    oplans = {int(n): O.Plan(kind, int(n), *params) for n in np.unique(ns)}
    As = [O.fill(kind, int(n), full_M(kind, int(n)), 5150 + i) for i, n in enumerate(ns)]
    _say("C5", kind, params, "oracle")
    R = _oracle_many([(oplans[int(n)], 0, A.copy(order="F")) for n, A in zip(ns, As)])
    T = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
    _say("C5", kind, params, "batched launch")
    ft.gpu_execute_batch([plans[int(n)] for n in ns], 0, T)
    ft.synchronize()
    for i, (t, r) in enumerate(zip(T, R)):
        assert np.array_equal(t.cpu().numpy().T, r), (i, int(ns[i]))


def test_two_threads_one_plan(ft):
    """The reference's executors only read the plan, so concurrent calls on
    one plan with distinct matrices are safe (drivers.c:120-124).  Two host
    threads call the drop-in executor on one plan with host matrices of the
    same M; every result is bit-identical to the oracle."""
    n = 300
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    inputs = [O.fill("sph", n, 2 * n - 1, 900 + t) for t in range(2)]
    expect = [O.execute(Q, 0, A.copy(order="F")) for A in inputs]
    bad = []

    def work(t):
        for it in range(6):
            G = ft.execute_sph_hi2lo(P, inputs[t].copy(order="F"))
            if not np.array_equal(G, expect[t]):
                bad.append((t, it))

    _say("two threads, one plan")
    th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not bad, bad


def test_batch_cache_survives_plan_recreation(ft):
    """The batch schedule cache is keyed on every plan (with its serial
    number and n), matrix pointer and M: plans destroyed and re-created between
    calls (often at the same addresses), with the same M, get a fresh schedule
    and exact results."""
    import torch
    for rnd, ns in enumerate(([60, 90, 40], [61, 90, 41], [60, 90, 40])):
        plans = [ft.plan_rotsphere(n) for n in ns]
        M = 79  # valid for every n here, so only the plans tell the rounds apart
        As = [O.fill("sph", n, 2 * n - 1, 40 + n)[:, :M].copy(order="F") for n in ns]
        T = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]  # (M, n): row j = column j
        _say("batch cache round", rnd, ns)
        ft.gpu_execute_batch(plans, 0, T, Ms=[M] * len(ns))
        ft.synchronize()
        for n, A, t in zip(ns, As, T):
            R = O.execute(O.Plan("sph", n), 0, A.copy(order="F"), M=M)
            assert np.array_equal(t.cpu().numpy().T, R), (rnd, n)
        del plans
        import gc
        gc.collect()


def test_single_vector_schedules_cached(ft):
    """ft_kernel_* keeps a schedule per (direction, sequences, stride): the
    reference's per-order loop (test_rotations.c:25-39) run twice, with host
    and with device buffers, gives the oracle's bits every time and the
    round trip within the reference's bound."""
    import torch
    n = 64
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    err = 0.0
    for rep in range(2):
        for m in range(2, n):
            base = np.zeros(n)
            base[: n - m] = 1.0
            G = ft.kernel_sph_hi2lo(P, m % 2, m, base.copy(), 1)
            R = O.kernel(Q, 0, m % 2, m, base.copy(), 1)
            assert np.array_equal(G, R), (rep, m)
            if rep == 1:
                T = torch.from_numpy(base.copy()).cuda()
                ft.kernel_sph_hi2lo(P, m % 2, m, T, 1)
                assert np.array_equal(T.cpu().numpy(), R), ("device", m)
            B = ft.kernel_sph_lo2hi(P, m % 2, m, G.copy(), 1)
            err += (O.rel2(B, base)) ** 2
    assert np.sqrt(err / 2) < 2 * n * EPS


def test_async_error_is_reported_by_synchronize(ft):
    """An asynchronous launch whose bounded waits expire (forced from the host
    with a one-poll limit) is reported by ft_gpu_synchronize on its stream,
    once; the next launch with the limit restored is exact and clean."""
    import torch
    n = 300
    P = ft.plan_rotsphere(n)
    A = O.fill("sph", n, 2 * n - 1, 12)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    try:
        ft.set_engine(ft.ENGINE_SYSTOLIC)
        ft.set_spin_limit(1)
        _say("async timeout path")
        ft.gpu_execute(P, 0, T)
        with pytest.raises(ft.GPUError, match="wait expired"):
            ft.synchronize()
        ft.synchronize()  # reported once
    finally:
        ft.set_spin_limit(0)
        ft.set_engine(ft.ENGINE_AUTO)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    ft.gpu_execute(P, 0, T)
    ft.synchronize()
    assert np.array_equal(T.cpu().numpy().T, O.execute(O.Plan("sph", n), 0, A.copy(order="F")))
