"""GPU parity of the HIP rotation path against the pinned CPU oracle.

The bar: relative 2-norm <= 1e-14 against the oracle (the reference's
algorithm in the fused FP64 form of its SIMD kernels) for every geometry,
direction and size here, and hi2lo -> lo2hi round trips within the
reference's own bounds (test_drivers.c:52-57).  The GPU kernel evaluates the
same fused rotations in the same dependency order, so the results are also
expected to be bit-identical; that is asserted too.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-14
EPS = 2.0 ** -52


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]


def _plans(ft, kind, n, params):
    maker = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]
    return maker(n, *params), O.Plan(kind, n, *params)


def _exec(ft, kind, d, P, A, M=None):
    fn = getattr(ft, "execute_%s_%s" % (kind, ("hi2lo", "lo2hi")[d]))
    return fn(P, A, M)


def _check(ft, kind, params, n, d, seed=17, M=None):
    M = full_M(kind, n) if M is None else M
    P, Q = _plans(ft, kind, n, params)
    A = O.fill(kind, n, full_M(kind, n), seed)[:, :M].copy(order="F")
    _say("case", kind, params, "n=%d M=%d dir=%d" % (n, M, d))
    G = _exec(ft, kind, d, P, A.copy(order="F"), M)
    R = O.execute(Q, d, A.copy(order="F"), M)
    err = O.rel2(G, R)
    assert err <= TOL, err
    assert np.array_equal(G, R), "not bitwise: rel2 %g" % err
    return A, G


CASES = [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)), ("disk", (0.0, 0.0))]


@pytest.mark.parametrize("kind,params", CASES)
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 16, 33, 64, 100])
@pytest.mark.parametrize("d", [0, 1])
def test_small_sizes(ft, kind, params, n, d):
    _check(ft, kind, params, n, d)


@pytest.mark.parametrize("kind,params", CASES + [("tri", (0.25, -0.5, 1.5)), ("disk", (-0.5, 2.0))])
@pytest.mark.parametrize("d", [0, 1])
def test_medium_sizes(ft, kind, params, d):
    _check(ft, kind, params, 257, d)


@pytest.mark.parametrize("kind,params", CASES)
def test_partial_M(ft, kind, params):
    n = 70
    for M in ((1, 2, 5, 37, 69, 70) if kind == "tri" else (1, 3, 5, 37, 99, 131)):
        for d in (0, 1):
            _check(ft, kind, params, n, d, M=M)


@pytest.mark.parametrize("n", [1024, 2048])
def test_sph_full_sizes(ft, n):
    """The headline sizes: n = 1024 (1024 x 2047) and 2048 (2048 x 4095)."""
    for d in (0, 1):
        _check(ft, "sph", (), n, d, seed=1000 + n)


def test_roundtrip_sph_1024_device_resident(ft):
    import torch
    n, M = 1024, 2047
    P = ft.plan_rotsphere(n)
    A = O.fill("sph", n, M, 3)
    _say("roundtrip sph n=1024 device tensor")
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()  # (M, n): row j = column j
    ft.execute_sph_hi2lo(P, T)
    ft.execute_sph_lo2hi(P, T)
    X = T.cpu().numpy().T
    assert O.rel2(X, A) < 4 * np.sqrt(n) * EPS
    assert O.relinf(X, A) < 2 * n * EPS


@pytest.mark.parametrize("kind,params", CASES)
def test_geometries_at_1024(ft, kind, params):
    _check(ft, kind, params, 1024, 0, seed=5)


@pytest.mark.parametrize("kind,params", CASES)
def test_single_column_kernels(ft, kind, params):
    """ft_kernel_* against the oracle, including strides (test_rotations.c:25-39)."""
    n = 48
    P, Q = _plans(ft, kind, n, params)
    hi = getattr(ft, "kernel_%s_hi2lo" % kind)
    lo = getattr(ft, "kernel_%s_lo2hi" % kind)
    for S in (1, 2):
        for m1, m2 in ((0, 8), (1, 9), (2, 17), (1, 2), (3, 30), (0, 0), (0, 47)):
            _say("kernel", kind, S, m1, m2)
            base = np.arange((n - 1) * S + 1, dtype=np.float64) / 7.0 - 1.3
            for d, fn in ((0, hi), (1, lo)):
                G = fn(P, m1, m2, base.copy(), S)
                R = O.kernel(Q, d, m1, m2, base.copy(), S)
                assert np.array_equal(G, R), (kind, S, m1, m2, d)


@pytest.mark.parametrize("L,W", [(2, 8), (4, 4), (4, 16), (8, 4), (8, 8), (2, 16)])
def test_tuning_shapes_agree(ft, L, W):
    """Every pipeline shape computes the same bits."""
    try:
        ft.set_engine(ft.ENGINE_PIPELINE)
        ft.set_tuning(L, W, 0)
        for kind, params in CASES:
            for d in (0, 1):
                _check(ft, kind, params, 150, d, seed=21)
    finally:
        ft.set_tuning(0, 0, 0)
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("engine", ["pipeline", "systolic"])
@pytest.mark.parametrize("kind,params", CASES)
def test_engines(ft, engine, kind, params):
    """Each executor engine, forced, against the oracle: sizes that give one
    band, several bands, column masks across band edges, and partial M."""
    try:
        ft.set_engine(ft.ENGINE_PIPELINE if engine == "pipeline" else ft.ENGINE_SYSTOLIC)
        for n, M in ((40, None), (129, None), (300, None), (300, 77 if kind == "tri" else 155)):
            for d in (0, 1):
                _check(ft, kind, params, n, d, seed=31, M=M)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("engine", ["pipeline", "systolic"])
def test_engines_at_limits(ft, engine):
    """Both engines forced at the headline shape (sph n = 1024, 16 bands per
    chain in the systolic engine) and tri/disk at n = 1024 (32 bands per
    chain: the systolic engine runs them in parts on several workgroups)."""
    try:
        ft.set_engine(ft.ENGINE_PIPELINE if engine == "pipeline" else ft.ENGINE_SYSTOLIC)
        for kind, params, n in (("sph", (), 1024), ("tri", (0.0, 0.0, 0.0), 1024), ("disk", (0.0, 0.0), 1024)):
            for d in (0, 1):
                _check(ft, kind, params, n, d, seed=77)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("maxpart", [1, 3, 16])
def test_systolic_chain_parts(ft, maxpart):
    """Chains split into parts of at most `maxpart` bands (one band per part:
    every band hands its rows to the next through A and a progress counter,
    across workgroups) give the oracle's bits, both directions, every
    geometry, full and partial M."""
    try:
        ft.set_engine(ft.ENGINE_SYSTOLIC)
        ft.set_systolic(maxpart)
        for kind, params in CASES:
            for n, M in ((129, None), (300, 77 if kind == "tri" else 155), (300, None)):
                for d in (0, 1):
                    _check(ft, kind, params, n, d, seed=41 + maxpart, M=M)
    finally:
        ft.set_systolic(0)
        ft.set_engine(ft.ENGINE_AUTO)


def test_systolic_wait_timeout_is_reported(ft):
    """The band-systolic engine's waits are bounded: with a host-set limit of
    one poll, the waits between bands expire, every wave leaves its loop, and
    the executor reports FT_GPU_ETIMEOUT; with the limit restored the same
    plan gives exact results again."""
    n = 300
    P, Q = _plans(ft, "sph", n, ())
    A = O.fill("sph", n, 2 * n - 1, 11)
    try:
        ft.set_engine(ft.ENGINE_SYSTOLIC)
        ft.set_spin_limit(1)
        _say("systolic timeout path")
        with pytest.raises(ft.GPUError, match="wait expired"):
            ft.execute_sph_hi2lo(P, A.copy(order="F"))
    finally:
        ft.set_spin_limit(0)
        ft.set_engine(ft.ENGINE_AUTO)
    G = ft.execute_sph_hi2lo(P, A.copy(order="F"))
    assert np.array_equal(G, O.execute(Q, 0, A.copy(order="F")))


def test_async_device_api_and_stream(ft):
    import torch
    n = 300
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = O.fill("sph", n, 2 * n - 1, 8)
    s = torch.cuda.Stream()
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        ft.gpu_execute(P, "hi2lo", T, stream=s)
    s.synchronize()
    R = O.execute(Q, 0, A.copy(order="F"))
    assert np.array_equal(T.cpu().numpy().T, R)


@pytest.mark.parametrize("engine", ["auto", "systolic"])
def test_batch_uneven_degrees(ft, engine):
    """ft_gpu_execute_batch: transforms of log-uniform degree in one launch,
    each bit-identical to the oracle; the automatic choice runs batches on the
    engine the tuning table names; forced, the band-systolic engine runs them (tri
    n = 600 in FP64 with chains split over several workgroups)."""
    import torch
    rng = np.random.default_rng(4)
    ns = np.unique(np.round(np.exp(rng.uniform(np.log(4), np.log(300), 24))).astype(int))
    cases = [("sph", (), ns), ("tri", (0.0, 0.0, 0.0), ns), ("disk", (0.0, 0.0), ns),
             ("tri", (0.0, 0.0, 0.0), np.append(ns[:6], 600))]
    try:
        ft.set_engine(ft.ENGINE_AUTO if engine == "auto" else ft.ENGINE_SYSTOLIC)
        for kind, params, sizes in cases:
            for d in (0, 1):
                _say("batch", engine, kind, d, list(sizes))
                maker = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]
                plans = [maker(int(n), *params) for n in sizes]
                As = [O.fill(kind, int(n), full_M(kind, int(n)), 100 + int(n)) for n in sizes]
                Ts = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
                ft.gpu_execute_batch(plans, d, Ts)
                used = ft.last_engine()
                torch.cuda.synchronize()
                want = ft.tuned_choice(kind, 0, d, 1, int(max(sizes)))[0] if engine == "auto" else ft.ENGINE_SYSTOLIC
                assert used == want, (used, want)
                for n, A, T in zip(sizes, As, Ts):
                    R = O.execute(O.Plan(kind, int(n), *params), d, A.copy(order="F"))
                    assert np.array_equal(T.cpu().numpy().T, R), (kind, d, n)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_fp32_path_error(ft, n):
    """FP32 storage and arithmetic, s/c rounded from the FP64 plan: error
    against the FP64 oracle (reported) within an FP32 rounding bound."""
    import torch
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = O.fill("sph", n, 2 * n - 1, 77)
    for d in (0, 1):
        T = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()
        ft.gpu_execute(P, d, T, precision=ft.FP32)
        torch.cuda.synchronize()
        G = T.double().cpu().numpy().T
        R = O.execute(Q, d, A.copy(order="F"))
        e2, einf = O.rel2(G, R), O.relinf(G, R)
        _say("fp32 n=%d dir=%d rel2=%.3e relinf=%.3e" % (n, d, e2, einf))
        assert e2 < 8 * np.sqrt(n) * 2.0 ** -23


def test_fp32_wide_bands_engines_agree(ft):
    """FP32 shapes that the band-systolic engine runs with 4 sequences per
    lane (more than 16 bands of 2): it and the pipelined engine evaluate the
    same fused FP32 rotations in the same order, so they agree bit for bit,
    and both stay within the FP32 bound of the FP64 oracle."""
    import torch
    cases = (("sph", (), 2048), ("tri", (0.0, 0.0, 0.0), 1024), ("disk", (0.0, 0.0), 700))
    try:
        for kind, params, n in cases:
            P, Q = _plans(ft, kind, n, params)
            A = O.fill(kind, n, full_M(kind, n), 88)
            R = O.execute(Q, 0, A.copy(order="F"))
            got = {}
            for name, eng in (("systolic", ft.ENGINE_SYSTOLIC), ("pipeline", ft.ENGINE_PIPELINE)):
                ft.set_engine(eng)
                T = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()
                _say("fp32", name, kind, n)
                ft.gpu_execute(P, 0, T, precision=ft.FP32)
                assert ft.last_engine() == eng
                torch.cuda.synchronize()
                got[name] = T.cpu().numpy().T
                e2 = O.rel2(got[name].astype(np.float64), R)
                _say("fp32 %s %s n=%d rel2=%.3e" % (name, kind, n, e2))
                assert e2 < 8 * np.sqrt(n) * 2.0 ** -23
            assert np.array_equal(got["systolic"], got["pipeline"]), (kind, n)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


def test_batch_graph_replays(ft):
    """A batch plan's launch captured in a HIP graph: every replay (with eager
    work and synchronisation between replays) equals eager execution."""
    import torch
    ns = [40, 77, 130, 64]
    plans = [ft.plan_rotsphere(n) for n in ns]
    As = [O.fill("sph", n, 2 * n - 1, 300 + n) for n in ns]
    G = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
    E = [t.clone() for t in G]
    BG = ft.BatchPlan(plans, "hi2lo", G)
    BE = ft.BatchPlan(plans, "hi2lo", E)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        BG.execute(s)
    torch.cuda.synchronize()
    for rep in range(3):
        _say("graph replay", rep)
        g.replay()
        torch.cuda.synchronize()
        BE.execute()
        torch.cuda.synchronize()
        for a, b in zip(G, E):
            assert torch.equal(a, b), rep
        _ = (E[0] * 2).sum().item()  # unrelated eager work between replays
    R = [O.execute(O.Plan("sph", n), 0, A.copy(order="F")) for n, A in zip(ns, As)]
    for _ in range(2):
        R = [O.execute(O.Plan("sph", n), 0, r.copy(order="F")) for n, r in zip(ns, R)]
    for r, t in zip(R, G):
        assert np.array_equal(t.cpu().numpy().T, r)
    del g
    torch.cuda.synchronize()
