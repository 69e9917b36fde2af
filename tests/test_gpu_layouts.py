"""GPU parity of the vector-field (sphv), rectangle-disk (rectdisk) and
annulus executors against the pinned oracle.

Both run the sphere's loop nests on other columns (drivers.c:204-220,
496-508): sphv pairs the columns 2m+1, 2m+2 of order m on a sphere plan,
rectdisk applies order m to column m on a rectangle-disk plan.  The GPU
evaluates the same fused rotations in the same dependency order as the
fused oracle, so results are asserted bit for bit (which implies the
relative 2-norm bar of 1e-14), on both engines, both directions, full and
partial widths, single and batched launches, FP64 and FP32.  Annulus plans
(QR-built on the host) run the disk executors (drivers.c:1073-1121).
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _setup(ft, layout, n, beta=0.5):
    """(library plan, oracle plan, full width, oracle executor)."""
    if layout == "sphv":
        return ft.plan_rotsphere(n), O.Plan("sph", n), 2 * n - 1, O.execute_sphv
    return ft.plan_rotrectdisk(n, beta), O.Plan("rectdisk", n, beta), n, O.execute


def _input(n, M, seed):
    return np.asfortranarray(np.random.default_rng(seed).uniform(-1.0, 1.0, (n, M)))


def _check(ft, layout, n, d, M=None, seed=3, beta=0.5):
    P, Q, full, oracle_exec = _setup(ft, layout, n, beta)
    M = full if M is None else M
    A = _input(n, M, seed)
    _say("case", layout, "n=%d M=%d dir=%d beta=%g" % (n, M, d, beta))
    G = getattr(ft, "execute_%s_%s" % (layout, ("hi2lo", "lo2hi")[d]))(P, A.copy(order="F"), M)
    R = oracle_exec(Q, d, A.copy(order="F"), M)
    assert np.array_equal(G, R), "not bitwise: rel2 %g" % O.rel2(G, R)


LAYOUTS = ["sphv", "rectdisk"]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("engine", ["auto", "pipeline", "systolic"])
def test_sizes_and_engines(ft, layout, engine):
    eng = {"auto": ft.ENGINE_AUTO, "pipeline": ft.ENGINE_PIPELINE, "systolic": ft.ENGINE_SYSTOLIC}[engine]
    try:
        ft.set_engine(eng)
        for n in (1, 2, 3, 4, 5, 9, 16, 33, 64, 100, 257):
            for d in (0, 1):
                _check(ft, layout, n, d, seed=n)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_partial_widths(ft, layout):
    """Widths that end inside an order's pair (sphv) or well before n (rectdisk),
    and the widest sphv matrix the orders allow (M = 2n+1)."""
    n = 70
    widths = (1, 3, 5, 7, 37, 99, 139, 141) if layout == "sphv" else (1, 2, 3, 17, 37, 69, 70)
    for engine in (ft.ENGINE_PIPELINE, ft.ENGINE_SYSTOLIC):
        try:
            ft.set_engine(engine)
            for M in widths:
                for d in (0, 1):
                    _check(ft, layout, n, d, M=M, seed=M)
        finally:
            ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("engine", ["pipeline", "systolic"])
def test_at_1024(ft, layout, engine):
    """n = 1024 (sphv 1024 x 2047, rectdisk 1024 x 1024) on each engine; the
    band-systolic engine with whole chains and with chains in parts of 3 bands."""
    try:
        ft.set_engine(ft.ENGINE_PIPELINE if engine == "pipeline" else ft.ENGINE_SYSTOLIC)
        for maxpart in ((0, 3) if engine == "systolic" else (0,)):
            ft.set_systolic(maxpart)
            for d in (0, 1):
                _check(ft, layout, 1024, d, seed=1024 + maxpart)
    finally:
        ft.set_systolic(0)
        ft.set_engine(ft.ENGINE_AUTO)


@pytest.mark.parametrize("beta", [0.0, -0.25, 2.5])
def test_rectdisk_parameters(ft, beta):
    for d in (0, 1):
        _check(ft, "rectdisk", 300, d, seed=7, beta=beta)


def test_rectdisk_single_column_kernels(ft):
    """ft_kernel_rectdisk_* against the oracle's kernel, with strides."""
    n = 48
    P, Q = ft.plan_rotrectdisk(n, 1.5), O.Plan("rectdisk", n, 1.5)
    for S in (1, 2):
        for m1, m2 in ((0, 8), (1, 9), (2, 17), (1, 2), (3, 30), (0, 0), (0, 47)):
            _say("kernel rectdisk", S, m1, m2)
            base = np.arange((n - 1) * S + 1, dtype=np.float64) / 7.0 - 1.3
            for d, fn in ((0, ft.kernel_rectdisk_hi2lo), (1, ft.kernel_rectdisk_lo2hi)):
                G = fn(P, m1, m2, base.copy(), S)
                R = O.kernel(Q, d, m1, m2, base.copy(), S)
                assert np.array_equal(G, R), (S, m1, m2, d)


def _f32_oracle(Q, layout, d, A32, M):
    """The oracle's FP32 loop nest on the executor's (column, order) pairs."""
    X = A32.copy(order="F")
    pairs = [(2 * m + t, m) for m in range(2, M // 2) for t in (1, 2)] if layout == "sphv" else \
        [(m, m) for m in range(2, M)]
    for j, m in pairs:
        col = np.ascontiguousarray(X[:, j])
        O.kernel_f32(Q, d, m % 2, m, col)
        X[:, j] = col
    return X


@pytest.mark.parametrize("layout", LAYOUTS)
def test_device_resident_fp32_and_fp64(ft, layout):
    """ft_gpu_execute on device tensors with the executor kind named: FP64
    bit for bit, FP32 bit for bit against the oracle's FP32 loop nest."""
    import torch
    n = 200
    P, Q, M, oracle_exec = _setup(ft, layout, n)
    A = _input(n, M, 99)
    for d in (0, 1):
        for engine in (ft.ENGINE_PIPELINE, ft.ENGINE_SYSTOLIC):
            try:
                ft.set_engine(engine)
                _say("device", layout, d, engine)
                T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
                ft.gpu_execute(P, d, T, kind=layout)
                T32 = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()
                ft.gpu_execute(P, d, T32, kind=layout, precision=ft.FP32)
                ft.synchronize()
            finally:
                ft.set_engine(ft.ENGINE_AUTO)
            assert np.array_equal(T.cpu().numpy().T, oracle_exec(Q, d, A.copy(order="F"), M))
            want32 = _f32_oracle(Q, layout, d, A.astype(np.float32, order="F"), M)
            assert np.array_equal(T32.cpu().numpy().T, want32)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("engine", ["auto", "systolic"])
def test_batch_uneven_degrees(ft, layout, engine):
    """Batches of sphv or rectdisk transforms of log-uniform degree: each
    bit-identical to the oracle; a batch plan launches the same work."""
    import torch
    rng = np.random.default_rng(11)
    ns = np.unique(np.round(np.exp(rng.uniform(np.log(4), np.log(300), 20))).astype(int))
    try:
        ft.set_engine(ft.ENGINE_AUTO if engine == "auto" else ft.ENGINE_SYSTOLIC)
        for d in (0, 1):
            _say("batch", layout, engine, d, list(ns))
            setups = [_setup(ft, layout, int(n)) for n in ns]
            As = [_input(int(n), s[2], 500 + int(n)) for n, s in zip(ns, setups)]
            Ts = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
            ft.gpu_execute_batch([s[0] for s in setups], d, Ts, kind=layout)
            torch.cuda.synchronize()
            Us = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
            bp = ft.BatchPlan([s[0] for s in setups], d, Us, kind=layout)
            bp.execute()
            ft.synchronize()
            for s, A, T, U in zip(setups, As, Ts, Us):
                R = s[3](s[1], d, A.copy(order="F"), s[2])
                assert np.array_equal(T.cpu().numpy().T, R), (layout, d, s[1].n)
                assert np.array_equal(U.cpu().numpy().T, R), (layout, d, s[1].n)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


def test_batch_rejects_mixed_constructors(ft):
    """A batch shares one closed form: a sphere plan and a rectangle-disk plan
    (same table layout, other tables) do not batch together."""
    import torch
    plans = [ft.plan_rotsphere(40), ft.plan_rotrectdisk(40, 0.0)]
    Ts = [torch.zeros(40, 40, dtype=torch.float64, device="cuda") for _ in plans]
    with pytest.raises(ft.GPUError):
        ft.gpu_execute_batch(plans, 0, Ts, Ms=[40, 40], kind="rectdisk")


def test_roundtrip_rectdisk_1024(ft):
    """hi2lo then lo2hi on the device within the reference's round-trip bounds."""
    import torch
    n = 1024
    P = ft.plan_rotrectdisk(n, 0.0)
    A = O.fill("tri", n, n, 5)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    ft.gpu_execute(P, 0, T)
    ft.gpu_execute(P, 1, T)
    ft.synchronize()
    X = T.cpu().numpy().T
    assert O.rel2(X, A) < 4 * np.sqrt(n) * 2.0 ** -52
    assert O.relinf(X, A) < 2 * n * 2.0 ** -52


@pytest.mark.parametrize("engine", ["auto", "pipeline", "systolic"])
def test_annulus_through_disk_kernels(ft, engine):
    """An annulus plan (ft_plan_rotannulus, QR-built on the host) runs the disk
    executors (drivers.c:1073-1121): bit for bit against the oracle's disk
    loop nests on the same tables, both directions, full and partial widths."""
    eng = {"auto": ft.ENGINE_AUTO, "pipeline": ft.ENGINE_PIPELINE, "systolic": ft.ENGINE_SYSTOLIC}[engine]
    try:
        ft.set_engine(eng)
        for n, M, rho in ((5, None, 0.3), (64, None, np.sqrt(1 / 3)), (300, None, 0.6), (300, 457, 0.2),
                          (1024, None, 0.5)):
            P = ft.plan_rotannulus(n, 0.0, 0.0, 0.0, rho)
            Q = O.Plan("annulus", n, 0.0, 0.0, rho=rho)
            M = 4 * n - 3 if M is None else M
            A = O.fill("disk", n, 4 * n - 3, n)[:, :M].copy(order="F")
            for d in (0, 1):
                _say("annulus", engine, n, M, d)
                G = (ft.execute_disk_hi2lo if d == 0 else ft.execute_disk_lo2hi)(P, A.copy(order="F"), M)
                assert np.array_equal(G, O.execute(Q, d, A.copy(order="F"), M)), (n, M, d)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)


def test_annulus_batch(ft):
    """A batch of annulus plans of uneven degree (one inner radius) in one
    launch, each bit for bit; plans of another radius do not batch with them."""
    import torch
    ns = [7, 40, 129, 300]
    plans = [ft.plan_rotannulus(n, 0.0, 0.0, 0.0, 0.5) for n in ns]
    As = [O.fill("disk", n, 4 * n - 3, 70 + n) for n in ns]
    for d in (0, 1):
        Ts = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in As]
        ft.gpu_execute_batch(plans, d, Ts)
        ft.synchronize()
        for n, A, T in zip(ns, As, Ts):
            R = O.execute(O.Plan("annulus", n, 0.0, 0.0, rho=0.5), d, A.copy(order="F"))
            assert np.array_equal(T.cpu().numpy().T, R), (n, d)
    other = plans[:2] + [ft.plan_rotannulus(40, 0.0, 0.0, 0.0, 0.25)]
    Ts = [torch.zeros(4 * P.n - 3, P.n, dtype=torch.float64, device="cuda") for P in other]
    with pytest.raises(ft.GPUError):
        ft.gpu_execute_batch(other, 0, Ts)
