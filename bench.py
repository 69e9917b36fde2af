"""Benchmark of the rotation stage on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu-baseline]
                    [--no-batch] [--calib-copy] [--engine auto|pipeline|systolic]

A step is one ft_execute_sph_hi2lo of the 1024 x 2047 FP64 spherical-harmonic
matrix (degree 1023) held in HBM, launched with ft_gpu_execute on the current
stream.  The timed region is K back-to-back steps between two HIP events on
that stream, bracketed by device synchronisation; after every timed region
ft_gpu_synchronize reads the launches' device error words, and the run exits
non-zero if any launch reported invalid results.  Before timing, the GPU's
hi2lo output is compared with the reference's own dispatched executor
(oracle/_ref) on the same seeded input: relative 2-norm above 1e-14 exits
non-zero too.

Besides the headline the JSON line carries, under "config": lo2hi at n = 1024;
the n = 256 round trip; hi2lo at n = 2048; FP32 at n = 512, 1024, 2048 with
its error against the FP64 result; tri (0,0,0), tri (-1/2,-1/2,-1/2) and disk
(0,0) at n = 1024; the vector-field (sphv), rectangle-disk (beta 0) and
annulus (rho = sqrt(1/3)) executors at n = 1024; the harmonic transforms
(sph2fourier, tri2cheb, disk2cxf and their inverses) at n = 1024 beside the
reference's published times; the batch of 512 sph transforms of log-uniform degree in
[32, 1024] (one launch, a captured HIP graph, and a host loop over the
single-transform entry point); and batches of 64 tri/disk transforms with n
uniform on [128, 1024].  It also carries the roofline of the rotation kernel
and the reference's CPU executors (their own compiled kernels) timed on this
host on bounded samples.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector peak (AMD specification)


class BenchError(RuntimeError):
    pass


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=3.0, help="time budget of each CPU baseline row")
    p.add_argument("--no-batch", action="store_true", help="skip the secondary measurements")
    p.add_argument("--engine", choices=("auto", "pipeline", "systolic"), default="auto",
                   help="engine of the headline launches (auto: the library's own choice)")
    p.add_argument("--case", help="time one case only, KIND:N:DIR[:fp32][:batch] (for profiling runs; KIND one of "
                                  "sph, sphv, tri, disk, rectdisk, annulus, or tet on the N x N x N array); "
                                  "grid:WHICH:N[:TRANS] for the grid transforms (WHICH e.g. sph_synthesis), "
                                  "spinsph:N:DIR:S for the spin-weighted sphere, "
                                  "e.g. tri:1024:hi2lo or sph:1024:hi2lo:batch (the 512-transform batch)")
    p.add_argument("--calib-copy", action="store_true",
                   help="also run a plain 64 MiB device copy (untimed), to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return p.parse_args()


def events(ft, torch, stream, fn, steps, warmup):
    """ms per call of fn() on `stream`, between HIP events on that stream;
    the launches' error words are read afterwards."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    try:
        ft.synchronize(stream)
    except ft.GPUError as e:
        raise BenchError("a timed launch reported invalid results: %s" % e)
    return start.elapsed_time(stop) / steps


def _plan(ft, kind, n, params=()):
    return {"sph": ft.plan_rotsphere, "sphv": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk,
            "rectdisk": ft.plan_rotrectdisk, "annulus": ft.plan_rotannulus}[kind](n, *params)


# parameters of the secondary kinds' plans: rectangle-disk beta = 0; annulus
# alpha = beta = gamma = 0 with the reference test's inner radius sqrt(1/3)
# (test_drivers.c:978)
PARAMS = {"sph": (), "sphv": (), "tri": (0.0, 0.0, 0.0), "disk": (0.0, 0.0), "rectdisk": (0.0,),
          "annulus": (0.0, 0.0, 0.0, float(np.sqrt(1.0 / 3.0)))}


def single(ft, torch, kind, n, direction, precision, steps, warmup, params=(), seed=2024, keep=False):
    from fasttransforms_amd.utilities import fill, full_M, rotations
    M = full_M(kind, n)
    P = _plan(ft, kind, n, params)
    dtype = torch.float64 if precision == ft.FP64 else torch.float32
    T = torch.from_numpy(np.ascontiguousarray(fill(kind, n, M, seed).T)).to(dtype).cuda()  # (M, n): row j = column j
    stream = torch.cuda.current_stream()
    ex = "sphv" if kind == "sphv" else None  # the other kinds run their plan's own executor
    ms = events(ft, torch, stream,
                lambda: ft.gpu_execute(P, direction, T, precision=precision, stream=stream, kind=ex), steps, warmup)
    out = (ms, 6.0 * rotations(kind, n, M))
    return out + ((P, T),) if keep else out


def tet(ft, torch, N, direction, steps, warmup):
    """ft_gpu_execute_tet on the N x N x N array (the reference's timing shape,
    test_drivers.c:1376-1380, its parameters alpha = beta = gamma = delta = 0)."""
    from fasttransforms_amd.utilities import tet_rotations
    L = M = N
    P1, P2 = ft.plan_rottriangle(N, 0.0, 0.0, 1.0), ft.plan_rottriangle(N, 0.0, 0.0, 0.0)
    g = torch.Generator(device="cpu").manual_seed(N)
    T = (2 * torch.rand(N * L * M, dtype=torch.float64, generator=g) - 1).cuda()
    stream = torch.cuda.current_stream()
    ms = events(ft, torch, stream, lambda: ft.gpu_execute_tet(P1, P2, direction, T, L, M, stream=stream), steps, warmup)
    return ms, 6.0 * tet_rotations(N, L, M)


def spin_rows(ft, torch, n=1024, spins=(0, 1, 2, -2, 4)):
    """ft_gpu_execute_spinsph on the n x (2n-1) complex array (the
    reference's timing shape, test_drivers.c:1591-1600)."""
    out = {}
    M = 2 * n - 1
    for s in spins:
        P = ft.plan_rotspinsphere(n, s)
        g = torch.Generator(device="cpu").manual_seed(n + s)
        T = torch.complex(2 * torch.rand(M, n, dtype=torch.float64, generator=g) - 1,
                          2 * torch.rand(M, n, dtype=torch.float64, generator=g) - 1).cuda()
        stream = torch.cuda.current_stream()
        for d in ("hi2lo", "lo2hi"):
            ms = events(ft, torch, stream, lambda: ft.gpu_execute_spinsph(P, d, T, stream=stream), 5, 2)
            out["spinsph_n%d_s%d_%s_ms" % (n, s, d)] = round(ms, 4)
    return out


def grid_rows(ft, torch, n=1024):
    """Synthesis and analysis on the n x (2n-1) grid (ft_gpu_execute_sph_grid),
    and the whole spherical harmonic transform on the device: sph2fourier then
    synthesis (coefficients to grid values), analysis then fourier2sph."""
    M = 2 * n - 1
    PS, PA = ft.plan_sph_synthesis(n, M), ft.plan_sph_analysis(n, M)
    H = ft.plan_sph2fourier(n)
    g = torch.Generator(device="cpu").manual_seed(n)
    T = (2 * torch.rand(M, n, dtype=torch.float64, generator=g) - 1).cuda()
    stream = torch.cuda.current_stream()
    out = {}
    out["sph_synthesis_n%d_ms" % n] = round(events(
        ft, torch, stream, lambda: ft.gpu_execute_sph_grid("sph_synthesis", "N", PS, T, stream=stream), 10, 2), 4)
    out["sph_analysis_n%d_ms" % n] = round(events(
        ft, torch, stream, lambda: ft.gpu_execute_sph_grid("sph_analysis", "N", PA, T, stream=stream), 10, 2), 4)

    def sht():
        ft.gpu_execute_harmonic("sph2fourier", "N", H, T, stream=stream)
        ft.gpu_execute_sph_grid("sph_synthesis", "N", PS, T, stream=stream)

    def isht():
        ft.gpu_execute_sph_grid("sph_analysis", "N", PA, T, stream=stream)
        ft.gpu_execute_harmonic("fourier2sph", "N", H, T, stream=stream)

    out["sht_synthesis_n%d_ms" % n] = round(events(ft, torch, stream, sht, 10, 2), 4)
    out["sht_analysis_n%d_ms" % n] = round(events(ft, torch, stream, isht, 10, 2), 4)
    return out


def tet_rows(ft, torch, sizes=(256, 512)):
    out = {}
    for N in sizes:
        for d in ("hi2lo", "lo2hi"):
            ms, fl = tet(ft, torch, N, d, 5, 2)
            out["tet_n%d_%s_ms" % (N, d)] = round(ms, 4)
            out["tet_n%d_%s_gflops" % (N, d)] = round(fl / (ms * 1e-3) / 1e9, 2)
    return out


def gpu_vs_reference(ft, torch, n):
    """rel-2 of the GPU's hi2lo and lo2hi outputs against the reference's own
    dispatched executor on the same seeded input (None if oracle/_ref is absent)."""
    import oracle as O
    if not O.have_ref():
        return None
    P = ft.plan_rotsphere(n)
    Q = O.Plan("sph", n)
    A = O.fill("sph", n, 2 * n - 1, 2024)
    out = {}
    for d, name in ((0, "hi2lo"), (1, "lo2hi")):
        T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        ft.gpu_execute(P, d, T)
        ft.synchronize()
        R = O.ref_execute(Q, d, A.copy(order="F"))
        out[name] = float(O.rel2(T.cpu().numpy().T, R))
    return out


def roundtrip_256(ft, torch, steps=20):
    """C1: hi2lo then lo2hi at n = 256, timed as one step; round-trip error."""
    from fasttransforms_amd.utilities import fill, rotations
    n, M = 256, 511
    P = ft.plan_rotsphere(n)
    A = fill("sph", n, M, 256)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    stream = torch.cuda.current_stream()

    def step():
        ft.gpu_execute(P, 0, T, stream=stream)
        ft.gpu_execute(P, 1, T, stream=stream)

    T0 = T.clone()
    step()
    torch.cuda.synchronize()
    X = T.cpu().numpy().T
    e2 = float(np.linalg.norm(X - A) / np.linalg.norm(A))
    einf = float(np.max(np.abs(X - A)) / np.max(np.abs(A)))
    T.copy_(T0)
    ms = events(ft, torch, stream, step, steps, 3)
    fl = 2 * 6.0 * rotations("sph", n, M)
    return {"ms": round(ms, 5), "gflops": round(fl / (ms * 1e-3) / 1e9, 2), "rel2": e2, "relinf": einf}


def fp32_rows(ft, torch, sizes):
    """C3: FP32 throughput and its error against the FP64 GPU result (itself
    bit-identical to the oracle) on the same input."""
    rows = {}
    for nn in sizes:
        ms32, fl, (P, T32) = single(ft, torch, "sph", nn, "hi2lo", ft.FP32, 10, 2, keep=True)
        ms64, _, (_, T64) = single(ft, torch, "sph", nn, "hi2lo", ft.FP64, 5, 1, keep=True)
        # fresh copies transformed once (the timed tensors went through many steps)
        from fasttransforms_amd.utilities import fill, full_M
        A = fill("sph", nn, full_M("sph", nn), 2024)
        t32 = torch.from_numpy(np.ascontiguousarray(A.T)).float().cuda()
        t64 = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        ft.gpu_execute(P, 0, t32, precision=ft.FP32)
        ft.gpu_execute(P, 0, t64)
        ft.synchronize()
        a = t32.double().cpu().numpy()
        b = t64.cpu().numpy()
        rows["n%d" % nn] = {"fp32_ms": round(ms32, 5), "fp64_ms": round(ms64, 5),
                            "fp32_gflops": round(fl / (ms32 * 1e-3) / 1e9, 2),
                            "rel2_vs_fp64": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                            "relinf_vs_fp64": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}
    return rows


def batch(ft, torch, count=512, seed=4242, steps=5):
    """C4: `count` sph hi2lo transforms, n = round(exp(U(ln 32, ln 1024))).

    Batched: one ft_gpu_execute_batch launch; also captured in a HIP graph.
    Looped: one ft_gpu_execute per transform, called from a host loop with
    prepared arguments (no Python wrapper work inside the timed loop)."""
    from fasttransforms_amd.utilities import fill, rotations
    lib = ft._lib
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(32), np.log(1024), count))).astype(int)
    plans = {int(n): ft.plan_rotsphere(int(n)) for n in np.unique(ns)}
    rot = {n: rotations("sph", n, 2 * n - 1) for n in plans}
    flops = 6.0 * sum(rot[int(n)] for n in ns)
    T = [torch.from_numpy(np.ascontiguousarray(fill("sph", int(n), 2 * int(n) - 1, seed + i).T)).cuda()
         for i, n in enumerate(ns)]
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    vp = ctypes.c_void_p
    pls = (vp * count)(*[plans[int(n)].ptr for n in ns])
    ptrs = (vp * count)(*[t.data_ptr() for t in T])
    ms_ = (ctypes.c_int * count)(*[2 * int(n) - 1 for n in ns])

    def run_batch():
        rc = lib.ft_gpu_execute_batch(ctypes.cast(pls, vp), ft.SPH, 0, ft.FP64, ctypes.cast(ptrs, vp),
                                      ctypes.cast(ms_, vp), count, sp)
        if rc:
            raise BenchError("ft_gpu_execute_batch failed: %d" % rc)

    args = [(vp(plans[int(n)].ptr), vp(t.data_ptr()), 2 * int(n) - 1) for n, t in zip(ns, T)]
    fn = lib.ft_gpu_execute

    def run_loop():
        for p, a, m in args:
            rc = fn(p, ft.SPH, 0, ft.FP64, a, m, sp)
            if rc:
                raise BenchError("ft_gpu_execute failed: %d" % rc)

    ms_batch = events(ft, torch, stream, run_batch, steps, 2)
    ms_loop = events(ft, torch, stream, run_loop, max(2, steps // 2), 1)
    ms_graph = None
    try:
        # a batch plan owns its schedule, so the captured launch stays valid
        bp = ft.BatchPlan([plans[int(n)] for n in ns], "hi2lo", T)
        g = torch.cuda.CUDAGraph()
        cs = torch.cuda.Stream()
        cs.wait_stream(stream)
        with torch.cuda.graph(g, stream=cs):
            bp.execute(cs)
        torch.cuda.synchronize()
        ft.synchronize(cs)
        ms_graph = events(ft, torch, stream, g.replay, steps, 2)
        ft.synchronize(cs)
    except ft.GPUError as e:
        raise BenchError("graph-replayed batch reported invalid results: %s" % e)
    except RuntimeError as e:  # capture unavailable: the graph figure is optional
        print("bench.py: graph capture skipped: %s" % e, file=sys.stderr)
    return {"count": count, "seed": seed, "n_min": int(ns.min()), "n_max": int(ns.max()),
            "n_mean": round(float(ns.mean()), 1), "gflop": round(flops / 1e9, 3),
            "batched_ms": round(ms_batch, 4), "looped_ms": round(ms_loop, 4),
            "graph_ms": None if ms_graph is None else round(ms_graph, 4),
            "batched_transforms_per_s": round(count / (ms_batch * 1e-3), 1),
            "looped_transforms_per_s": round(count / (ms_loop * 1e-3), 1),
            "batched_over_looped": round(ms_loop / ms_batch, 3),
            "batched_gflops": round(flops / (ms_batch * 1e-3) / 1e9, 2)}


def batch64(ft, torch, kind, params, seed=5150, steps=3):
    """C5: 64 transforms with n uniform on [128, 1024], one launch vs a host loop."""
    from fasttransforms_amd.utilities import fill, full_M, rotations
    lib = ft._lib
    rng = np.random.default_rng(seed)
    ns = rng.integers(128, 1025, 64)
    plans = {int(n): _plan(ft, kind, int(n), params) for n in np.unique(ns)}
    flops = 6.0 * sum(rotations(kind, int(n), full_M(kind, int(n))) for n in ns)
    T = [torch.from_numpy(np.ascontiguousarray(fill(kind, int(n), full_M(kind, int(n)), seed + i).T)).cuda()
         for i, n in enumerate(ns)]
    stream = torch.cuda.current_stream()
    pl = [plans[int(n)] for n in ns]
    kid = {"sph": ft.SPH, "tri": ft.TRI, "disk": ft.DISK}[kind]
    sp = ctypes.c_void_p(stream.cuda_stream)
    args = [(ctypes.c_void_p(p.ptr), ctypes.c_void_p(t.data_ptr()), full_M(kind, p.n)) for p, t in zip(pl, T)]

    def run_loop():
        for p, a, m in args:
            rc = lib.ft_gpu_execute(p, kid, 0, ft.FP64, a, m, sp)
            if rc:
                raise BenchError("ft_gpu_execute failed: %d" % rc)

    ms_b = events(ft, torch, stream, lambda: ft.gpu_execute_batch(pl, 0, T, stream=stream), steps, 1)
    ms_l = events(ft, torch, stream, run_loop, steps, 1)
    return {"params": list(params), "gflop": round(flops / 1e9, 3), "batched_ms": round(ms_b, 4),
            "looped_ms": round(ms_l, 4), "batched_gflops": round(flops / (ms_b * 1e-3) / 1e9, 2),
            "batched_over_looped": round(ms_l / ms_b, 3)}


# ------------------------------------------------------------ CPU baseline --

def host_cpu():
    """CPU model and physical core count (lscpu), and the threads this run may use."""
    model, cores = None, None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        pairs = {tuple(l.split(",")[:2]) for l in out.splitlines() if l and not l.startswith("#")}
        cores = len(pairs) or None
        info = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in info.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    share = int(env) if env and env.isdigit() else allowed
    threads = max(1, min(cores or allowed, allowed, share))
    return {"model": model, "physical_cores": cores, "cpus_allowed": allowed, "threads": threads}


def _time_ref(O, P, kind, d, A, seconds, min_runs=5):
    """Median seconds of the reference's executor after FT_TIME's two warm-ups
    (ftutilities.h:220-226), over at least min_runs runs or `seconds`."""
    B = np.zeros(A.size)
    X = A.copy(order="F")
    for _ in range(2):
        O.ref_execute(P, d, X, B=B)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < min_runs:
        X = A.copy(order="F")
        t0 = time.perf_counter()
        O.ref_execute(P, d, X, B=B)
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    return float(np.median(times)), len(times)


def cpu_baseline(n, seconds):
    """The reference's dispatched executors (its compiled kernels) on this host."""
    import oracle as O
    from fasttransforms_amd.utilities import rotations
    if not O.have_ref():
        return None
    cpu = host_cpu()
    T = cpu["threads"]
    rows = {}

    def run(name, kind, params, nn, d, threads):
        O.ref_set_num_threads(threads)
        M = {"sph": 2 * nn - 1, "tri": nn, "disk": 4 * nn - 3}[kind]
        P = O.Plan(kind, nn, *params)
        A = O.fill(kind, nn, M, 99)
        med, runs = _time_ref(O, P, kind, d, A, seconds)
        fl = 6.0 * rotations(kind, nn, M)
        rows[name] = {"median_ms": round(med * 1e3, 3), "runs": runs, "threads": threads,
                      "gflops": round(fl / med / 1e9, 2)}
        return med, fl

    med, fl = run("sph_hi2lo_n%d" % n, "sph", (), n, 0, T)
    run("sph_hi2lo_n%d_1thread" % n, "sph", (), n, 0, 1)
    run("sph_lo2hi_n%d" % n, "sph", (), n, 1, T)
    run("sph_hi2lo_n256", "sph", (), 256, 0, T)
    run("tri_hi2lo_n%d" % n, "tri", (0.0, 0.0, 0.0), n, 0, T)
    run("disk_hi2lo_n%d" % n, "disk", (0.0, 0.0), n, 0, T)
    # C4's host loop: 512 transforms, plans built beforehand (untimed), one timed pass
    O.ref_set_num_threads(T)
    rng = np.random.default_rng(4242)
    ns = np.round(np.exp(rng.uniform(np.log(32), np.log(1024), 512))).astype(int)
    plans = {int(k): O.Plan("sph", int(k)) for k in np.unique(ns)}
    mats = [O.fill("sph", int(k), 2 * int(k) - 1, 4242 + i) for i, k in enumerate(ns)]
    t0 = time.perf_counter()
    for k, A in zip(ns, mats):
        O.ref_execute(plans[int(k)], 0, A)
    loop_s = time.perf_counter() - t0
    rows["batch512_host_loop"] = {"seconds": round(loop_s, 3), "threads": T,
                                  "transforms_per_s": round(512 / loop_s, 1)}
    return {"median_s": med, "flops": fl, "threads": T, "cpu": cpu, "rows": rows,
            "isa": O.REF_ISA_NAMES[O.ref_isa()]}


def profiled_traffic(engine, n):
    """HBM bytes per hi2lo launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a copy of
    known size) for this engine and n, from the summary committed under profiles/, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "rot_%s_sph_n%d.md" % (engine, n))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read(), re.M)
    return (int(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


TUNING_CASES = (("sph", 1024, "hi2lo", ()), ("sph", 1024, "lo2hi", ()), ("sph", 2048, "hi2lo", ()),
                ("sph", 2048, "lo2hi", ()), ("tri", 1024, "lo2hi", (0.0, 0.0, 0.0)),
                ("disk", 1024, "lo2hi", (0.0, 0.0)))


def tuning_report(ft, torch):
    """The tuned table's choice against the built-in one (band-systolic
    engine with whole chains of up to 16 bands; pipelined L 4, W 16), per
    case: both times and the tuned choice."""
    out = {}
    names = {ft.ENGINE_PIPELINE: "pipeline", ft.ENGINE_SYSTOLIC: "systolic"}
    for kind, nn, d, params in TUNING_CASES:
        row = {}
        for label, table in (("tuned", "compiled"), ("builtin", None)):
            ft.set_tuning_table(table)
            e, mp, L, W, wg = ft.tuned_choice(kind, 0, d, 0, nn)
            m_, _ = single(ft, torch, kind, nn, d, ft.FP64, 10, 2, params)
            row[label + "_ms"] = round(m_, 4)
            row[label + "_choice"] = {"engine": names[e], "maxpart": mp} if e == ft.ENGINE_SYSTOLIC else \
                {"engine": names[e], "L": L, "W": W}
        ft.set_tuning_table("compiled")
        row["speedup"] = round(row["builtin_ms"] / row["tuned_ms"], 3)
        out["%s_n%d_%s" % (kind, nn, d)] = row
    return out


def calib_copy(torch):
    """A 64 MiB device copy whose FETCH_SIZE / WRITE_SIZE calibrate the HBM counters."""
    src = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()


# the reference's published forward / backward times of the harmonic transforms
# (README.md:51-76, 18-thread Xeon W-2191B, (alpha, beta, gamma) = 0): the
# column "degree 1023" is n = 1024
PUBLISHED_S = {
    "sph2fourier": {512: 0.006277, 1024: 0.026590, 2048: 0.141962},
    "fourier2sph": {512: 0.006357, 1024: 0.027123, 2048: 0.145149},
    "tri2cheb": {1024: 0.019123, 2048: 0.113299},
    "cheb2tri": {1024: 0.019529, 2048: 0.114718},
    "disk2cxf": {1024: 0.064321, 2048: 0.369272},
    "cxf2disk": {1024: 0.064544, 2048: 0.372493},
}


def harmonic(ft, torch, n, transform, trans, steps, warmup):
    """ms per ft_gpu_execute_harmonic on a device-resident FP64 matrix of the transform's full width."""
    from fasttransforms_amd.utilities import fill
    if transform in ("tri2cheb", "cheb2tri"):
        H, kind = ft.plan_tri2cheb(n, 0.0, 0.0, 0.0), "tri"
    elif transform in ("disk2cxf", "cxf2disk"):
        H, kind = ft.plan_disk2cxf(n, 0.0, 0.0), "disk"
    elif transform in ("rectdisk2cheb", "cheb2rectdisk"):
        H, kind = ft.plan_rectdisk2cheb(n, 0.0), "tri"
    elif transform in ("ann2cxf", "cxf2ann"):  # the reference test's annulus (test_drivers.c:21)
        H, kind = ft.plan_ann2cxf(n, 0.0, 0.0, 0.0, float(np.sqrt(1.0 / 3.0))), "disk"
    else:
        H, kind = ft.plan_sph2fourier(n), "sph"
    M = {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]
    T = torch.from_numpy(np.ascontiguousarray(fill(kind, n, M, 2024).T)).cuda()
    stream = torch.cuda.current_stream()
    return events(ft, torch, stream, lambda: ft.gpu_execute_harmonic(transform, trans, H, T, stream=stream),
                  steps, warmup)


def harmonic_rows(ft, torch, n):
    """The harmonic transforms at n against the reference's published times."""
    out = {}
    for transform in ("sph2fourier", "fourier2sph", "tri2cheb", "cheb2tri", "disk2cxf", "cxf2disk", "ann2cxf",
                      "cxf2ann"):
        ms = harmonic(ft, torch, n, transform, "N", 10, 2)
        out["%s_ms" % transform] = round(ms, 4)
        pub = PUBLISHED_S.get(transform, {}).get(n)
        if pub:
            out["%s_published_ms" % transform] = round(pub * 1e3, 3)
            out["%s_vs_published" % transform] = round(pub * 1e3 / ms, 1)
    return out


def run_case(ft, torch, spec, steps, warmup):
    """--case: one configuration only, for profiling runs."""
    parts = spec.split(":")
    if parts[0] in PUBLISHED_S or parts[0] in ("sphv2fourier", "fourier2sphv", "rectdisk2cheb", "cheb2rectdisk",
                                               "ann2cxf", "cxf2ann"):
        ms = harmonic(ft, torch, int(parts[1]), parts[0], parts[2] if len(parts) > 2 else "N", steps, warmup)
        print(json.dumps({"case": spec, "ms": round(ms, 5)}), flush=True)
        return
    if parts[0] == "grid":
        # grid:WHICH:N[:TRANS], WHICH one of sph_synthesis, sph_analysis, disk_synthesis, ...
        which, nn = parts[1], int(parts[2])
        trans = parts[3] if len(parts) > 3 else "N"
        geo = which.split("_")[0]
        M = 4 * nn - 3 if geo == "disk" else 2 * nn - 1
        syn = which.endswith("synthesis") == (trans == "N")
        P = getattr(ft, "plan_%s_%s" % (geo, "synthesis" if syn else "analysis"))(nn, M)
        T = torch.rand(M, nn, dtype=torch.float64, device="cuda")
        stream = torch.cuda.current_stream()
        ms = events(ft, torch, stream, lambda: ft.gpu_execute_sph_grid(which, trans, P, T, stream=stream), steps, warmup)
        print(json.dumps({"case": spec, "ms": round(ms, 5), "GB/s": round(16.0 * nn * M / (ms * 1e-3) / 1e9, 1)}),
              flush=True)
        return
    if parts[0] == "spinsph":
        # spinsph:N:DIR:S
        nn, d, sp = int(parts[1]), parts[2], int(parts[3]) if len(parts) > 3 else 0
        P = ft.plan_rotspinsphere(nn, sp)
        T = torch.rand(2 * nn - 1, nn, dtype=torch.complex128, device="cuda")
        stream = torch.cuda.current_stream()
        ms = events(ft, torch, stream, lambda: ft.gpu_execute_spinsph(P, d, T, stream=stream), steps, warmup)
        print(json.dumps({"case": spec, "ms": round(ms, 5)}), flush=True)
        return
    kind, nn, d = parts[0], int(parts[1]), parts[2]
    if kind == "tet":
        m_, fl = tet(ft, torch, nn, d, steps, warmup)
        print(json.dumps({"case": spec, "ms": round(m_, 5), "gflops": round(fl / (m_ * 1e-3) / 1e9, 2)}), flush=True)
        return
    precision = ft.FP32 if "fp32" in parts[3:] else ft.FP64
    params = PARAMS[kind]
    if "batch" in parts[3:]:
        res = batch(ft, torch, steps=steps)
        print(json.dumps({"case": spec, "batched_ms": res["batched_ms"]}), flush=True)
        return
    m_, fl = single(ft, torch, kind, nn, d, precision, steps, warmup, params)
    print(json.dumps({"case": spec, "ms": round(m_, 5), "gflops": round(fl / (m_ * 1e-3) / 1e9, 2),
                      "engine": ft.last_engine()}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the rotation stage runs on one GPU; "
              "--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import fasttransforms_amd as ft

    if args.case:
        run_case(ft, torch, args.case, args.steps, args.warmup)
        if args.calib_copy:
            calib_copy(torch)
        return

    n = args.n
    M = 2 * n - 1
    engines = {"auto": ft.ENGINE_AUTO, "pipeline": ft.ENGINE_PIPELINE, "systolic": ft.ENGINE_SYSTOLIC}
    names = {ft.ENGINE_PIPELINE: "pipeline", ft.ENGINE_SYSTOLIC: "systolic"}
    ft.set_engine(engines[args.engine])

    parity = gpu_vs_reference(ft, torch, n)
    if parity is not None and max(parity.values()) > 1e-14:
        print("bench.py: GPU output differs from the reference's executor: %s" % parity, file=sys.stderr)
        sys.exit(3)

    ms, flops = single(ft, torch, "sph", n, "hi2lo", ft.FP64, args.steps, args.warmup)
    engine_used = names.get(ft.last_engine(), "none")
    ms_lo, _ = single(ft, torch, "sph", n, "lo2hi", ft.FP64, max(5, args.steps // 2), 2)
    gflops = flops / (ms * 1e-3) / 1e9
    achieved_tflops = flops / (ms * 1e-3) / 1e12
    compulsory_bytes = 16.0 * n * M + 16.0 * n * (n + 1) / 2

    extra = {}
    if not args.no_batch:
        # both engines on the headline shape and at n/2
        for eng in ("pipeline", "systolic"):
            ft.set_engine(engines[eng])
            for nn in (n, n // 2):
                try:
                    m_e, _ = single(ft, torch, "sph", nn, "hi2lo", ft.FP64, 10, 2)
                    extra["sph_n%d_hi2lo_ms_%s" % (nn, eng)] = round(m_e, 4)
                except ft.GPUError as e:  # an engine that does not apply to the shape
                    extra["sph_n%d_hi2lo_ms_%s" % (nn, eng)] = None
                    print("bench.py: %s engine at n=%d: %s" % (eng, nn, e), file=sys.stderr)
        ft.set_engine(engines[args.engine])
        ms2, fl2 = single(ft, torch, "sph", 2 * n, "hi2lo", ft.FP64, 5, 2)
        ms2l, _ = single(ft, torch, "sph", 2 * n, "lo2hi", ft.FP64, 5, 2)
        mst, flt = single(ft, torch, "tri", n, "hi2lo", ft.FP64, 10, 2, (0.0, 0.0, 0.0))
        msth, _ = single(ft, torch, "tri", n, "hi2lo", ft.FP64, 10, 2, (-0.5, -0.5, -0.5))
        msd, fld = single(ft, torch, "disk", n, "hi2lo", ft.FP64, 10, 2, (0.0, 0.0))
        for kind in ("sphv", "rectdisk", "annulus"):
            m_k, fl_k = single(ft, torch, kind, n, "hi2lo", ft.FP64, 10, 2, PARAMS[kind])
            extra["%s_hi2lo_ms" % kind] = round(m_k, 4)
            extra["%s_hi2lo_gflops" % kind] = round(fl_k / (m_k * 1e-3) / 1e9, 2)
        extra.update({
            "sph_n256_roundtrip": roundtrip_256(ft, torch),
            "sph_n%d_hi2lo_ms" % (2 * n): round(ms2, 4),
            "sph_n%d_hi2lo_gflops" % (2 * n): round(fl2 / (ms2 * 1e-3) / 1e9, 2),
            "sph_n%d_lo2hi_ms" % (2 * n): round(ms2l, 4),
            "fp32": fp32_rows(ft, torch, (n // 2, n, 2 * n)),
            "tri_hi2lo_ms": round(mst, 4), "tri_hi2lo_gflops": round(flt / (mst * 1e-3) / 1e9, 2),
            "tri_m05_hi2lo_ms": round(msth, 4), "tri_m05_hi2lo_gflops": round(flt / (msth * 1e-3) / 1e9, 2),
            "disk_hi2lo_ms": round(msd, 4), "disk_hi2lo_gflops": round(fld / (msd * 1e-3) / 1e9, 2),
            "harmonic_n%d" % n: harmonic_rows(ft, torch, n),
            "tet": tet_rows(ft, torch),
            "spinsph": spin_rows(ft, torch),
            "grid": grid_rows(ft, torch, n),
            "batch_sph_hi2lo_512": batch(ft, torch),
            "batch64_tri": batch64(ft, torch, "tri", (0.0, 0.0, 0.0)),
            "batch64_tri_m05": batch64(ft, torch, "tri", (-0.5, -0.5, -0.5)),
            "batch64_disk": batch64(ft, torch, "disk", (0.0, 0.0)),
            "tuning": tuning_report(ft, torch),
        })

    if args.calib_copy:
        calib_copy(torch)

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, args.cpu_seconds)
    cpu_obj = None
    if cpu:
        c = cpu["cpu"]
        cpu_obj = {"value": round(cpu["flops"] / cpu["median_s"] / 1e9, 2), "unit": "GFLOP/s", "cores": cpu["threads"],
                   "kind": "reference",
                   "sample": "the reference's ft_execute_sph_hi2lo executor (%s kernels compiled from its sources), "
                             "n=%d, median of %d runs after 2 warm-ups: %.1f ms on %d threads; host %s, %s physical "
                             "cores, %d CPUs allowed to this run" % (
                                 cpu["isa"], n, cpu["rows"]["sph_hi2lo_n%d" % n]["runs"], cpu["median_s"] * 1e3,
                                 cpu["threads"], c["model"], c["physical_cores"], c["cpus_allowed"]),
                   "rows": cpu["rows"]}
    config = {"workload": "ft_execute_sph_hi2lo, n=%d (%dx%d column-major FP64), device-resident" % (n, n, M),
              "n": n, "M": M, "precision": "fp64", "engine": engine_used, "lo2hi_ms": round(ms_lo, 5),
              "lo2hi_gflops": round(flops / (ms_lo * 1e-3) / 1e9, 2), "rel2_vs_reference": parity}
    config.update(extra)
    traffic, traffic_src = profiled_traffic(engine_used, n)
    line = {
        "metric": "sph_hi2lo FP64 GFLOP/s & % FP64 peak at n=1024; batched vs looped transforms/s",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U(-1,1) valid-triangle fill)",
        "config": config,
        "roofline": {"bound": "fp64_valu", "achieved": round(achieved_tflops, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "note": "compute-bound (51 flop/B against a 9.8 flop/B ridge) on the FP64 vector ALU "
                             "(v_mul_f64/v_fma_f64, no MFMA); 6 algorithmic flops per rotation take 4 FP64 "
                             "instructions, so the instruction mix caps the rate at 75% of peak",
                     "compulsory_bytes": compulsory_bytes,
                     "algorithmic_flops": flops},
        "cpu_baseline": cpu_obj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    try:
        main()
    except BenchError as e:
        print("bench.py: %s" % e, file=sys.stderr)
        sys.exit(4)
