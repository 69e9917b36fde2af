"""Benchmark of the rotation stage on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu-baseline]
                    [--no-batch] [--calib-copy]

A step is one ft_execute_sph_hi2lo of the 1024 x 2047 FP64 spherical-harmonic
matrix (degree 1023) held in HBM, launched with ft_gpu_execute on the current
stream.  The timed region is K back-to-back steps between two HIP events on
that stream, bracketed by device synchronisation.  Besides that headline the
JSON line carries, under "config": lo2hi at n = 1024, hi2lo at n = 2048, the
FP32 variant at n = 1024, tri and disk at n = 1024, and the batch of 512
transforms of log-uniform degree in [32, 1024] run as one launch (plain and
captured in a HIP graph) against a host loop over the single-transform entry
point.  It also carries the roofline of the rotation kernel and the reference's
CPU executors (their own compiled kernels) timed on this host on a bounded
sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 peak (vector and matrix alike)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-batch", action="store_true", help="skip the secondary measurements")
    p.add_argument("--engine", choices=("auto", "pipeline", "systolic"), default="auto",
                   help="engine of the headline launches (auto: the library's own choice)")
    p.add_argument("--calib-copy", action="store_true",
                   help="also run a plain 64 MiB device copy (untimed), to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return p.parse_args()


def events(torch, stream, fn, steps, warmup):
    """ms per call of fn() on `stream`, between HIP events on that stream."""
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
    return start.elapsed_time(stop) / steps


def single(ft, torch, kind, n, direction, precision, steps, warmup, params=()):
    from fasttransforms_amd.utilities import fill, full_M, rotations
    M = full_M(kind, n)
    P = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind](n, *params)
    dtype = torch.float64 if precision == ft.FP64 else torch.float32
    T = torch.from_numpy(np.ascontiguousarray(fill(kind, n, M, 2024).T)).to(dtype).cuda()  # (M, n): row j = column j
    stream = torch.cuda.current_stream()
    ms = events(torch, stream, lambda: ft.gpu_execute(P, direction, T, precision=precision, stream=stream),
                steps, warmup)
    return ms, 6.0 * rotations(kind, n, M)


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
            raise RuntimeError("ft_gpu_execute_batch failed: %d" % rc)

    args = [(vp(plans[int(n)].ptr), vp(t.data_ptr()), 2 * int(n) - 1) for n, t in zip(ns, T)]
    fn = lib.ft_gpu_execute

    def run_loop():
        for p, a, m in args:
            rc = fn(p, ft.SPH, 0, ft.FP64, a, m, sp)
            if rc:
                raise RuntimeError("ft_gpu_execute failed: %d" % rc)

    ms_batch = events(torch, stream, run_batch, steps, 2)
    ms_loop = events(torch, stream, run_loop, max(2, steps // 2), 1)
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
        ms_graph = events(torch, stream, g.replay, steps, 2)
    except Exception as e:  # the graph figure is optional for the line
        print("bench.py: graph capture skipped: %s" % e, file=sys.stderr)
    return {"count": count, "seed": seed, "n_min": int(ns.min()), "n_max": int(ns.max()),
            "n_mean": round(float(ns.mean()), 1), "gflop": round(flops / 1e9, 3),
            "batched_ms": round(ms_batch, 4), "looped_ms": round(ms_loop, 4),
            "graph_ms": None if ms_graph is None else round(ms_graph, 4),
            "batched_transforms_per_s": round(count / (ms_batch * 1e-3), 1),
            "looped_transforms_per_s": round(count / (ms_loop * 1e-3), 1),
            "batched_over_looped": round(ms_loop / ms_batch, 3)}


def cpu_baseline(n, seconds):
    """The reference's dispatched executor (its compiled kernels) on host cores."""
    import oracle as O
    if not O.have_ref():
        return None
    threads = min(16, os.cpu_count() or 1)
    O.ref_set_num_threads(threads)
    P = O.Plan("sph", n)
    M = 2 * n - 1
    A = O.fill("sph", n, M, 99)
    B = np.zeros(n * M)
    for _ in range(2):  # FT_TIME's two warm-ups (ftutilities.h:220-226)
        O.ref_execute(P, 0, A, B=B)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 5:
        t0 = time.perf_counter()
        O.ref_execute(P, 0, A, B=B)
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    return {"median_s": float(np.median(times)), "runs": len(times), "threads": threads,
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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the rotation stage runs on one GPU; "
              "--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import fasttransforms_amd as ft

    n = args.n
    M = 2 * n - 1
    engines = {"auto": ft.ENGINE_AUTO, "pipeline": ft.ENGINE_PIPELINE, "systolic": ft.ENGINE_SYSTOLIC}
    names = {ft.ENGINE_PIPELINE: "pipeline", ft.ENGINE_SYSTOLIC: "systolic"}
    ft.set_engine(engines[args.engine])
    ms, flops = single(ft, torch, "sph", n, "hi2lo", ft.FP64, args.steps, args.warmup)
    engine_used = names.get(ft.last_engine(), "none")
    ms_lo, _ = single(ft, torch, "sph", n, "lo2hi", ft.FP64, max(5, args.steps // 2), 2)
    gflops = flops / (ms * 1e-3) / 1e9
    achieved_tflops = flops / (ms * 1e-3) / 1e12
    compulsory_bytes = 16.0 * n * M + 16.0 * n * (n + 1) / 2

    extra = {}
    if not args.no_batch:
        # both engines on the headline shape and at n/2 (where auto picks the systolic one)
        for eng in ("pipeline", "systolic"):
            ft.set_engine(engines[eng])
            for nn in (n, n // 2):
                try:
                    m_e, _ = single(ft, torch, "sph", nn, "hi2lo", ft.FP64, 10, 2)
                    extra["sph_n%d_hi2lo_ms_%s" % (nn, eng)] = round(m_e, 4)
                except Exception as e:  # an engine that does not apply to the shape
                    extra["sph_n%d_hi2lo_ms_%s" % (nn, eng)] = None
                    print("bench.py: %s engine at n=%d: %s" % (eng, nn, e), file=sys.stderr)
        ft.set_engine(engines[args.engine])
        ms2, fl2 = single(ft, torch, "sph", 2 * n, "hi2lo", ft.FP64, 5, 2)
        ms32, _ = single(ft, torch, "sph", n, "hi2lo", ft.FP32, 10, 2)
        mst, flt = single(ft, torch, "tri", n, "hi2lo", ft.FP64, 10, 2, (0.0, 0.0, 0.0))
        msd, fld = single(ft, torch, "disk", n, "hi2lo", ft.FP64, 10, 2, (0.0, 0.0))
        extra.update({"sph_n%d_hi2lo_ms" % (2 * n): round(ms2, 4),
                 "sph_n%d_hi2lo_gflops" % (2 * n): round(fl2 / (ms2 * 1e-3) / 1e9, 2),
                 "sph_fp32_hi2lo_ms": round(ms32, 4),
                 "tri_hi2lo_ms": round(mst, 4), "tri_hi2lo_gflops": round(flt / (mst * 1e-3) / 1e9, 2),
                 "disk_hi2lo_ms": round(msd, 4), "disk_hi2lo_gflops": round(fld / (msd * 1e-3) / 1e9, 2),
                 "batch_sph_hi2lo_512": batch(ft, torch)})

    if args.calib_copy:
        src = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, args.cpu_seconds)
    cpu_obj = None
    if cpu:
        cpu_obj = {"value": flops / cpu["median_s"] / 1e9, "unit": "GFLOP/s", "cores": cpu["threads"],
                   "kind": "reference",
                   "sample": "%d runs of the reference's ft_execute_sph_hi2lo executor (%s kernels compiled "
                             "from its sources) at n=%d, median %.1f ms" % (cpu["runs"], cpu["isa"], n,
                                                                           cpu["median_s"] * 1e3)}
    config = {"workload": "ft_execute_sph_hi2lo, n=%d (%dx%d column-major FP64), device-resident" % (n, n, M),
              "n": n, "M": M, "precision": "fp64", "engine": engine_used, "lo2hi_ms": round(ms_lo, 5),
              "lo2hi_gflops": round(flops / (ms_lo * 1e-3) / 1e9, 2)}
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
        "roofline": {"bound": "mfma", "achieved": round(achieved_tflops, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "note": "compute-bound (51 flop/B against a 9.8 flop/B ridge) on the FP64 vector ALU; "
                             "MI355X's FP64 peak is 78.6 TFLOP/s for vector and matrix alike",
                     "compulsory_bytes": compulsory_bytes,
                     "algorithmic_flops": flops},
        "cpu_baseline": cpu_obj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
