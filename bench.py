"""Benchmark of the rotation stage on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu-baseline]

A step is one ft_execute_sph_hi2lo of the 1024 x 2047 FP64 spherical-harmonic
matrix (degree 1023) held in HBM, launched with ft_gpu_execute on the current
stream.  The timed region is K back-to-back steps between two HIP events on
that stream, bracketed by device synchronisation.  Prints one JSON line with
the GFLOP/s, the roofline of the rotation kernel and the reference's CPU
executors (their own compiled kernels) timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--calib-copy", action="store_true",
                   help="also run a plain 64 MiB device copy (untimed), to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return p.parse_args()


def time_gpu(ft, torch, P, T, direction, steps, warmup):
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        ft.gpu_execute(P, direction, T, stream=stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        ft.gpu_execute(P, direction, T, stream=stream)
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per launch


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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the rotation stage runs on one GPU; "
              "--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import fasttransforms_amd as ft
    from fasttransforms_amd.utilities import fill, rotations

    n = args.n
    M = 2 * n - 1
    flops = 6.0 * rotations("sph", n, M)
    P = ft.plan_rotsphere(n)
    A = fill("sph", n, M, 2024)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()  # (M, n): row j = column j
    ms = time_gpu(ft, torch, P, T, "hi2lo", args.steps, args.warmup)
    ms_lo = time_gpu(ft, torch, P, T, "lo2hi", max(5, args.steps // 2), 2)
    gflops = flops / (ms * 1e-3) / 1e9
    achieved_tflops = flops / (ms * 1e-3) / 1e12
    compulsory_bytes = 16.0 * n * M + 16.0 * n * (n + 1) / 2

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
        "config": {"workload": "ft_execute_sph_hi2lo, n=%d (%dx%d column-major FP64), device-resident" % (n, n, M),
                   "n": n, "M": M, "precision": "fp64", "lo2hi_ms": round(ms_lo, 5),
                   "lo2hi_gflops": round(flops / (ms_lo * 1e-3) / 1e9, 2)},
        "roofline": {"bound": "fp64_vector", "achieved": round(achieved_tflops, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": None,
                     "compulsory_bytes": compulsory_bytes,
                     "algorithmic_flops": flops},
        "cpu_baseline": cpu_obj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
