"""Autotuner of the rotation stage on the GPU.

For every kind (sph, tri, disk), precision (FP64, FP32), direction and a set
of degrees it times the candidate engines and shapes -- the band-systolic
engine with several chain-part sizes, the pipelined engine with several
(L, W) -- on the same synthetic matrix, checks that every candidate gives
bitwise the same result as the first, and keeps the fastest.  Batches of
transforms of log-uniform degree (the shape of the reference's per-order
callers) are tuned the same way, per largest degree.  The result is a table
of rows (kind, precision, direction, batch, nlo, nhi, engine, maxpart, L, W,
wg_per_cu): written as JSON (the measurements) and as the C header the
library compiles in (csrc/ft_tuned_gfx950.h), and installed in the running
library (ft_gpu_set_tuning_table) so it takes effect at once.

    python -m fasttransforms_amd.tune --json out.json --header out.h [--quick]

No reference counterpart: the reference selects its SIMD kernel by CPU
feature (src/drivers.c:58-110) and has no tuning table.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

import fasttransforms_amd as ft

KINDS = {"sph": (ft.SPH, ()), "tri": (ft.TRI, (0.0, 0.0, 0.0)), "disk": (ft.DISK, (0.0, 0.0))}
SIZES = [64, 128, 256, 512, 1024, 2048]
QUICK_SIZES = [256, 1024]
BATCH_NMAX = [256, 1024]
SYS_PARTS = [16, 12, 8]
PIPE_SHAPES = [(4, 16), (2, 16), (8, 8), (2, 8)]
BATCH_SHAPES = [(4, 16, 2), (4, 16, 1), (2, 16, 2), (8, 8, 2)]
NMAX_INT = 2 ** 31 - 1


def _plan(kind, n):
    code, params = KINDS[kind]
    maker = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]
    return maker(n, *params)


def _full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]


def _apply(cand):
    """Sets the knobs of a candidate: ("systolic", maxpart) or ("pipeline", L, W, wg_per_cu)."""
    if cand[0] == "systolic":
        ft.set_engine(ft.ENGINE_SYSTOLIC)
        ft.set_systolic(cand[1])
        ft.set_tuning(0, 0, 0)
    else:
        ft.set_engine(ft.ENGINE_PIPELINE)
        ft.set_systolic(0)
        ft.set_tuning(cand[1], cand[2], cand[3] if len(cand) > 3 else 0)


def _reset():
    ft.set_engine(ft.ENGINE_AUTO)
    ft.set_systolic(0)
    ft.set_tuning(0, 0, 0)


def _time(torch, run, min_ms=3.0, reps=3):
    """Median over reps of the mean time of enough launches to fill min_ms."""
    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    run()
    e1.record()
    torch.cuda.synchronize()
    k = max(2, min(50, int(math.ceil(min_ms / max(e0.elapsed_time(e1), 1e-3)))))
    out = []
    for _ in range(reps):
        e0.record()
        for _ in range(k):
            run()
        e1.record()
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / k)
    ft.synchronize()
    return float(np.median(out))


def _sweep(torch, label, cands, make_inputs, run):
    """Times every candidate on the same inputs; all must agree bitwise."""
    ref = None
    rows = []
    for cand in cands:
        _apply(cand)
        try:
            data = make_inputs()
            run(data)
            ft.synchronize()
            got = [t.cpu() for t in data]
            same = ref is None or all(torch.equal(a, b) for a, b in zip(got, ref))
            if ref is None:
                ref = got
            ms = _time(torch, lambda: run(data))
            rows.append({"cand": list(cand), "ms": ms, "bitwise_same": bool(same)})
        except ft.GPUError as e:
            rows.append({"cand": list(cand), "ms": None, "error": str(e)})
        finally:
            _reset()
        print("%s %-22s %s" % (label, cand, "%.4f ms" % rows[-1]["ms"] if rows[-1].get("ms") else rows[-1]),
              flush=True)
    return rows


def _best(rows):
    ok = [r for r in rows if r.get("ms") is not None and r.get("bitwise_same")]
    if len(ok) != len([r for r in rows if r.get("ms") is not None]):
        raise RuntimeError("a candidate differs from the others: %r" % rows)
    return min(ok, key=lambda r: r["ms"]) if ok else None


def _row(kind, precision, direction, batch, nlo, nhi, cand):
    code = KINDS[kind][0]
    if cand[0] == "systolic":
        return [code, precision, direction, batch, nlo, nhi, ft.ENGINE_SYSTOLIC, cand[1], 0, 0, 0]
    return [code, precision, direction, batch, nlo, nhi, ft.ENGINE_PIPELINE, 0, cand[1], cand[2],
            cand[3] if len(cand) > 3 else 0]


def _bands(sizes):
    """Degree ranges nearest (geometrically) to each measured size."""
    out = []
    for k, s in enumerate(sizes):
        lo = 1 if k == 0 else int(math.sqrt(sizes[k - 1] * s)) + 1
        hi = NMAX_INT if k == len(sizes) - 1 else int(math.sqrt(s * sizes[k + 1]))
        out.append((lo, hi))
    return out


def tune(torch, sizes=SIZES, batch_nmax=BATCH_NMAX, kinds=("sph", "tri", "disk"), precisions=(0, 1),
         directions=(0, 1), batch_count=256):
    props = torch.cuda.get_device_properties(0)
    record = {"device": "%s, %s, %d CUs" % (torch.cuda.get_device_name(0), getattr(props, "gcnArchName", "?"),
                                            props.multi_processor_count), "single": [], "batch": []}
    table = []
    g = torch.Generator().manual_seed(2024)
    for kind in kinds:
        for precision in precisions:
            dt = torch.float64 if precision == 0 else torch.float32
            for direction in directions:
                bands = _bands(sizes)
                for n, (lo, hi) in zip(sizes, bands):
                    P = _plan(kind, n)
                    M = _full_M(kind, n)
                    X = (torch.rand(M, n, generator=g, dtype=torch.float64) - 0.5).to(dt).cuda()
                    cands = [("systolic", p) for p in SYS_PARTS] + [("pipeline", L, W) for L, W in PIPE_SHAPES]
                    rows = _sweep(torch, "%s p%d d%d n=%d" % (kind, precision, direction, n), cands,
                                  lambda: [X.clone()],
                                  lambda d: ft.gpu_execute(P, direction, d[0], precision=precision))
                    best = _best(rows)
                    record["single"].append({"kind": kind, "precision": precision, "direction": direction, "n": n,
                                             "candidates": rows, "best": best})
                    if best:
                        table.append(_row(kind, precision, direction, 0, lo, hi, best["cand"]))
                bands = _bands(batch_nmax)
                for nmax, (lo, hi) in zip(batch_nmax, bands):
                    rng = np.random.default_rng(4242 + nmax)
                    ns = np.round(np.exp(rng.uniform(np.log(32), np.log(nmax), batch_count))).astype(int)
                    plans = [_plan(kind, int(n)) for n in ns]
                    Xs = [(torch.rand(_full_M(kind, int(n)), int(n), generator=g, dtype=torch.float64) - 0.5)
                          .to(dt).cuda() for n in ns]
                    cands = [("pipeline",) + s for s in BATCH_SHAPES] + [("systolic", p) for p in (16, 8)]
                    rows = _sweep(torch, "%s p%d d%d batch nmax=%d" % (kind, precision, direction, nmax), cands,
                                  lambda: [x.clone() for x in Xs],
                                  lambda d: ft.gpu_execute_batch(plans, direction, d, precision=precision))
                    best = _best(rows)
                    record["batch"].append({"kind": kind, "precision": precision, "direction": direction,
                                            "nmax": nmax, "count": batch_count, "candidates": rows, "best": best})
                    if best:
                        table.append(_row(kind, precision, direction, 1, lo, hi, best["cand"]))
    record["table"] = table
    return record


HEADER = """/* ft_tuned_gfx950.h -- the autotuning table for gfx950 (MI355X), written by
 * fasttransforms_amd/tune.py from a sweep on the device (%s); the row with
 * kind -1 ends it.  Columns: kind (0 sph, 1 tri, 2 disk), precision
 * (0 FP64, 1 FP32), direction (0 hi2lo, 1 lo2hi), batch, degree range
 * nlo..nhi, engine (1 pipelined, 2 band-systolic), band-systolic bands per
 * chain part, pipelined L and W, workgroups per CU (0: built-in). */
#define FT_TUNED_ARCH "gfx950"
static const ft_tune_row ft_tuned_gfx950[] = {
%s    {-1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
};
"""


def header_text(table, device):
    body = "".join("    {%s},\n" % ", ".join(str(v) for v in row) for row in table)
    return HEADER % (device, body)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--json", help="measurement record to write")
    ap.add_argument("--header", help="C header to write (csrc/ft_tuned_gfx950.h)")
    ap.add_argument("--quick", action="store_true", help="two degrees, one batch size")
    args = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("tune: no GPU", file=sys.stderr)
        return 2
    t0 = time.time()
    rec = tune(torch, sizes=QUICK_SIZES if args.quick else SIZES,
               batch_nmax=[1024] if args.quick else BATCH_NMAX)
    rec["seconds"] = time.time() - t0
    ft.set_tuning_table(rec["table"])
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rec, f, indent=1)
    if args.header:
        with open(args.header, "w") as f:
            f.write(header_text(rec["table"], rec["device"]))
    print("tune: %d rows in %.0f s" % (len(rec["table"]), rec["seconds"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
