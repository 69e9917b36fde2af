"""Write the golden vectors of the rotation stage from the reference's own code.

    make -C oracle all ref        # needs the reference tree (development container)
    python tests/golden/make_golden.py

For every case the input is the seeded valid-triangle fill (oracle.fill, a
SplitMix64 stream, so it is regenerated in the tests, not stored; its SHA-256
is stored to detect drift).  The outputs come from oracle/_ref/libftref.so,
i.e. the reference's compiled kernels dispatched as src/drivers.c dispatches
them:

* ``default``: the scalar executors (drivers.c:94-110, 314-326, 414-430),
  compiled exactly as the reference's Make.inc compiles them -- host
  independent bits;
* ``simd``: the executor the host's cpuid selects (AVX512F on AVX-512 hosts,
  drivers.c:58-74); the ISA name is recorded.

Output: tests/golden/ref_vectors.npz and tests/golden/manifest.json.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

CASES = [
    ("sph", (), 5), ("sph", (), 16), ("sph", (), 33), ("sph", (), 64),
    ("tri", (0.0, 0.0, 0.0), 5), ("tri", (0.0, 0.0, 0.0), 16), ("tri", (0.0, 0.0, 0.0), 33),
    ("tri", (-0.5, -0.5, -0.5), 16), ("tri", (-0.5, -0.5, -0.5), 33),
    ("disk", (0.0, 0.0), 5), ("disk", (0.0, 0.0), 16), ("disk", (0.0, 0.0), 33),
]
SEED = 20240611


def full_M(kind, n):
    return {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]


def case_key(kind, params, n):
    tag = "" if not params else "_" + "_".join("%g" % p for p in params)
    return "%s%s_n%d" % (kind, tag, n)


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libftref.so is missing: run `make -C oracle ref` first")
    isa = O.ref_isa()
    arrays, manifest = {}, {"seed": SEED, "simd_isa": O.REF_ISA_NAMES[isa], "cases": {}}
    for kind, params, n in CASES:
        M = full_M(kind, n)
        P = O.Plan(kind, n, *params)
        A = O.fill(kind, n, M, SEED + n)
        key = case_key(kind, params, n)
        entry = {"kind": kind, "params": list(params), "n": n, "M": M, "seed": SEED + n,
                 "input_sha256": hashlib.sha256(A.tobytes(order="F")).hexdigest()}
        for d, dname in ((0, "hi2lo"), (1, "lo2hi")):
            for variant, isa_sel in (("default", 0), ("simd", -1)):
                X = A.copy(order="F")
                O.ref_execute(P, d, X, isa=isa_sel)
                k = "%s_%s_%s" % (key, dname, variant)
                arrays[k] = np.ascontiguousarray(X.T)  # stored (M, n): row j is column j
                entry[k] = hashlib.sha256(arrays[k].tobytes()).hexdigest()
        manifest["cases"][key] = entry
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **arrays)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote %d arrays (simd = %s)" % (len(arrays), manifest["simd_isa"]))


if __name__ == "__main__":
    main()
