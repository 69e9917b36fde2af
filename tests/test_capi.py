"""A C program compiled against include/fasttransforms_amd.h and linked with
-lfasttransforms_amd, as the reference's callers are: it builds and links
here (every declared entry point resolves), and on the GPU it runs the
reference's own round-trip and single-column accuracy checks
(test_drivers.c:37-57, test_rotations.c:25-39) and its harmonic-transform
round trips (test_drivers.c:226-248, 700-720, 937-957, 1128-1156,
1234-1254)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "fasttransforms_amd")


def _build(tmp_path):
    exe = str(tmp_path / "test_capi")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "test_capi.c"), "-L", LIBDIR, "-lfasttransforms_amd",
           "-Wl,-rpath," + LIBDIR, "-lm", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_c_caller_builds_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_c_caller_round_trips(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=100)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
