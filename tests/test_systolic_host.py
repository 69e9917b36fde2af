"""Host side of the band-systolic engine, checked on the CPU: the classes,
band streams, step-major (s, c) tables and chains the library builds for an
n x M matrix, run through a step-by-step emulation of the kernel
(tests/sys_emulator.py), must reproduce the oracle's executors.  Covers one
and several bands per chain, partial last bands, column masks across band
edges, partial M, both directions and both band widths (S = 2, 4).  The
emulation rounds c*x + s*y in two steps where the kernel fuses them, so the
bar is a relative 2-norm of 1e-14 here; the GPU tests hold the kernel itself
to the oracle's bits."""
import numpy as np
import pytest

import oracle as O
import sys_emulator as se

CASES = [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)), ("disk", (0.0, 0.0))]


def _plan(ft, kind, n, params):
    maker = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]
    return maker(n, *params), O.Plan(kind, n, *params)


@pytest.mark.parametrize("kind,params", CASES)
@pytest.mark.parametrize("n,S,M", [(3, 2, None), (20, 2, None), (100, 2, None), (100, 4, None), (77, 2, 31),
                                   (150, 2, None), (150, 4, 91)])
@pytest.mark.parametrize("d", [0, 1])
def test_host_tables_and_chains(ft, kind, params, n, S, M, d):
    full = {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]
    M = full if M is None else (M if kind != "tri" else min(M, n))
    P, Q = _plan(ft, kind, n, params)
    A = np.asfortranarray(np.random.default_rng(n + 7 * S + d).standard_normal((n, M)))
    bands, tab, chains = se.host_build(ft._lib, P.ptr, d, S, M)
    got = se.emulate(bands, tab, chains, S, A.copy(order="F"))
    ref = O.execute(Q, d, A.copy(order="F"), M)
    assert O.rel2(got, ref) <= 1e-14
    # every chain fits a workgroup and every band's stream fits the staging rows
    assert all(1 <= c.nb <= 16 for c in chains) or not chains
    # the last sequence emits position R at step R + (S+1)*15 + S
    assert all(b.R < 1024 and b.steps % 8 == 0 and b.steps >= b.R + (S + 1) * 15 + S + 1 for b in bands)
