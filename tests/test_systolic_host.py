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


def _check_parts(chains, bands, maxpart):
    """Parts of at most maxpart bands, each chain's parts consecutive and in
    chain order, linked by distinct progress counters 0 .. links-1."""
    assert all(1 <= c.nb <= maxpart for c in chains)
    links = set()
    for k, c in enumerate(chains):
        if c.split_in >= 0:
            p = chains[k - 1]
            assert k > 0 and p.split_out == c.split_in
            assert p.band0 + p.nb == c.band0 and list(p.col) == list(c.col) and p.prio == c.prio
        if c.split_out >= 0:
            assert c.split_out not in links
            links.add(c.split_out)
            assert k + 1 < len(chains) and chains[k + 1].split_in == c.split_out
    assert links == set(range(len(links)))


@pytest.mark.parametrize("kind,params", CASES)
@pytest.mark.parametrize("n,S,M,maxpart", [(3, 2, None, 16), (20, 2, None, 16), (100, 2, None, 16),
                                           (100, 4, None, 16), (77, 2, 31, 16), (150, 2, None, 16),
                                           (150, 4, 91, 16), (150, 2, None, 1), (150, 2, 91, 2), (100, 2, None, 3),
                                           (300, 2, None, 4)])
@pytest.mark.parametrize("d", [0, 1])
def test_host_tables_and_chains(ft, kind, params, n, S, M, maxpart, d):
    full = {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]
    M = full if M is None else (M if kind != "tri" else min(M, n))
    P, Q = _plan(ft, kind, n, params)
    A = np.asfortranarray(np.random.default_rng(n + 7 * S + d).standard_normal((n, M)))
    bands, tab, chains = se.host_build(ft._lib, P.ptr, d, S, M, maxpart)
    got = se.emulate(bands, tab, chains, S, A.copy(order="F"))
    ref = O.execute(Q, d, A.copy(order="F"), M)
    assert O.rel2(got, ref) <= 1e-14
    _check_parts(chains, bands, maxpart)
    # the last sequence emits position R at step R + (S+1)*15 + S
    assert all(b.steps % 8 == 0 and b.steps >= b.R + (S + 1) * 15 + S + 1 for b in bands)


@pytest.mark.parametrize("kind,n,maxpart,nxcd,nsx", [("sph", 300, 2, 8, 4), ("tri", 300, 4, 8, 32), ("sph", 1024, 8, 8, 32),
                                                     ("tri", 1024, 8, 8, 32), ("disk", 600, 3, 4, 8), ("sph", 100, 1, 1, 8)])
@pytest.mark.parametrize("d", [0, 1])
def test_slot_assignment(ft, kind, n, maxpart, nxcd, nsx, d):
    """Every part runs once; the waves it takes exist; the parts of a chain run
    on one XCD in strictly increasing slots (a part waits only on earlier
    slots of its XCD, which the kernel's in-order tickets make deadlock-free);
    no slot holds more than its share of work by much."""
    params = {"sph": (), "tri": (0.0, 0.0, 0.0), "disk": (0.0, 0.0)}[kind]
    P, _ = _plan(ft, kind, n, params)
    M = {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]
    bands, _, chains = se.host_build(ft._lib, P.ptr, d, 2, M, maxpart)
    items, slot_items, xcd_slot = se.assign(ft._lib, bands, chains, nxcd, nsx)
    assert xcd_slot == [x * nsx for x in range(nxcd + 1)] and slot_items[-1] == len(chains)
    slot_of = {}
    for s in range(nxcd * nsx):
        for k in range(slot_items[s], slot_items[s + 1]):
            c, w0 = items[k]
            assert c not in slot_of and 0 <= w0 and w0 + chains[c].nb <= 16
            slot_of[c] = s
    assert sorted(slot_of) == list(range(len(chains)))
    for k, c in enumerate(chains):
        if c.split_in >= 0:
            a, b = slot_of[k - 1], slot_of[k]
            assert a // nsx == b // nsx and a < b
    work = np.zeros(nxcd * nsx)
    for k, c in enumerate(chains):
        work[slot_of[k]] += sum(bands[c.band0 + w].steps for w in range(c.nb))
    biggest = max(sum(bands[c.band0 + w].steps for w in range(c.nb)) for c in chains)
    assert work.max() <= work.mean() + 2 * biggest
