"""The autotuning table (ft_tuning.c, fasttransforms_amd/tune.py): the table
compiled into the library covers every (kind, precision, direction, batch)
with contiguous, non-overlapping degree ranges and valid choices; knobs
override it; tables installed at run time are validated.  On the GPU: every
tuned choice gives bitwise the same result as the built-in engines (and the
oracle where it is cheap)."""
import numpy as np
import pytest

import oracle as O

NMAX = 2 ** 31 - 1


def _groups(rows):
    g = {}
    for r in rows:
        g.setdefault(tuple(r[:4]), []).append(r)
    return g


def test_compiled_table_covers_every_shape(ft):
    ft.set_tuning_table("compiled")
    rows = ft.tuning_table()
    assert rows, "the library carries a tuned table"
    g = _groups(rows)
    assert set(g) == {(k, p, d, b) for k in (0, 1, 2) for p in (0, 1) for d in (0, 1) for b in (0, 1)}
    for key, rs in g.items():
        rs = sorted(rs, key=lambda r: r[4])
        assert rs[0][4] == 1 and rs[-1][5] == NMAX, key
        for a, b in zip(rs, rs[1:]):
            assert b[4] == a[5] + 1, (key, a, b)
        for r in rs:
            assert r[6] in (ft.ENGINE_PIPELINE, ft.ENGINE_SYSTOLIC)
            if r[6] == ft.ENGINE_SYSTOLIC:
                assert 1 <= r[7] <= 16
            else:
                assert r[8] in (2, 4, 8) and r[9] in (4, 8, 16) and 0 <= r[10] <= 8


def test_choice_follows_table_then_knobs(ft):
    ft.set_tuning_table("compiled")
    for r in ft.tuning_table():
        n = r[4] if r[5] == NMAX else (r[4] + r[5]) // 2
        e, mp, L, W, wg = ft.tuned_choice(r[0], r[1], r[2], r[3], n)
        assert e == r[6]
        if e == ft.ENGINE_SYSTOLIC:
            assert mp == r[7]
        else:
            assert (L, W) == (r[8], r[9])
    try:
        ft.set_engine(ft.ENGINE_PIPELINE)
        ft.set_tuning(2, 8, 3)
        ft.set_systolic(5)
        assert ft.tuned_choice("sph", 0, 0, 0, 1024) == (ft.ENGINE_PIPELINE, 5, 2, 8, 3)
    finally:
        ft.set_engine(ft.ENGINE_AUTO)
        ft.set_tuning(0, 0, 0)
        ft.set_systolic(0)


def test_installed_tables(ft):
    try:
        row = [0, 0, 0, 0, 1, NMAX, ft.ENGINE_PIPELINE, 0, 8, 8, 0]
        ft.set_tuning_table([row])
        assert ft.tuning_table() == [row]
        assert ft.tuned_choice("sph", 0, 0, 0, 1024)[:4] == (ft.ENGINE_PIPELINE, 16, 8, 8)
        # shapes the table does not cover take the built-in choice
        assert ft.tuned_choice("tri", 0, 0, 0, 1024) == (ft.ENGINE_SYSTOLIC, 16, 4, 16, 0)
        assert ft.tuned_choice("tri", 0, 0, 1, 1024) == (ft.ENGINE_PIPELINE, 16, 4, 16, 2)
        for bad in ([0, 0, 0, 0, 5, 4, 2, 8, 0, 0, 0], [3, 0, 0, 0, 1, 9, 2, 8, 0, 0, 0],
                    [0, 0, 0, 0, 1, 9, 7, 8, 0, 0, 0], [0, 0, 0, 0, 1, 9, 1, 0, 3, 8, 0]):
            with pytest.raises(ValueError):
                ft.set_tuning_table([bad])
        assert ft.tuning_table() == [row]
        ft.set_tuning_table(None)
        assert ft.tuning_table() == []
        assert ft.tuned_choice("sph", 0, 0, 0, 2048) == (ft.ENGINE_SYSTOLIC, 16, 4, 16, 0)
    finally:
        ft.set_tuning_table("compiled")


def _rep(r):
    """A degree of the row's range, at most 700 (oracle-cheap)."""
    return max(r[4], min(r[5], 700, int(np.sqrt(r[4] * min(r[5], 4096)))))


@pytest.mark.gpu
def test_tuned_choices_bitwise(ft):
    """Each single-transform row of the compiled table, at a degree of its
    range: the tuned launch (engine recorded) equals the pipelined engine
    with its built-in shape bit for bit, and the oracle in FP64."""
    import torch
    ft.set_tuning_table("compiled")
    makers = {0: ("sph", ()), 1: ("tri", (0.0, 0.0, 0.0)), 2: ("disk", (0.0, 0.0))}
    seen = set()
    for r in ft.tuning_table():
        if r[3] != 0:
            continue
        kind, params = makers[r[0]]
        n = _rep(r)
        if (r[0], r[1], r[2], n) in seen:
            continue
        seen.add((r[0], r[1], r[2], n))
        M = {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind]
        P = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind](n, *params)
        A = O.fill(kind, n, M, 900 + n)
        dt = torch.float64 if r[1] == 0 else torch.float32
        X = torch.from_numpy(np.ascontiguousarray(A.T)).to(dt).cuda()
        T1 = X.clone()
        ft.gpu_execute(P, r[2], T1, precision=r[1])
        assert ft.last_engine() == r[6], r
        try:
            ft.set_engine(ft.ENGINE_PIPELINE)
            T2 = X.clone()
            ft.gpu_execute(P, r[2], T2, precision=r[1])
        finally:
            ft.set_engine(ft.ENGINE_AUTO)
        ft.synchronize()
        assert torch.equal(T1, T2), (r, n)
        if r[1] == 0 and n <= 300:
            R = O.execute(O.Plan(kind, n, *params), r[2], A.copy(order="F"))
            assert np.array_equal(T1.cpu().numpy().T, R), (r, n)


@pytest.mark.gpu
def test_tuned_batches_bitwise(ft):
    """Each batch row of the compiled table on a batch of 12 transforms of
    degree in its range: tuned launch equals the built-in pipelined batch."""
    import torch
    ft.set_tuning_table("compiled")
    makers = {0: ("sph", ()), 1: ("tri", (0.0, 0.0, 0.0)), 2: ("disk", (0.0, 0.0))}
    rng = np.random.default_rng(77)
    for r in ft.tuning_table():
        if r[3] != 1:
            continue
        kind, params = makers[r[0]]
        big = _rep(r)
        ns = [int(v) for v in np.unique(rng.integers(16, big + 1, 11))] + [big]
        plans = [{"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind](n, *params)
                 for n in ns]
        dt = torch.float64 if r[1] == 0 else torch.float32
        Xs = [torch.from_numpy(np.ascontiguousarray(
            O.fill(kind, n, {"sph": 2 * n - 1, "tri": n, "disk": 4 * n - 3}[kind], 50 + n).T)).to(dt).cuda()
              for n in ns]
        A1 = [x.clone() for x in Xs]
        ft.gpu_execute_batch(plans, r[2], A1, precision=r[1])
        try:
            ft.set_engine(ft.ENGINE_PIPELINE)
            A2 = [x.clone() for x in Xs]
            ft.gpu_execute_batch(plans, r[2], A2, precision=r[1])
        finally:
            ft.set_engine(ft.ENGINE_AUTO)
        ft.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(A1, A2)), (r, ns)
