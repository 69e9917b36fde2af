"""Step-by-step CPU emulation of the band-systolic engine (ft_systolic.hpp)
from the bands, tables and chains the library's host code builds
(``ftamd_sys_debug_host``).  Test infrastructure: it checks the host-side
construction (class geometry, band streams, step-major tables, chains,
column masks) against the oracle without a GPU; the kernel itself is checked
on the GPU by tests/test_gpu.py.
"""
import ctypes

import numpy as np

LANES = 16


class Band(ctypes.Structure):
    _fields_ = [("tab", ctypes.c_longlong), ("R", ctypes.c_int), ("steps", ctypes.c_int),
                ("gbase", ctypes.c_int), ("nseq", ctypes.c_int)]


class Chain(ctypes.Structure):
    _fields_ = [("band0", ctypes.c_int), ("nb", ctypes.c_int), ("dir", ctypes.c_int), ("rstep", ctypes.c_int),
                ("rpar", ctypes.c_int), ("prio", ctypes.c_int), ("lda", ctypes.c_int), ("split_in", ctypes.c_int),
                ("split_out", ctypes.c_int), ("col", ctypes.c_int * 8),
                ("glo", ctypes.c_int * 8), ("ghi", ctypes.c_int * 8), ("aoff", ctypes.c_longlong)]


# column layouts of a sphere-geometry matrix (FTAMD_LAYOUT_* in ft_hip_abi.h)
LAYOUT = {"std": 0, "sphv": 1, "rectdisk": 2}


def host_build(lib, plan_ptr, direction, S, M, maxpart=16, layout="std"):
    """(bands, table (entries x 2: s, c), chain parts of at most maxpart bands)
    as the library builds them for the column layout (std: the plan's own;
    sphv or rectdisk on a sphere-geometry plan)."""
    fn = lib.ftamd_sys_debug_host
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.POINTER(Band)), ctypes.POINTER(ctypes.c_int),
                   ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.c_longlong),
                   ctypes.POINTER(ctypes.POINTER(Chain)), ctypes.POINTER(ctypes.c_int)]
    pb, nb = ctypes.POINTER(Band)(), ctypes.c_int()
    pt, ne = ctypes.POINTER(ctypes.c_double)(), ctypes.c_longlong()
    pc, nc = ctypes.POINTER(Chain)(), ctypes.c_int()
    rc = fn(ctypes.c_void_p(plan_ptr), direction, S, M, LAYOUT[layout], maxpart, ctypes.byref(pb), ctypes.byref(nb), ctypes.byref(pt),
            ctypes.byref(ne), ctypes.byref(pc), ctypes.byref(nc))
    if rc != 0:
        raise RuntimeError("ftamd_sys_debug_host failed: %d" % rc)
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    try:
        bands = [Band.from_buffer_copy(pb[i]) for i in range(nb.value)]
        tab = np.ctypeslib.as_array(pt, shape=(max(ne.value, 1) * 2,)).copy().reshape(-1, 2)
        chains = [Chain.from_buffer_copy(pc[i]) for i in range(nc.value)]
    finally:
        for p in (pb, pt, pc):
            libc.free(ctypes.cast(p, ctypes.c_void_p))
    return bands, tab, chains


def emulate(bands, tab, chains, S, A):
    """Runs every chain part on the Fortran-ordered array A (in place), in
    list order (the parts of a chain are listed in chain order): element e of
    column c of a chain is A's entry aoff + c*lda + e in Fortran order, as the
    kernel addresses it.  A part after the first of its chain (split_in >= 0)
    reads its whole stream from A, where the part before it (split_out >= 0)
    stored every row it emitted: what the kernel hands over through A and a
    progress counter."""
    B = LANES * S
    F = A.reshape(-1, order="F")
    assert np.shares_memory(F, A)
    for ch in chains:
        cols = np.array(list(ch.col)).reshape(4, 2)
        valid = cols >= 0
        rstep, rpar, d = ch.rstep, ch.rpar, ch.dir
        prev = None
        for w in range(ch.nb):
            bd = bands[ch.band0 + w]
            R, steps = bd.R, bd.steps
            first, last = w == 0, w == ch.nb - 1
            Rprev = bands[ch.band0 + w - 1].R if not first else 0
            Rnext = bands[ch.band0 + w + 1].R if not last else -1
            npa = R + 1 if first else (R - Rprev if d == 0 else 0)
            npa_next = 0 if last else (Rnext - R if d == 0 else 0)
            row_of = (lambda p: R - p) if d == 0 else (lambda p: p)
            inp = np.zeros((steps + 2, 4, 2))
            for p in range(min(steps + 2, R + 1)):
                if p < npa:
                    e = rstep * row_of(p) + rpar
                    inp[p][valid] = F[ch.aoff + cols[valid] * ch.lda + e]
                else:
                    inp[p] = prev[p]
            q = np.arange(B)
            g = bd.gbase + q
            glo = np.array(list(ch.glo)).reshape(4, 2)
            ghi = np.array(list(ch.ghi)).reshape(4, 2)
            act = (g[:, None, None] >= glo[None]) & (g[:, None, None] <= ghi[None])
            Y = np.zeros((B, 4, 2))
            E = np.zeros((B, 4, 2))
            nxt = {}
            for t in range(steps):
                # FP64 tables are slot-major within a step: sequence S*i + s at s*16 + i
                sc = tab[bd.tab + t * B + (q % S) * LANES + q // S]
                sn = np.where(act, sc[:, 0, None, None], 0.0)
                cs = np.where(act, sc[:, 1, None, None], 1.0)
                # lane i's first sequence takes what the lane before emitted at
                # the previous step (lane 0: the stream); sequence s > 0 of a
                # lane takes what sequence s-1 emits in this step
                X0 = np.concatenate([inp[t][None], E[S - 1::S][:-1]], axis=0)
                Yn, En = np.empty_like(Y), np.empty_like(E)
                for s in range(S):
                    k = np.arange(LANES) * S + s
                    x = X0 if s == 0 else En[k - 1]
                    Yn[k] = cs[k] * x + sn[k] * Y[k]
                    En[k] = cs[k] * Y[k] - sn[k] * x
                Y, E = Yn, En
                po = t - ((S + 1) * (LANES - 1) + S)
                if 0 <= po <= R:
                    P = po + npa_next
                    if not last and P <= Rnext:
                        nxt[P] = E[B - 1].copy()
                    else:
                        e = rstep * row_of(po) + rpar
                        F[ch.aoff + cols[valid] * ch.lda + e] = E[B - 1][valid]
            prev = nxt
    return A


def host_build_tet(lib, plan_ptr, direction, S, part, N, L, m0, count, maxpart=16):
    """(bands, table, chain parts) of one of the tetrahedron's stages as the
    library builds them (ftamd_sys_debug_tet): part 0 the slab columns on the
    triangle plan of size N, part 1 the slab rows on the plan of size L."""
    fn = lib.ftamd_sys_debug_tet
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 8 + [
        ctypes.POINTER(ctypes.POINTER(Band)), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.c_longlong),
        ctypes.POINTER(ctypes.POINTER(Chain)), ctypes.POINTER(ctypes.c_int)]
    pb, nb = ctypes.POINTER(Band)(), ctypes.c_int()
    pt, ne = ctypes.POINTER(ctypes.c_double)(), ctypes.c_longlong()
    pc, nc = ctypes.POINTER(Chain)(), ctypes.c_int()
    rc = fn(ctypes.c_void_p(plan_ptr), direction, S, part, N, L, m0, count, maxpart, ctypes.byref(pb),
            ctypes.byref(nb), ctypes.byref(pt), ctypes.byref(ne), ctypes.byref(pc), ctypes.byref(nc))
    if rc != 0:
        raise RuntimeError("ftamd_sys_debug_tet failed: %d" % rc)
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    try:
        bands = [Band.from_buffer_copy(pb[i]) for i in range(nb.value)]
        tab = np.ctypeslib.as_array(pt, shape=(max(ne.value, 1) * 2,)).copy().reshape(-1, 2)
        chains = [Chain.from_buffer_copy(pc[i]) for i in range(nc.value)]
    finally:
        for p in (pb, pt, pc):
            libc.free(ctypes.cast(p, ctypes.c_void_p))
    return bands, tab, chains


def host_build_spin(lib, plan_ptr, direction, S, s, M, maxpart=16):
    """(bands, table, chain parts) of the spin-weighted sphere's second stage
    on a sphere plan as the library builds them (ftamd_sys_debug_spin)."""
    fn = lib.ftamd_sys_debug_spin
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [
        ctypes.POINTER(ctypes.POINTER(Band)), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.c_longlong),
        ctypes.POINTER(ctypes.POINTER(Chain)), ctypes.POINTER(ctypes.c_int)]
    pb, nb = ctypes.POINTER(Band)(), ctypes.c_int()
    pt, ne = ctypes.POINTER(ctypes.c_double)(), ctypes.c_longlong()
    pc, nc = ctypes.POINTER(Chain)(), ctypes.c_int()
    rc = fn(ctypes.c_void_p(plan_ptr), direction, S, s, M, maxpart, ctypes.byref(pb), ctypes.byref(nb),
            ctypes.byref(pt), ctypes.byref(ne), ctypes.byref(pc), ctypes.byref(nc))
    if rc != 0:
        raise RuntimeError("ftamd_sys_debug_spin failed: %d" % rc)
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    try:
        bands = [Band.from_buffer_copy(pb[i]) for i in range(nb.value)]
        tab = np.ctypeslib.as_array(pt, shape=(max(ne.value, 1) * 2,)).copy().reshape(-1, 2)
        chains = [Chain.from_buffer_copy(pc[i]) for i in range(nc.value)]
    finally:
        for p in (pb, pt, pc):
            libc.free(ctypes.cast(p, ctypes.c_void_p))
    return bands, tab, chains


class Item(ctypes.Structure):
    _fields_ = [("chain", ctypes.c_int), ("wave0", ctypes.c_int)]


def assign(lib, bands, chains, nxcd, nsx):
    """(items, slot_items, xcd_slot) as the library deals the chain parts to
    nxcd XCDs of nsx slots each (ftamd_sys_assign)."""
    fn = lib.ftamd_sys_assign
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    nch, nslots = len(chains), nxcd * nsx
    ch = (Chain * max(nch, 1))(*chains)
    bd = (Band * max(len(bands), 1))(*bands)
    items = (Item * max(nch, 1))()
    slot_items = (ctypes.c_int * (nslots + 1))()
    xcd_slot = (ctypes.c_int * (nxcd + 1))()
    ws_load = (ctypes.c_long * (nslots * 17 + nxcd))()
    ws_chain = (ctypes.c_int * max(2 * nch, 1))()
    rc = fn(ctypes.addressof(ch), nch, ctypes.addressof(bd), nxcd, nsx, ctypes.addressof(items),
            ctypes.addressof(slot_items), ctypes.addressof(xcd_slot), ctypes.addressof(ws_load),
            ctypes.addressof(ws_chain))
    if rc != 0:
        raise RuntimeError("ftamd_sys_assign failed: %d" % rc)
    return [(it.chain, it.wave0) for it in items[:nch]], list(slot_items), list(xcd_slot)
