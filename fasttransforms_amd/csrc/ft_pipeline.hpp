// ft_pipeline.hpp -- the pipelined rotation engine (templates), shared by the
// per-geometry kernel files.  See ft_rotations.hip for the design.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "ft_hip_abi.h"

namespace ftamd {

constexpr int kLanes = 64;

// Optional timeline instrumentation (-DFTAMD_TRACE, diagnostic builds only):
// per stage wave and task, start/end wall clock, tiles, and the shader-clock
// cycles spent in each part of a tile.  The product build compiles all of it
// away.  Slots: 0 LDS-ring input wait, 1 progress wait, 2 row stores and
// progress publication, 3 input rows arriving, 4 rotations, 5 LDS-ring
// output, 6 tiles.  -DFTAMD_STAMPS keeps only the per-tile clock stamps
// (input taken, rotated, emitted), which perturb the timing least.
#if defined(FTAMD_STAMPS) && !defined(FTAMD_TRACE)
#define FTAMD_TRACE
#define FTAMD_TRACE_LITE
#endif
#if defined(FTAMD_TRACE) && !defined(FTAMD_TRACE_LITE)
#define FT_TR_BEGIN() const unsigned long long ft_tr0 = __builtin_amdgcn_s_memtime()
#define FT_TR_END(slot) c.tr[slot] += __builtin_amdgcn_s_memtime() - ft_tr0
#define FT_TR_COUNT(slot) c.tr[slot] += 1
#define FT_TR_SYNC() __builtin_amdgcn_s_waitcnt(0)
#else
#define FT_TR_BEGIN() (void)0
#define FT_TR_END(slot) (void)0
#define FT_TR_COUNT(slot) (void)0
#define FT_TR_SYNC() (void)0
#endif

// ---------------------------------------------------------------- geometry --
// The engine works on "rows" e = 0 .. ne-1 of a sub-problem in which every
// rotation couples rows (e, e+1).  For the triangle and the disk that is the
// column itself (rotations_default.c:45-71 couple k and k+1).  The sphere
// couples k and k+2 (rotations_default.c:31-43), so its even rows and its odd
// rows are two independent chains: each sphere column is run as two
// sub-problems, rpar = 0 (rows 0,2,4,...) and rpar = 1 (rows 1,3,5,...).
//   ne(n, rpar)        rows of the sub-problem;
//   top(n, rpar, j)    largest e whose pair (e, e+1) sequence j rotates;
//   tab(N, rpar, j, e) table entry of that rotation, for a table built with
//                      n = N (packed index of rotations.c:113-114, 170-171;
//                      the sphere's device table stores each column's even
//                      rows first, then its odd rows, so both are contiguous);
//   real(rpar, e)      row of the column.
template <int KIND> struct Geo;
template <> struct Geo<FTAMD_SPH> {
    static constexpr int STEP = 2;
    __device__ static int ne(int n, int rpar) { return (n - rpar + 1) / 2; }
    __device__ static int top(int n, int rpar, int j) {
        const int k = n - 3 - j - rpar;  // largest row k = rpar (mod 2) of sequence j, minus rpar
        return k < 0 ? -1 : k / 2;
    }
    __device__ static long long tab(int N, int rpar, int j, int e) {
        return (long long)j * (2 * N + 1 - j) / 2 + (rpar ? (N - j + 1) / 2 : 0) + e;
    }
    __device__ static int real(int rpar, int e) { return 2 * e + rpar; }
};
template <> struct Geo<FTAMD_TRI> {
    static constexpr int STEP = 1;
    __device__ static int ne(int n, int) { return n; }
    __device__ static int top(int n, int, int j) { return n - 2 - j; }
    __device__ static long long tab(int N, int, int j, int e) { return (long long)j * (2 * N + 1 - j) / 2 + e; }
    __device__ static int real(int, int e) { return e; }
};
template <> struct Geo<FTAMD_DISK> {
    static constexpr int STEP = 2;
    __device__ static int ne(int n, int) { return n; }
    __device__ static int top(int n, int, int j) { return n - 2 - (j + 1) / 2; }
    __device__ static long long tab(int N, int, int j, int e) {
        return (long long)j * N - (long long)(j / 2) * ((j + 1) / 2) + e;
    }
    __device__ static int real(int, int e) { return e; }
};

template <typename T> using gptr = __attribute__((address_space(1))) T *;

// One table entry: the interleaved (s, c) of one rotation, loaded as one
// 2-vector (16 bytes in FP64).
template <typename T> using SCPair = T __attribute__((ext_vector_type(2)));

template <typename T> __device__ __forceinline__ T fmaT(T a, T b, T c);
template <> __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// One rotation of the (lower, upper) virtual row pair.
template <typename T> __device__ __forceinline__ void givens(T &x, T &y, T s, T c) {
    T nx = fmaT<T>(c, x, s * y);
    T ny = fmaT<T>(c, y, -(s * x));
    x = nx;
    y = ny;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane N of every row of 16 lanes, broadcast to the whole row (DPP
// row_newbcast).  The sines and cosines of a tile are held one per lane
// (lane l: entry l mod 16) and broadcast where they are used, so they come
// from ordinary vector loads -- returned in order, prefetched a tile ahead --
// rather than from scalar loads, whose scalar-cache misses serialise.
template <int N, typename T> __device__ __forceinline__ T bcast(T v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + N, 0xf, 0xf, true);
}
// n must fold to a constant (every caller is fully unrolled).
template <typename T> __device__ __forceinline__ T bcast_n(T v, int n) {
    switch (n & 15) {
        case 0: return bcast<0>(v);
        case 1: return bcast<1>(v);
        case 2: return bcast<2>(v);
        case 3: return bcast<3>(v);
        case 4: return bcast<4>(v);
        case 5: return bcast<5>(v);
        case 6: return bcast<6>(v);
        case 7: return bcast<7>(v);
        case 8: return bcast<8>(v);
        case 9: return bcast<9>(v);
        case 10: return bcast<10>(v);
        case 11: return bcast<11>(v);
        case 12: return bcast<12>(v);
        case 13: return bcast<13>(v);
        case 14: return bcast<14>(v);
        default: return bcast<15>(v);
    }
}

// Bounded waits.  ~2^22 polls of >= 64 cycles each is far beyond any
// legitimate wait; on expiry the caller abandons its task.
constexpr int kSpinLimit = 1 << 22;

// Inter-task hand-off.  The last stage of a task stores the rows it emits in a
// row-major boundary buffer ([row][lane]: one row of 64 lanes is 8 whole
// 64-byte lines) and the first stage of the next task reads them back.  Both
// use device-scope relaxed atomics (sc1 stores and loads), coherent per
// location between CUs without any L2 write-back or invalidation; the
// producer waits until its stores are acknowledged (s_waitcnt vmcnt(0))
// before it publishes its progress with a device-scope atomic store, and the
// consumer polls that with device-scope atomic loads.  Chained tasks also run
// on one XCD, which keeps their traffic in one L2.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

__device__ __forceinline__ void publish(int *prog, int value) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    wait_vmem();
    __hip_atomic_store(prog, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until rows >= target are final.
__device__ __forceinline__ bool global_wait_le(const int *prog, int target, int *known, const int *abort_flag) {
    if (*known > target) {
        int it = 0;
        for (;; ++it) {
            int v = uniform(__hip_atomic_load(prog, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (v <= target) {
                *known = v;
                break;
            }
            if (it >= kSpinLimit ||
                uniform(__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
                return false;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return true;
}

// Issue priority of this wave against the other waves of its SIMD: the
// stages of the longest chains (the launch's critical path) go first.
__device__ __forceinline__ void set_priority(int prio) {
    switch (prio) {
        case 3: __builtin_amdgcn_s_setprio(3); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        default: __builtin_amdgcn_s_setprio(0); break;
    }
}

// LDS accesses of the stage-to-stage ring.  Relaxed atomics so that the
// compiler neither caches nor reorders them; LDS executes the instructions of
// one wave in issue order, which is the only ordering the ring relies on.
template <typename T> __device__ __forceinline__ T lds_ld(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T> __device__ __forceinline__ void lds_st(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ------------------------------------------------------------------ kernel --
template <typename T, int KIND, int DIR, int L, int W, int R>
struct Pipeline {
    using G = Geo<KIND>;
    static constexpr int STEP = G::STEP;
    static constexpr int WS = 2 * L;           // window rows per lane
    static constexpr int NP = L * L;           // rotations (table entries) per tile
    static constexpr int NS = (NP + 15) / 16;  // lane-distributed registers holding them

    // The table entries of one tile, lane-distributed: register g of lane l
    // holds entry p = (16 g + l mod 16) mod NP of the tile, p = lj*L + i
    // (sequence lj, the i-th of its L consecutive entries).
    struct SC {
        T s[NS], c[NS];
    };

    // A task's workgroup: WC compute stages (waves 0 .. WC-1), an importer
    // (wave WC) that feeds stage 0 the rows it reads from A or from the
    // previous task, and an exporter (wave WC+1) that takes the last stage's
    // rows to the next task or to A.  Compute stages talk only through LDS
    // rings; the global-memory latencies of the hand-off between tasks stay
    // with the two I/O waves.
    static constexpr int WC = W - 2;

    struct Shared {
        T ring[WC + 1][R][L][kLanes];  // ring r: written by stage r-1 (the importer for r = 0)
        int emitted[WC + 1];
        int consumed[WC + 1];
        int task;
        int abort_flag;
#ifdef FTAMD_TRACE
        unsigned long long tr[W][8];
#endif
    };

    // Everything one stage needs, uniform except the per-lane data.
    // Rows are "virtual": v = e for hi2lo, v = ne-1-e for lo2hi (lo2hi runs as
    // hi2lo on reversed rows: the transposed rotation of the reversed pair is
    // the forward rotation).
    struct Stage {
        gptr<const SCPair<T>> tab;  // table entry 0
        gptr<T> A;          // the transform's matrix
        int coff;           // this lane's column: element offset in A (-1: no column)
        gptr<T> bin;        // boundary buffer the task reads (row 0 of it), nullptr in the first pass
        gptr<T> bout;       // boundary buffer the task writes, nullptr in the last pass
        int rs;             // row stride of A
        int n, N;           // rows of the transform, n of the table
        int rpar;           // sphere row parity of the sub-problem (-1: none)
        int ne;             // rows of the sub-problem
        int jmax;           // largest j this lane needs
        int q0, nq;         // first sequence and count
        int j0;             // j of sequence q0
        int K, Bot;         // first tile position, lowest virtual row to emit
        int ntile;          // tiles
        int jfull;
        bool lanes_all;     // every lane active for every sequence of the stage
        int lo0;            // lo2hi: lowest virtual row of sequence q0
        int lo_next;        // lo2hi: emitted rows below it are final (no later task touches them)
        int lo_valid;       // lowest virtual row the task reads or writes
        int emit_off;       // tiles before the first emission
        int sci[NS];        // this lane's table indices for tile 0
    };

    // Per-stage control of one sweep.
    struct Ctl {
        int prev_emits;     // emissions the ring this stage reads will carry
        int consumed_seen;  // tiles the next stage was last seen to have taken
        int prio;           // issue priority of the stage's other tiles
#ifdef FTAMD_TRACE
        unsigned long long tr[8];
        unsigned long long *stamps;  // per tile: input ready, rotated, emitted (first 160 tiles)
#endif
    };

    __device__ static int seq_j(int j_first, int q) { return DIR == 0 ? j_first - STEP * q : j_first + STEP * q; }
    __device__ static int real_row(const Stage &st, int v) { return G::real(st.rpar, DIR == 0 ? v : st.ne - 1 - v); }
    __device__ static int top_v(const Stage &st, int j) { return G::top(st.n, st.rpar, j); }

    template <int PH> __device__ static constexpr int slot(int i) { return (PH * L + i) % WS; }

    // Tile t sits at virtual position k = K - t*L; sequence lj uses rows
    // k+lj-(L-1) .. k+lj of its table column (hi2lo) or the mirrored rows
    // (lo2hi), so every lane's index moves by L per tile.
    __device__ static void init_sc_index(Stage &st, int lane) {
#pragma unroll
        for (int g = 0; g < NS; ++g) {
            const int p = (16 * g + (lane & 15)) % NP;
            int lj = p / L;
            const int i = p % L;
            if (lj >= st.nq) lj = 0;  // no such sequence: any valid entry (never used)
            const int j = DIR == 0 ? st.j0 - STEP * lj : st.j0 + STEP * lj;
            st.sci[g] = (int)(DIR == 0 ? G::tab(st.N, st.rpar, j, st.K + lj) - (L - 1) + i
                                       : G::tab(st.N, st.rpar, j, st.ne - 2 - (st.K + lj)) + i);
        }
    }
    __device__ static void load_sc(const Stage &st, int t, SC &sc) {
        const int d = DIR == 0 ? -t * L : t * L;
#pragma unroll
        for (int g = 0; g < NS; ++g) {
            const SCPair<T> e = st.tab[st.sci[g] + d];
            sc.s[g] = e.x;
            sc.c[g] = e.y;
        }
    }

    // One L x L tile at position k (virtual), phase PH of the register ring:
    // sequence lj rotates the pairs (r, r+1), r = k - lk + lj, with entry
    // i = L-1-lk (hi2lo) or lk (lo2hi) of its L.  The checked form (tiles at
    // the ends of a sweep, short stages, stages where some columns need fewer
    // sequences) computes every rotation and keeps the rows of the ones that
    // do not apply through selects: it stays one basic block that the
    // compiler interleaves like the plain form, whereas a branch per rotation
    // would serialise the FP64 latency chain of the tile.
    template <int PH, bool CHECKED>
    __device__ static void rotate_tile(const Stage &st, T (&x)[WS], int k, const SC &sc) {
#pragma unroll
        for (int lj = 0; lj < L; ++lj) {
            const int j = DIR == 0 ? st.j0 - STEP * lj : st.j0 + STEP * lj;
            int lo = 0, hi = 0;
            bool act = true;
            if (CHECKED) {
                const int top = top_v(st, j);
                lo = DIR == 0 ? 0 : st.ne - 2 - top;
                hi = DIR == 0 ? top : st.ne - 2;
                act = lj < st.nq && (st.lanes_all || j <= st.jmax);
            }
#pragma unroll
            for (int lk = 0; lk < L; ++lk) {
                const int r = k - lk + lj;
                const int p = lj * L + (DIR == 0 ? L - 1 - lk : lk);
                // broadcast with every lane active
                const T s = bcast_n(sc.s[p / 16], p % 16);
                const T c = bcast_n(sc.c[p / 16], p % 16);
                const int px = slot<PH>(L + lk - lj);
                const int py = slot<PH>(L + lk - lj - 1);
                if (CHECKED) {
                    const bool apply = act && r >= lo && r <= hi;
                    T nx = x[px], ny = x[py];
                    givens<T>(nx, ny, s, c);
                    x[px] = apply ? nx : x[px];
                    x[py] = apply ? ny : x[py];
                } else {
                    givens<T>(x[px], x[py], s, c);
                }
            }
        }
    }

    // hi2lo tile that reaches below row 0 (k - (L-1) < 0) in a stage whose
    // every lane needs all L sequences: only rotations with r = KB - lk + lj
    // >= 0 apply, and with KB a compile-time constant the others vanish from
    // the code.  These tiles carry the bottom rows from stage to stage, the
    // critical path of the launch, so they get straight-line code with no
    // wasted rotation.
    template <int PH, int KB>
    __device__ static void rotate_bottom(T (&x)[WS], const SC &sc) {
#pragma unroll
        for (int lj = 0; lj < L; ++lj)
#pragma unroll
            for (int lk = 0; lk < L; ++lk) {
                if (KB - lk + lj < 0) continue;
                const int p = lj * L + (L - 1 - lk);
                const T s = bcast_n(sc.s[p / 16], p % 16);
                const T c = bcast_n(sc.c[p / 16], p % 16);
                givens<T>(x[slot<PH>(L + lk - lj)], x[slot<PH>(L + lk - lj - 1)], s, c);
            }
    }
    // Dispatch k in [-L, L-2] to its specialisation.
    template <int PH, int KB>
    __device__ static void bottom_at(int k, T (&x)[WS], const SC &sc) {
        if constexpr (KB <= L - 2) {
            if (k == KB) rotate_bottom<PH, KB>(x, sc);
            else bottom_at<PH, KB + 1>(k, x, sc);
        }
    }

    // Row v handed over by the previous task, or read from A in the first pass.
    __device__ static T load_in(const Stage &st, int v, int lane) {
        if (v < st.lo_valid || v > st.ne - 1) return T(0);
        if (st.bin != nullptr)
            return __hip_atomic_load(st.bin + (long long)v * kLanes + lane, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        if (st.coff < 0) return T(0);
        return st.A[st.coff + real_row(st, v) * st.rs];
    }
    // Emitted row v: every one goes to the next task's boundary buffer (which
    // has FTAMD_BND_PAD spare rows on both sides, so each emission is exactly
    // L stores), and the final ones to A (the last pass; in lo2hi also the
    // rows no later pass touches).  Returns whether it wrote A.
    __device__ static bool store_out(const Stage &st, int v, T val, int lane) {
        if (st.bout != nullptr)
            __hip_atomic_store(st.bout + (long long)v * kLanes + lane, val, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (v < st.lo_valid || v > st.ne - 1) return false;
        if ((st.bout == nullptr || v < st.lo_next) && st.coff >= 0) {
            st.A[st.coff + real_row(st, v) * st.rs] = val;
            return true;
        }
        return false;
    }

    // XCD (accelerator complex die) this workgroup runs on.  Each XCD has its
    // own L2; a group's chained tasks are all handed to one XCD so that the
    // rows they pass on meet in one L2.
    __device__ static int xcd_id(int nxcd) {
        const int id = (int)__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));  // HW_REG_XCC_ID[3:0]
        return nxcd > 1 ? id % nxcd : 0;
    }

    __device__ static void run(const FtamdLaunch &a, Shared &sh) {
        const int lane = threadIdx.x & (kLanes - 1);
        const int w = uniform(threadIdx.x / kLanes);
        const int x = uniform(xcd_id(a.nxcd));
        const int first = uniform(a.xcd_off[x]);
        const int count = uniform(a.xcd_off[x + 1]) - first;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) {
                sh.task = (int)atomicAdd(a.ticket + x, 1u);
                sh.abort_flag = 0;
            }
            if (threadIdx.x < WC + 1) {
                sh.emitted[threadIdx.x] = 0;
                sh.consumed[threadIdx.x] = 0;
            }
            __syncthreads();
            const int t = uniform(sh.task);
            if (t >= count) break;
            const int task = first + t;
            const int g = uniform(a.tasks[2 * task]);
            const int p = uniform(a.tasks[2 * task + 1]);
#ifdef FTAMD_TRACE
            if (lane == 0)
                for (int i = 0; i < 8; ++i) sh.tr[w][i] = 0;
            const unsigned long long ft_w0 = wall_clock64();
            const unsigned long long ft_c0 = __builtin_amdgcn_s_memtime();
#endif
            if (w < WC) {
                set_priority(uniform(a.groups[g].prio));
                run_stage(a, sh, g, p, w, lane);
            } else {
                __builtin_amdgcn_s_setprio(3);
                if (w == WC) run_importer(a, sh, g, p, lane);
                else run_exporter(a, sh, g, p, lane);
            }
            set_priority(0);
#ifdef FTAMD_TRACE
            if (a.trace && lane == 0) {
                unsigned long long *rec = a.trace + ((long long)task * 16 + w) * 512;
                rec[0] = ft_w0;
                rec[1] = wall_clock64();
                rec[2] = __builtin_amdgcn_s_memtime() - ft_c0;
                for (int i = 0; i < 8; ++i) rec[3 + i] = sh.tr[w][i];
                rec[11] = ((unsigned long long)x << 32) | (unsigned)(blockIdx.x);
                rec[12] = ((unsigned long long)g << 32) | (unsigned)p;
            }
#endif
            __syncthreads();
            if (threadIdx.x == 0) atomicAdd(a.completed, 1u);
        }
        // Last workgroup out checks that every task ran and resets the tickets
        // and progress slots so the next launch (or graph replay) starts clean.
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned d = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (d == gridDim.x - 1) {
                if (__hip_atomic_load(a.completed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != (unsigned)a.ntasks)
                    atomicOr(a.err, FTAMD_ERR_INCOMPLETE);
                for (int i = 0; i < a.nprog; ++i)
                    __hip_atomic_store(a.prog + i, FTAMD_PROG_NONE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int i = 0; i < a.nxcd; ++i)
                    __hip_atomic_store(a.ticket + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(a.completed, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(a.done, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }

#ifdef FTAMD_TRACE
    __device__ static int first_task_of(const FtamdLaunch &a, int g, int p) {
        for (int i = 0; i < a.ntasks; ++i)
            if (a.tasks[2 * i] == g && a.tasks[2 * i + 1] == p) return i;
        return -1;
    }
#endif

    __device__ static void fail(const FtamdLaunch &a, Shared &sh) {
        __hip_atomic_store(&sh.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        atomicOr(a.err, FTAMD_ERR_TIMEOUT);
    }
    __device__ static bool aborted(Shared &sh) { return uniform(lds_ld(&sh.abort_flag)) != 0; }

    // Geometry of compute stage c of pass p of group g (false if the group
    // has no such stage).
    __device__ static bool setup(const FtamdLaunch &a, int g, int p, int c, int lane, Stage &st) {
        const FtamdGroup &gd = a.groups[g];
        const int nstages = uniform(gd.nstages);
        const int s = p * WC + c;
        if (s >= nstages) return false;
        const int npass = uniform(gd.npass);
        st.tab = (gptr<const SCPair<T>>)(a.sc);
        st.n = uniform(gd.n);
        st.N = a.ntab;
        st.rpar = uniform(gd.rpar);
        st.ne = G::ne(st.n, st.rpar);
        st.rs = uniform(gd.rs);
        st.A = (gptr<T>)(a.A ? a.A : gd.A);
        st.coff = a.lane_off[g * kLanes + lane];
        {
            gptr<T> bnd = (gptr<T>)(a.bnd) + gd.bnd;
            const long long per = (long long)(st.ne + 2 * FTAMD_BND_PAD) * kLanes;
            st.bin = p > 0 ? bnd + (long long)(p - 1) * per + FTAMD_BND_PAD * kLanes : nullptr;
            st.bout = p + 1 < npass ? bnd + (long long)p * per + FTAMD_BND_PAD * kLanes : nullptr;
        }
        st.jmax = a.lane_jmax[g * kLanes + lane];
        const int Q = uniform(gd.Q);
        const int j_first = uniform(gd.j_first);
        st.q0 = s * L;
        st.nq = min(L, Q - st.q0);
        st.j0 = seq_j(j_first, st.q0);
        st.jfull = uniform(gd.jfull);
        {
            const int jmax_stage = DIR == 0 ? st.j0 : seq_j(j_first, st.q0 + st.nq - 1);
            st.lanes_all = jmax_stage <= st.jfull;
        }
        if (DIR == 0) {
            st.K = top_v(st, st.j0);
            st.Bot = 0;
            st.lo0 = 0;
            st.lo_next = 0;
        } else {
            st.K = st.ne - 1;
            st.Bot = st.ne - 2 - top_v(st, seq_j(j_first, p * WC * L));
            st.lo0 = st.ne - 2 - top_v(st, st.j0);
            st.lo_next = p + 1 < npass ? max(st.ne - 2 - top_v(st, seq_j(j_first, (p + 1) * WC * L)), 0) : 0;
        }
        {
            const int E = st.K + 1 - st.Bot;
            st.ntile = E <= 0 ? 1 : (E + L - 1) / L + 1;
        }
        st.lo_valid = DIR == 0 ? 0 : max(st.Bot, 0);
        st.emit_off = DIR == 0 ? 0 : 1;
        return true;
    }

    // Last compute stage of pass p.
    __device__ static int last_stage(const FtamdLaunch &a, int g, int p) {
        return uniform(min(WC, a.groups[g].nstages - p * WC) - 1);
    }

    // ------------------------------------------------------------ importer --
    // Feeds stage 0 the rows of each of its tiles: from A in the first pass,
    // else from the previous task's boundary buffer once its progress shows
    // them final.  The next tile's rows are in flight while this tile's wait
    // for a free ring slot.
    __device__ static void run_importer(const FtamdLaunch &a, Shared &sh, int g, int p, int lane) {
        Stage st;
        if (!setup(a, g, p, 0, lane, st)) return;
        const int *prog = p > 0 ? a.prog + uniform(a.groups[g].prog_base) + p - 1 : nullptr;
        int known = FTAMD_PROG_NONE, consumed_seen = 0;
        T cur[L], nxt[L];
#ifdef FTAMD_TRACE
        unsigned long long *stamps = nullptr;
        if (a.trace) {
            const int task = first_task_of(a, g, p);
            if (task >= 0) stamps = a.trace + ((long long)task * 16 + WC) * 512 + 16;
        }
#endif
        if (prog && !global_wait_le(prog, max(st.K - L + 1, st.lo_valid), &known, &sh.abort_flag)) {
            fail(a, sh);
            return;
        }
#pragma unroll
        for (int i = 0; i < L; ++i) cur[i] = load_in(st, st.K - i, lane);
        for (int t = 0; t < st.ntile; ++t) {
            const int k = st.K - t * L;
            const bool more = t + 1 < st.ntile;
            const int next_target = max(k - 2 * L + 1, st.lo_valid);
            bool later = false;
            if (more) {
                if (!prog || known <= next_target) {
#pragma unroll
                    for (int i = 0; i < L; ++i) nxt[i] = load_in(st, k - L - i, lane);
                } else {
                    later = true;
                }
            }
            bool ok = t - R + 1 <= consumed_seen;
            for (int it = 0; !ok && it < kSpinLimit; ++it) {
                consumed_seen = uniform(lds_ld(&sh.consumed[0]));
                if (consumed_seen >= t - R + 1) {
                    ok = true;
                    break;
                }
                if (aborted(sh)) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok) {
                fail(a, sh);
                return;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            T *dst = &sh.ring[0][t % R][0][lane];
#pragma unroll
            for (int i = 0; i < L; ++i) lds_st(dst + i * kLanes, cur[i]);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            lds_st(&sh.emitted[0], t + 1);
#ifdef FTAMD_TRACE
            if (stamps && t < 160 && lane == 0) stamps[3 * t] = wall_clock64();
            if (stamps && t < 160 && lane == 0) stamps[3 * t + 2] = later ? 1 : 0;
#endif
            if (later) {
                if (!global_wait_le(prog, next_target, &known, &sh.abort_flag)) {
                    fail(a, sh);
                    return;
                }
#ifdef FTAMD_TRACE
                if (stamps && t < 160 && lane == 0) stamps[3 * t + 1] = wall_clock64();
#endif
#pragma unroll
                for (int i = 0; i < L; ++i) nxt[i] = load_in(st, k - L - i, lane);
            }
#pragma unroll
            for (int i = 0; i < L; ++i) cur[i] = nxt[i];
        }
    }

    // ------------------------------------------------------------ exporter --
    // Stores the last stage's emissions (boundary buffer and/or A) and
    // publishes the progress of the next task's input: the previous
    // emission's, once its stores are acknowledged, leaving this emission's L
    // stores in flight (vmcnt(L): the exporter issues no other vector-memory
    // instructions); an emission that also writes A is waited for whole.
    __device__ static void run_exporter(const FtamdLaunch &a, Shared &sh, int g, int p, int lane) {
        const int last = last_stage(a, g, p);
        Stage st;
        if (!setup(a, g, p, last, lane, st)) return;
        const int npass = uniform(a.groups[g].npass);
        int *prog = p + 1 < npass ? a.prog + uniform(a.groups[g].prog_base) + p : nullptr;
        const int nemit = st.ntile - st.emit_off;
        const int r = last + 1;
        bool prev_plain = false;
#ifdef FTAMD_TRACE
        unsigned long long *stamps = nullptr;
        if (a.trace) {
            const int task = first_task_of(a, g, p);
            if (task >= 0) stamps = a.trace + ((long long)task * 16 + WC + 1) * 512 + 16;
        }
#endif
        for (int e = 0; e < nemit; ++e) {
            const int k = st.K - (e + st.emit_off) * L;
            const T *src = &sh.ring[r][e % R][0][lane];
            T in[L];
            bool ok = false;
            for (int it = 0; it < kSpinLimit; ++it) {
                if (uniform(lds_ld(&sh.emitted[r])) >= e + 1) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
                    for (int i = 0; i < L; ++i) in[i] = lds_ld(src + i * kLanes);
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    ok = true;
                    break;
                }
                if (aborted(sh)) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok) {
                fail(a, sh);
                return;
            }
            lds_st(&sh.consumed[r], e + 1);
#ifdef FTAMD_TRACE
            if (stamps && e < 160 && lane == 0) stamps[3 * e] = wall_clock64();
#endif
            bool wrote_a = false;
#pragma unroll
            for (int i = 0; i < L; ++i) wrote_a |= store_out(st, k + L - i, in[i], lane);
            if (prog) {
                const bool plain = !(DIR == 1 && k + 1 < st.lo_next);
                if (plain && prev_plain) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_s_waitcnt(0x0F70 | L);  // vmcnt(L)
#ifdef FTAMD_TRACE
                    if (stamps && e < 160 && lane == 0) stamps[3 * e + 1] = wall_clock64();
#endif
                    __hip_atomic_store(prog, max(k + L + 1, st.lo_valid), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    publish(prog, max(k + 1, st.lo_valid));
                }
                prev_plain = plain;
            }
#ifdef FTAMD_TRACE
            if (stamps && e < 160 && lane == 0) stamps[3 * e + 2] = wall_clock64();
#endif
            (void)wrote_a;
        }
        if (prog) publish(prog, -1);
    }

    // --------------------------------------------------------- compute stage
    __device__ static void run_stage(const FtamdLaunch &a, Shared &sh, int g, int p, int w, int lane) {
        Stage st;
        if (!setup(a, g, p, w, lane, st)) return;
        init_sc_index(st, lane);
        Ctl c;
        if (w == 0) {
            c.prev_emits = st.ntile;  // the importer feeds every tile
        } else {
            const int jp = seq_j(uniform(a.groups[g].j_first), st.q0 - L);
            const int Kp = DIR == 0 ? top_v(st, jp) : st.ne - 1;
            const int E = Kp + 1 - st.Bot;
            const int Tp = E <= 0 ? 1 : (E + L - 1) / L + 1;
            c.prev_emits = Tp - st.emit_off;
        }
        c.consumed_seen = 0;
        c.prio = min(uniform(a.groups[g].prio), 2);
#ifdef FTAMD_TRACE
        for (int i = 0; i < 8; ++i) c.tr[i] = 0;
        c.stamps = nullptr;
        if (a.trace) {
            const int task = first_task_of(a, g, p);
            if (task >= 0) c.stamps = a.trace + ((long long)task * 16 + w) * 512 + 16;
        }
#endif

        T x[WS];
#pragma unroll
        for (int i = 0; i < WS; ++i) x[i] = T(0);
        SC sc[2];
        load_sc(st, 0, sc[0]);

        // Initial carried rows (positions 0 .. L-1 = virtual rows K+L .. K+1).
        // hi2lo: untouched by every earlier stage and task, read from A.
        // lo2hi: above the top row, zero.
        if (DIR == 0) {
#pragma unroll
            for (int i = 0; i < L; ++i) {
                const int v = st.K + L - i;
                x[i] = (v > st.ne - 1 || st.coff < 0) ? T(0) : st.A[st.coff + real_row(st, v) * st.rs];
            }
        }

        int t = 0;
        for (;;) {
            if (!tile<0>(a, sh, st, c, x, sc, t, w, lane)) return;
            if (++t >= st.ntile) break;
            if (!tile<1>(a, sh, st, c, x, sc, t, w, lane)) return;
            if (++t >= st.ntile) break;
        }
#ifdef FTAMD_TRACE
        for (int i = 0; i < 8; ++i) sh.tr[w][i] = c.tr[i];
#endif
    }

    template <int PH>
    __device__ static bool tile(const FtamdLaunch &a, Shared &sh, const Stage &st, Ctl &c, T (&x)[WS], SC (&sc)[2],
                                int t, int w, int lane) {
        const int k = st.K - t * L;
        const bool more = t + 1 < st.ntile;
        // The rows every later stage waits for pass through a stage in its
        // last tiles (hi2lo: the bottom rows) or its first ones (lo2hi: the
        // top virtual rows); those tiles get the highest issue priority.
        if (DIR == 0 ? t == st.ntile - 2 || (t == 0 && st.ntile < 2) : t == 0) __builtin_amdgcn_s_setprio(3);
        if (DIR == 0 ? t == 0 && st.ntile >= 2 : t == 2) set_priority(c.prio);
        // ---- the next tile's table entries, in flight during this tile
        if (more) load_sc(st, t + 1, sc[1 - PH]);
        // ---- new rows: positions L .. 2L-1 = virtual rows k .. k-L+1
        if (t < c.prev_emits) {
            // The flag and the rows are read in one batch: the producer writes
            // its rows before its flag, so rows read after a flag that shows
            // them are the producer's.  If the batch finds the flag short,
            // later polls read only the flag and the rows once it is set.
            FT_TR_BEGIN();
            const T *src = &sh.ring[w][t % R][0][lane];
            T in[L];
            bool ok = false;
            {
                const int f = lds_ld(&sh.emitted[w]);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
                for (int i = 0; i < L; ++i) in[i] = lds_ld(src + i * kLanes);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                ok = uniform(f) >= t + 1;
            }
            for (int it = 0; !ok && it < kSpinLimit; ++it) {
                __builtin_amdgcn_s_sleep(1);
                if (uniform(lds_ld(&sh.emitted[w])) >= t + 1) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
                    for (int i = 0; i < L; ++i) in[i] = lds_ld(src + i * kLanes);
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    ok = true;
                    break;
                }
                if (aborted(sh)) break;
            }
            FT_TR_END(0);
            if (!ok) {
                fail(a, sh);
                return false;
            }
#pragma unroll
            for (int i = 0; i < L; ++i) x[slot<PH>(L + i)] = in[i];
            lds_st(&sh.consumed[w], t + 1);
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i) x[slot<PH>(L + i)] = T(0);
        }

        {
            FT_TR_BEGIN();
            FT_TR_SYNC();
            FT_TR_END(3);
        }
#ifdef FTAMD_TRACE
        if (c.stamps && t < 160 && lane == 0) c.stamps[3 * t] = wall_clock64();
#endif
        // ---- rotations
        {
            FT_TR_BEGIN();
            const bool full = st.nq == L && st.lanes_all;
            bool fast = full;
            if (DIR == 0) fast = fast && k - (L - 1) >= 0;
            else fast = fast && k + (L - 1) <= st.ne - 2 && k - (L - 1) >= st.lo0;
            if (fast) rotate_tile<PH, false>(st, x, k, sc[PH]);
            else if (DIR == 0 && full && k >= -L) bottom_at<PH, -L>(k, x, sc[PH]);
            else rotate_tile<PH, true>(st, x, k, sc[PH]);
            FT_TR_SYNC();
            FT_TR_END(4);
            FT_TR_COUNT(6);
        }
#ifdef FTAMD_TRACE
        if (c.stamps && t < 160 && lane == 0) c.stamps[3 * t + 1] = wall_clock64();
#endif

        // ---- emit: positions 0 .. L-1 = virtual rows k+L .. k+1, to the ring
        // the next stage (or the exporter) reads
        if (t >= st.emit_off) {
            const int e = t - st.emit_off;
            FT_TR_BEGIN();
            // ring slot free: the reader has taken emission e-R (the last
            // value read is kept, so most tiles read nothing)
            bool ok = e - R + 1 <= c.consumed_seen;
            for (int it = 0; !ok && it < kSpinLimit; ++it) {
                c.consumed_seen = uniform(lds_ld(&sh.consumed[w + 1]));
                if (c.consumed_seen >= e - R + 1) {
                    ok = true;
                    break;
                }
                if (aborted(sh)) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok) {
                fail(a, sh);
                return false;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            T *dst = &sh.ring[w + 1][e % R][0][lane];
#pragma unroll
            for (int i = 0; i < L; ++i) lds_st(dst + i * kLanes, x[slot<PH>(i)]);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            lds_st(&sh.emitted[w + 1], e + 1);
            FT_TR_END(5);
        }
#ifdef FTAMD_TRACE
        if (c.stamps && t < 160 && lane == 0) c.stamps[3 * t + 2] = wall_clock64();
#endif
        return true;
    }
};

template <typename T, int KIND, int DIR, int L, int W, int R>
__global__ void __launch_bounds__(W * kLanes) rot_pipeline(FtamdLaunch a) {
    using P = Pipeline<T, KIND, DIR, L, W, R>;
    __shared__ typename P::Shared sh;
    P::run(a, sh);
}

using KernelFn = void (*)(FtamdLaunch);

// LDS ring depth (tiles a stage may run ahead of the next).
constexpr int kRing = 4;

// Kernel for (L, W) if its LDS ring fits the 160 KB of a gfx950 CU.
template <typename T, int KIND, int DIR, int L, int W>
KernelFn instance() {
    if constexpr (sizeof(typename Pipeline<T, KIND, DIR, L, W, kRing>::Shared) <= 160 * 1024)
        return rot_pipeline<T, KIND, DIR, L, W, kRing>;
    else
        return nullptr;
}

// The (L, W) shapes built for every geometry and direction.
template <typename T, int KIND, int DIR>
KernelFn pick_shape(int L, int W) {
    switch (L * 100 + W) {
        case 208: return instance<T, KIND, DIR, 2, 8>();
        case 216: return instance<T, KIND, DIR, 2, 16>();
        case 404: return instance<T, KIND, DIR, 4, 4>();
        case 408: return instance<T, KIND, DIR, 4, 8>();
        case 416: return instance<T, KIND, DIR, 4, 16>();
        case 804: return instance<T, KIND, DIR, 8, 4>();
        case 808: return instance<T, KIND, DIR, 8, 8>();
        default: return nullptr;
    }
}

template <int KIND>
KernelFn pick_kind(int precision, int dir, int L, int W) {
    if (precision == 0) return dir == 0 ? pick_shape<double, KIND, 0>(L, W) : pick_shape<double, KIND, 1>(L, W);
    return dir == 0 ? pick_shape<float, KIND, 0>(L, W) : pick_shape<float, KIND, 1>(L, W);
}

KernelFn pick_sph(int precision, int dir, int L, int W);
KernelFn pick_tri(int precision, int dir, int L, int W);
KernelFn pick_disk(int precision, int dir, int L, int W);

}  // namespace ftamd
