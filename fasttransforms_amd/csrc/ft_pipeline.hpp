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
// cycles spent waiting on the LDS ring, on the previous task's progress and on
// its own row stores.  The product build compiles all of it away.
#ifdef FTAMD_TRACE
#define FT_TR_BEGIN() const unsigned long long ft_tr0 = __builtin_amdgcn_s_memtime()
#define FT_TR_END(slot) sh.tr[w][slot] += __builtin_amdgcn_s_memtime() - ft_tr0
#else
#define FT_TR_BEGIN() (void)0
#define FT_TR_END(slot) (void)0
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

// Address-space views: the plan table is read through the constant address
// space so that wave-uniform loads become scalar (s_load) loads into SGPRs;
// matrix columns are global memory.
template <typename T> using cptr = const __attribute__((address_space(4))) T *;
template <typename T> using gptr = __attribute__((address_space(1))) T *;

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

// Bounded waits.  ~2^22 polls of >= 64 cycles each is far beyond any
// legitimate wait; on expiry the caller abandons its task.
constexpr int kSpinLimit = 1 << 22;

__device__ __forceinline__ bool lds_wait_ge(const int *flag, int target, const int *abort_flag) {
    for (int it = 0; it < kSpinLimit; ++it) {
        int v = uniform(__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (v >= target) return true;
        if (uniform(__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// Inter-task hand-off through A.  The rows one task passes to the next are
// stored and loaded with device-scope relaxed atomics (sc1 global stores and
// loads), which are coherent per location across CUs and L2s without any
// L2 write-back or invalidation; the producer waits until its row stores are
// acknowledged (s_waitcnt vmcnt(0)) before publishing its progress with a
// device-scope atomic store, and the consumer polls that with device-scope
// atomic loads.  Plain cached accesses with a release/acquire pair would need
// buffer_wbl2 / buffer_inv of the whole L2 on every tile.  Chained tasks still
// run on one XCD, which keeps their traffic in one L2 / MALL slice.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

__device__ __forceinline__ void publish(int *prog, int value) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    wait_stores();
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

// ------------------------------------------------------------------ kernel --
template <typename T, int KIND, int DIR, int L, int W, int R>
struct Pipeline {
    using G = Geo<KIND>;
    static constexpr int STEP = G::STEP;
    static constexpr int WS = 2 * L;   // window rows per lane

    struct Shared {
        T ring[W > 1 ? W - 1 : 1][R][L][kLanes];
        int emitted[W];
        int consumed[W];
        int task;
        int abort_flag;
#ifdef FTAMD_TRACE
        unsigned long long tr[W][4];
#endif
    };

    // Everything one stage wave needs, uniform except the per-lane column data.
    // Rows are "virtual": v = e for hi2lo, v = ne-1-e for lo2hi (lo2hi runs as
    // hi2lo on reversed rows: the transposed rotation of the reversed pair is
    // the forward rotation).
    struct Stage {
        cptr<T> sc;         // table entry 0 (interleaved s, c)
        gptr<T> col;        // this lane's column (nullptr: no column)
        int rs;             // row stride
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
        int lo0;            // DIR 1: lowest virtual row of sequence q0
        bool coh_load;      // rows of A come from the previous task (device-scope loads)
        bool coh_store;     // rows of A go to the next task (device-scope stores)
    };

    // Per-stage control of one sweep.
    struct Ctl {
        bool is_first, is_last;
        int prev_emits, emit_off, lo_valid;
        int *prog_prev, *prog_mine;
        int known;          // progress of the previous task seen so far
        int pend;           // progress to publish once the last tile's stores are acknowledged
    };

    __device__ static int seq_j(int j_first, int q) { return DIR == 0 ? j_first - STEP * q : j_first + STEP * q; }
    __device__ static int real_row(const Stage &st, int v) { return G::real(st.rpar, DIR == 0 ? v : st.ne - 1 - v); }
    __device__ static int top_v(const Stage &st, int j) { return G::top(st.n, st.rpar, j); }

    // Table entry of rotation (virtual pair (v, v+1), sequence j).
    __device__ static long long tab_idx(const Stage &st, int j, int v) {
        return G::tab(st.N, st.rpar, j, DIR == 0 ? v : st.ne - 2 - v);
    }

    template <int PH> __device__ static constexpr int slot(int i) { return (PH * L + i) % WS; }

    // One L x L tile at position k (virtual), phase PH of the register ring:
    // sequence lj rotates the pairs (r, r+1), r = k - lk + lj.
    template <int PH, bool CHECKED>
    __device__ static void rotate_tile(const Stage &st, T (&x)[WS], int k) {
#pragma unroll
        for (int lj = 0; lj < L; ++lj) {
            if (CHECKED && lj >= st.nq) break;
            const int j = DIR == 0 ? st.j0 - STEP * lj : st.j0 + STEP * lj;
            // L consecutive table entries of column j (scalar loads).
            const long long i0 = DIR == 0 ? tab_idx(st, j, k + lj) - (L - 1) : tab_idx(st, j, k + lj);
            cptr<T> p = st.sc + 2 * i0;
            T sv[L], cv[L];
#pragma unroll
            for (int i = 0; i < L; ++i) {
                sv[i] = p[2 * i];
                cv[i] = p[2 * i + 1];
            }
            int lo = 0, hi = 0;
            if (CHECKED) {
                const int top = top_v(st, j);
                lo = DIR == 0 ? 0 : st.ne - 2 - top;
                hi = DIR == 0 ? top : st.ne - 2;
            }
            const bool act = !CHECKED || st.lanes_all || j <= st.jmax;
#pragma unroll
            for (int lk = 0; lk < L; ++lk) {
                const int r = k - lk + lj;
                if (CHECKED && (r < lo || r > hi)) continue;
                const int i = DIR == 0 ? L - 1 - lk : lk;
                const int px = slot<PH>(L + lk - lj);
                const int py = slot<PH>(L + lk - lj - 1);
                if (act) givens<T>(x[px], x[py], sv[i], cv[i]);
            }
        }
    }

    // Row v of this lane's column (0 if outside [lo_valid, ne-1]).  Rows handed
    // over by another task use device-scope accesses (see publish()); the
    // first pass reads what the host wrote and the last pass writes what only
    // the host reads, with ordinary cached accesses.
    __device__ static T load_row(const Stage &st, int v, int lo_valid) {
        if (v < lo_valid || v > st.ne - 1 || st.col == nullptr) return T(0);
        gptr<T> ptr = st.col + (long long)real_row(st, v) * st.rs;
        if (st.coh_load) return __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return *ptr;
    }
    __device__ static void store_row(const Stage &st, int v, int lo_valid, T val) {
        if (v < lo_valid || v > st.ne - 1 || st.col == nullptr) return;
        gptr<T> ptr = st.col + (long long)real_row(st, v) * st.rs;
        if (st.coh_store) __hip_atomic_store(ptr, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *ptr = val;
    }

    // XCD (accelerator complex die) this workgroup runs on.  Each XCD has its
    // own L2; a group's chained tasks are all handed to one XCD so that the
    // rows they pass through A meet in one L2.
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
            if (threadIdx.x < W) {
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
            if (lane == 0) sh.tr[w][0] = sh.tr[w][1] = sh.tr[w][2] = sh.tr[w][3] = 0;
            const unsigned long long ft_w0 = wall_clock64();
            const unsigned long long ft_c0 = __builtin_amdgcn_s_memtime();
#endif
            run_stage(a, sh, g, p, w, lane);
#ifdef FTAMD_TRACE
            if (a.trace && lane == 0) {
                unsigned long long *rec = a.trace + ((long long)task * 16 + w) * 8;
                rec[0] = ft_w0;
                rec[1] = wall_clock64();
                rec[2] = __builtin_amdgcn_s_memtime() - ft_c0;
                rec[3] = sh.tr[w][0];
                rec[4] = sh.tr[w][1];
                rec[5] = sh.tr[w][2];
                rec[6] = ((unsigned long long)x << 32) | (unsigned)(blockIdx.x);
                rec[7] = ((unsigned long long)g << 32) | (unsigned)p;
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

    __device__ static void fail(const FtamdLaunch &a, Shared &sh) {
        __hip_atomic_store(&sh.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        atomicOr(a.err, FTAMD_ERR_TIMEOUT);
    }

    __device__ static void run_stage(const FtamdLaunch &a, Shared &sh, int g, int p, int w, int lane) {
        const FtamdGroup &gd = a.groups[g];
        const int nstages = uniform(gd.nstages);
        const int s = p * W + w;
        if (s >= nstages) return;
        const int last = uniform(min(W, nstages - p * W) - 1);
        const bool is_first = w == 0, is_last = w == last;

        Stage st;
        st.sc = (cptr<T>)(a.sc);
        st.n = uniform(gd.n);
        st.N = a.ntab;
        st.rpar = uniform(gd.rpar);
        st.ne = G::ne(st.n, st.rpar);
        st.rs = uniform(gd.rs);
        const long long off = a.lane_off[(long long)g * kLanes + lane];
        st.col = off < 0 ? nullptr : (gptr<T>)(gd.A) + off;
        st.jmax = a.lane_jmax[(long long)g * kLanes + lane];
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
        } else {
            st.K = st.ne - 1;
            st.Bot = st.ne - 2 - top_v(st, seq_j(j_first, p * W * L));
            st.lo0 = st.ne - 2 - top_v(st, st.j0);
        }
        {
            const int E = st.K + 1 - st.Bot;
            st.ntile = E <= 0 ? 1 : (E + L - 1) / L + 1;
        }
        const int emit_off = DIR == 0 ? 0 : 1;
        // Emits the previous stage makes (only needed by consumers).
        int prev_emits = 0;
        if (!is_first) {
            const int qp = st.q0 - L;
            const int jp = seq_j(j_first, qp);
            const int Kp = DIR == 0 ? top_v(st, jp) : st.ne - 1;
            const int E = Kp + 1 - st.Bot;
            const int Tp = E <= 0 ? 1 : (E + L - 1) / L + 1;
            prev_emits = Tp - emit_off;
        }
        Ctl c;
        c.is_first = is_first;
        c.is_last = is_last;
        c.prev_emits = prev_emits;
        c.emit_off = emit_off;
        c.lo_valid = DIR == 0 ? 0 : max(st.Bot, 0);
        c.prog_prev = p > 0 ? a.prog + uniform(gd.prog_base) + p - 1 : nullptr;
        c.prog_mine = p + 1 < uniform(gd.npass) ? a.prog + uniform(gd.prog_base) + p : nullptr;
        c.known = FTAMD_PROG_NONE;
        c.pend = FTAMD_PROG_NONE;
        st.coh_load = c.prog_prev != nullptr;
        st.coh_store = c.prog_mine != nullptr;

        T x[WS];
#pragma unroll
        for (int i = 0; i < WS; ++i) x[i] = T(0);

        // Initial carried rows (positions 0 .. L-1 = virtual rows K+L .. K+1).
        // hi2lo: untouched by the earlier stages of this task, read from A.
        // lo2hi: above the top row, zero.
        if (DIR == 0) {
            if (c.prog_prev && !global_wait_le(c.prog_prev, st.K + 1, &c.known, &sh.abort_flag)) { fail(a, sh); return; }
#pragma unroll
            for (int i = 0; i < L; ++i) x[i] = load_row(st, st.K + L - i, c.lo_valid);
        }

        int t = 0;
        for (;;) {
            if (!tile<0>(a, sh, st, c, x, t, w, lane)) return;
            if (++t >= st.ntile) break;
            if (!tile<1>(a, sh, st, c, x, t, w, lane)) return;
            if (++t >= st.ntile) break;
        }
        if (is_last && c.prog_mine) publish(c.prog_mine, -1);
    }

    template <int PH>
    __device__ static bool tile(const FtamdLaunch &a, Shared &sh, const Stage &st, Ctl &c, T (&x)[WS], int t, int w,
                                int lane) {
        const int k = st.K - t * L;
        // ---- new rows: positions L .. 2L-1 = virtual rows k .. k-L+1
        if (c.is_first) {
            FT_TR_BEGIN();
            const bool ok = !c.prog_prev || global_wait_le(c.prog_prev, max(k - L + 1, c.lo_valid), &c.known,
                                                            &sh.abort_flag);
            FT_TR_END(1);
            if (!ok) {
                fail(a, sh);
                return false;
            }
#pragma unroll
            for (int i = 0; i < L; ++i) x[slot<PH>(L + i)] = load_row(st, k - i, c.lo_valid);
        } else if (t < c.prev_emits) {
            FT_TR_BEGIN();
            const bool ok = lds_wait_ge(&sh.emitted[w - 1], t + 1, &sh.abort_flag);
            FT_TR_END(0);
            if (!ok) { fail(a, sh); return false; }
            const T *src = &sh.ring[w - 1][t % R][0][lane];
#pragma unroll
            for (int i = 0; i < L; ++i) x[slot<PH>(L + i)] = src[i * kLanes];
            __hip_atomic_store(&sh.consumed[w - 1], t + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
#pragma unroll
            for (int i = 0; i < L; ++i) x[slot<PH>(L + i)] = T(0);
        }

        // ---- rotations
        bool fast = st.nq == L && st.lanes_all;
        if (DIR == 0) fast = fast && k - (L - 1) >= 0;
        else fast = fast && k + (L - 1) <= st.ne - 2 && k - (L - 1) >= st.lo0;
        if (fast) rotate_tile<PH, false>(st, x, k);
        else rotate_tile<PH, true>(st, x, k);

        // ---- emit: positions 0 .. L-1 = virtual rows k+L .. k+1
        if (t >= c.emit_off) {
            const int e = t - c.emit_off;
            if (c.is_last) {
#pragma unroll
                for (int i = 0; i < L; ++i) store_row(st, k + L - i, c.lo_valid, x[slot<PH>(i)]);
                if (c.prog_mine) {
                    // Publish the previous tile's progress once its stores are
                    // acknowledged, leaving this tile's L stores in flight
                    // (vmcnt counts only this wave's stores here).  Edge tiles,
                    // which may store fewer rows, wait for everything.
                    FT_TR_BEGIN();
                    const bool full = !c.is_first && k + 1 >= c.lo_valid && k + L <= st.ne - 1;
                    if (full) {
                        if (c.pend != FTAMD_PROG_NONE) {
                            __atomic_signal_fence(__ATOMIC_SEQ_CST);
                            __builtin_amdgcn_s_waitcnt(0x0F70 | L);  // vmcnt(L)
                            __hip_atomic_store(c.prog_mine, c.pend, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        c.pend = k + 1;
                    } else {
                        publish(c.prog_mine, k + 1);
                        c.pend = FTAMD_PROG_NONE;
                    }
                    FT_TR_END(2);
                }
            } else {
                FT_TR_BEGIN();
                const bool ok = e < R || lds_wait_ge(&sh.consumed[w], e - R + 1, &sh.abort_flag);
                FT_TR_END(0);
                if (!ok) { fail(a, sh); return false; }
                T *dst = &sh.ring[w][e % R][0][lane];
#pragma unroll
                for (int i = 0; i < L; ++i) dst[i * kLanes] = x[slot<PH>(i)];
                __hip_atomic_store(&sh.emitted[w], e + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
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

// Kernel for (L, W) if its LDS ring fits the 160 KB of a gfx950 CU.
template <typename T, int KIND, int DIR, int L, int W>
KernelFn instance() {
    constexpr int R = 4;
    if constexpr (sizeof(typename Pipeline<T, KIND, DIR, L, W, R>::Shared) <= 160 * 1024)
        return rot_pipeline<T, KIND, DIR, L, W, R>;
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
