// ft_systolic.hpp -- band-systolic rotation engine for the whole-matrix
// executors (ft_execute_{sph,tri,disk}_{hi2lo,lo2hi}, drivers.c:58-110,
// 314-326, 414-430).
//
// What is computed (src/rotations/rotations_default.c:31-71): every column
// applies a chain of rotation sequences; sequence j rotates the row pairs
// (e, e+1) of its class rows e = top(j) .. 0 (hi2lo) or 0 .. top(j) (lo2hi).
// Consecutive sequences of a chain differ by one in top(j), and every column
// of a class applies a suffix (hi2lo) or prefix (lo2hi) of one common list of
// sequences k = 0 .. K-1.
//
// How it maps onto the GPU:
//
//  * Lanes are sequences.  A wave runs one band: FTAMD_SYS_LANES*S
//    consecutive sequences, S per lane, for the 8 columns of a chain (4 rows
//    of 16 lanes, 2 columns per lane).  Each lane holds, per sequence and
//    column, the row its sequence carries; every step every sequence rotates
//    (carried, fresh) and emits the other row to the next sequence: inside a
//    lane in the same step, from lane to lane by a DPP row shift one step
//    later.  Rows enter the band as a stream at lane 0 of each row (rows
//    R .. 0 for hi2lo, 0 .. R for lo2hi) and leave it at lane 15, S+1 steps
//    per lane later.  Steps where a sequence has no rotation (before its rows
//    arrive, after its last row, or for a column that does not apply it) use
//    the identity (s, c) = (0, 1), which passes the stream through with the
//    same delay; so the whole band is one uniform loop, and the table of a
//    band is laid out step-major ([step][lane][slot]) with identities in the
//    gaps.
//  * s and c are per lane (each lane is a different sequence): no broadcast;
//    the 4 rows of a wave read the same 16 lanes' entries.
//  * A chain (one workgroup) runs its bands as waves; band b+1 reads band
//    b's output stream through an LDS ring as it is produced, so a chain's
//    bands overlap and the chain's latency is about one band's stream plus
//    (S+1)*LANES steps per band.  The first band reads A, the rows that leave
//    the chain for good are stored to A.
//
// Arithmetic: the reference's rotation in fused form, x' = fma(c, x, s*y),
// y' = fma(c, y, -(s*x)) (src/rotations/rotations_AVX512F.c:4-15).  With
// x = the fresh row and y = the carried one this is both the hi2lo rotation of
// rows (e, e+1) and the transposed lo2hi rotation of (e+1, e), so one loop
// serves both directions.  Compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "ft_hip_abi.h"

namespace ftsys {

constexpr int kLanes = FTAMD_SYS_LANES;
constexpr int kRows = 4;
constexpr int kCols = 2;
constexpr int kWaves = FTAMD_SYS_WAVES;
constexpr int kZ = 32;                 // ring positions between two bands
constexpr int kU = FTAMD_SYS_UNROLL;   // steps per loop iteration
constexpr int kDin = 2;                // input positions read ahead
constexpr int kChunk = 16;             // matrix rows fetched per lane-parallel load


template <typename T> using Pair = T __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T> __device__ __forceinline__ T lds_ld(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T> __device__ __forceinline__ void lds_st(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Lanes 1..15 of every row of 16 take src of the lane before; lane 0 of each
// row keeps old (DPP row_shr:1 without bound control).
template <typename T> __device__ __forceinline__ T shift_in(T old, T src) {
    return __builtin_amdgcn_update_dpp(old, src, 0x111, 0xf, 0xf, false);
}

template <typename T> __device__ __forceinline__ T fmaT(T a, T b, T c);
template <> __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T, int S, bool ROWMASK>
struct Engine {
    static constexpr int B = kLanes * S;  // sequences per band
    // a row that enters lane 0 at step p leaves lane 15 at step p + kOutDelay:
    // S+1 steps per lane (the lane's S sequences pass it on within a step,
    // the DPP shift to the next lane takes a step)
    static constexpr int kOutDelay = (S + 1) * (kLanes - 1) + S;

    // table bytes of one step (all 16 lanes) and steps per 1 KB LDS-DMA load
    static constexpr int kStepBytes = kLanes * S * 2 * (int)sizeof(T);
    static constexpr int kSPD = 1024 / kStepBytes;
    // steps of table entries staged ahead in LDS (4 DMA blocks)
    static constexpr int kD = 4 * kSPD;
    static_assert(kSPD >= 1 && kU % kSPD == 0 && kD <= FTAMD_SYS_PREFETCH,
                  "DMA blocks must tile the unroll and the tables' spare steps");
    // a block's table slots are at most one wrap of the LDS ring, and its
    // DMAs use the compile-time offsets 0, 2, 4, 6 (kSPD 2) or 0, 4 (kSPD 4)
    static_assert(kD >= kU && kD % kU == 0 && (kSPD == 2 || kSPD == 4),
                  "table staging supports 2 or 4 steps per DMA and a ring of whole blocks");


    static constexpr int kMaxRows = sizeof(T) == 8 ? FTAMD_SYS_MAX_ROWS_FP64 : FTAMD_SYS_MAX_ROWS_FP32;

    struct Shared {
        // table entries of the next kD steps of each wave, written by LDS DMA
        // (first in the block: the DMA's LDS address is taken from M0)
        Pair<T> scr[kWaves][kD][kLanes][S];
        T ring[kWaves][kZ][kRows][kCols];  // ring w: input stream of band (wave) w
        T stage[kMaxRows][kRows][kCols];   // rows leaving the chain, by class row, until stored to A
        int avail[kWaves];                 // positions of ring w written by wave w-1 (exclusive end)
        int freed[kWaves];                 // positions of ring w wave w has read (exclusive end)
        int abort_flag;
    };

    // The table entries of one lane for one step: S (s, c) pairs.
    struct SC {
        Pair<T> p[S];
        __device__ Pair<T> pair(int s) const { return p[s]; }
    };

    // A wave whose wait expires raises the workgroup's abort flag (every
    // other wait then ends too) and reports the error once its loop is left.
    __device__ static void fail(Shared &sh) { lds_st(&sh.abort_flag, 1); }

#ifdef FTAMD_TRACE
    // per-wave accumulators of wait clocks (instrumented builds)
    __device__ static unsigned long long &tr_in() {
        static __shared__ unsigned long long v[kWaves];
        return v[threadIdx.x >> 6];
    }
    __device__ static unsigned long long &tr_out() {
        static __shared__ unsigned long long v[kWaves];
        return v[threadIdx.x >> 6];
    }
    __device__ static unsigned long long &tr_dma() {
        static __shared__ unsigned long long v[kWaves];
        return v[threadIdx.x >> 6];
    }
#endif

    // Waits until *ctr >= need (seen caches the last value read), at most
    // `limit` polls.
    __device__ static bool wait_ge(Shared &sh, int *ctr, int need, int &seen, int limit) {
#ifdef FTAMD_TRACE
        struct Acc {
            unsigned long long t0, *dst;
            __device__ ~Acc() {
                if ((threadIdx.x & 63) == 0) *dst += __builtin_amdgcn_s_memtime() - t0;
            }
        } acc{__builtin_amdgcn_s_memtime(), (ctr >= &sh.avail[0] && ctr < &sh.avail[kWaves]) ? &tr_in() : &tr_out()};
#endif
        for (int it = 0; seen < need; ++it) {
            if (it >= limit || uniform(lds_ld(&sh.abort_flag))) return false;
            __builtin_amdgcn_s_sleep(1);
            seen = uniform(lds_ld(ctr));
        }
        return true;
    }

    // Table entries reach the waves by LDS DMA (global_load_lds_dwordx4: 64
    // lanes x 16 bytes, the kSPD steps of one block, straight into the
    // wave's LDS ring, no registers in flight), issued by hand with hand
    // counted waits: in this loop the compiler's own wait insertion loses
    // count and would drain every load at every step, defeating the
    // prefetch.  Each block's DMA is the only vector-memory instruction the
    // loop issues besides the first band's matrix fetches (which only add
    // to the count), so a block is in LDS once at most as many DMAs as were
    // issued after it are outstanding.
    // `after` is a value the caller computed from the entries last read out
    // of the block being refilled: tying it to the DMA keeps the DMA after
    // those reads have returned.
    __device__ static void dma_block(Shared &sh, int w, const char *tab_band, int step, unsigned lane16, T after) {
        const int slot = step % kD;
        const unsigned lds = (unsigned)(uintptr_t)&sh.scr[w][slot][0][0];
        const char *src = tab_band + (long long)uniform(step) * kStepBytes;
        // M0 (the DMA's LDS address) needs one wait state before the DMA
        // reads it; nothing else in this kernel uses M0.
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2"
                     :
                     : "s"(uniform((int)lds)), "v"(lane16), "s"(src), "v"(after)
                     : "memory");
    }
    // The same, with the step split into a uniform base and a compile-time
    // part K: the instruction's immediate offset is added to both the global
    // and the LDS address, so M0 stays at the base slot and offset:K*bytes
    // selects step base+K on both sides.
    template <int K>
    __device__ static void dma_block_at(unsigned lds_base, const char *src_base, unsigned lane16, T after) {
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2 offset:%4"
                     :
                     : "s"(lds_base), "v"(lane16), "s"(src_base), "v"(after), "n"(K * kStepBytes)
                     : "memory");
    }
    // Called before the first read of a block, one step before its first
    // step, when the two blocks after it have been requested and the third
    // not yet.
    __device__ static void wait_block() {
#ifdef FTAMD_TRACE
        const unsigned long long c0 = __builtin_amdgcn_s_memtime();
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kD / kSPD - 2) : "memory");
        if ((threadIdx.x & 63) == 0) tr_dma() += __builtin_amdgcn_s_memtime() - c0;
#else
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kD / kSPD - 2) : "memory");
#endif
    }
    // Lane 15 of each row of 16 (the row's last sequence) to every lane of
    // the row (DPP row_newbcast:15).
    __device__ static T row_end(T v) { return __builtin_amdgcn_update_dpp(v, v, 0x15f, 0xf, 0xf, false); }
    __device__ static unsigned lds_addr(const void *p) { return (unsigned)(uintptr_t)p; }

    __device__ static void read_sc(Shared &sh, int w, int step, int i, SC &sc) {
        const Pair<T> *p = &sh.scr[w][step % kD][i][0];
#pragma unroll
        for (int s = 0; s < S; ++s) sc.p[s] = p[s];
    }

    // One step: rotate every (sequence, column) of the lane.  X0: the fresh
    // rows of the lane's first sequence (the last sequence of the lane before
    // emitted them at the previous step, or the stream at lane 0); Y: carried
    // rows; sequence s > 0 takes as its fresh row what sequence s-1 emits in
    // this step.
    // MASK 0: every column applies every sequence of the band; 1: the two
    // columns of a lane row apply the same sequences (act[s][0]); 2: per
    // column.  A sequence a column does not apply is the identity (0, 1).
    template <int MASK>
    __device__ static void rotate(T (&Y)[S][kCols], T (&E)[S][kCols], const T (&X0)[kCols], const SC &sc,
                                  const bool (&act)[S][kCols]) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const Pair<T> e = sc.pair(s);
            const T sn_row = MASK == 1 ? (act[s][0] ? e.x : T(0)) : e.x;
            const T cs_row = MASK == 1 ? (act[s][0] ? e.y : T(1)) : e.y;
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                const T x = s == 0 ? X0[c] : E[s - 1][c];
                const T y = Y[s][c];
                T sn = sn_row, cs = cs_row;
                if (MASK == 2) {
                    sn = act[s][c] ? e.x : T(0);
                    cs = act[s][c] ? e.y : T(1);
                }
                Y[s][c] = fmaT<T>(cs, x, sn * y);
                E[s][c] = fmaT<T>(cs, y, -(sn * x));
            }
        }
    }

    template <int MASK>
    __device__ static void run_band(const FtamdSysLaunch &a, Shared &sh, const FtamdSysChain &ch, int w, int lane) {
        const int r = lane >> 4, i = lane & 15;
        const int nb = uniform(ch.nb);
        const int b = uniform(ch.band0) + w;
        const int R = uniform(a.bands[b].R);
        const int steps = uniform(a.bands[b].steps);
        const int gbase = uniform(a.bands[b].gbase);
        const bool first = w == 0, last = w == nb - 1;
        const int dir = uniform(ch.dir);
        const int Rprev = first ? 0 : uniform(a.bands[b - 1].R);
        const int Rnext = last ? -1 : uniform(a.bands[b + 1].R);
        // stream positions 0 .. npa-1 come from A, the rest from the previous band
        const int npa = first ? R + 1 : (dir == 0 ? R - Rprev : 0);
        // the next band's position of our output position p is p + npa_next;
        // outputs beyond the next band's stream leave the chain
        const int npa_next = last ? 0 : (dir == 0 ? Rnext - R : 0);
        const int rstep = uniform(ch.rstep), rpar = uniform(ch.rpar);
        const long long lda = uniform(ch.lda);
        const int spin = uniform(a.spin_limit);  // polls before a wait gives up
        T *A = (T *)a.A + ch.aoff;
        const int colv[kCols] = {ch.col[2 * r], ch.col[2 * r + 1]};
        const char *tab_band = (const char *)a.tab + a.bands[b].tab * (long long)sizeof(Pair<T>);
        const unsigned lane16 = (unsigned)lane * 16u;
        const unsigned r16 = (unsigned)r * (kCols * sizeof(T));  // row r's place in a ring or stage entry

        // matrix element of class row e of column c of this lane's row
        auto aptr = [&](int c, int e) -> T * { return A + (long long)colv[c] * lda + (long long)rstep * e + rpar; };
        // class row of stream position p
        auto row_of = [&](int p) { return dir == 0 ? R - p : p; };

        // with MASK 1 column 0 of a row carries the row's mask (column 1's
        // when column 0 is absent)
        bool act[S][kCols];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                const int g = gbase + S * i + s;
                const int k = MASK == 1 && c == 0 && colv[0] < 0 ? 2 * r + 1 : 2 * r + c;
                act[s][c] = g >= ch.glo[k] && g <= ch.ghi[k];
            }

        T(*in_ring)[kRows][kCols] = sh.ring[w];
        T(*out_ring)[kRows][kCols] = last ? sh.ring[w] : sh.ring[w + 1];
        int avail_seen = 0, freed_seen = 0;
        bool ok = true;
#ifdef FTAMD_TRACE
        if (lane == 0) tr_in() = tr_out() = tr_dma() = 0;
        const unsigned long long tr_start = wall_clock64(), tr_c0 = __builtin_amdgcn_s_memtime();
#endif

        // ---- stream positions from A: lane i of each row fetches position
        // p0 + i of a chunk of 16.  (Loads only: this wave issues
        // no global store before its loop ends, so its waits on the table
        // loads count exactly.)
        auto fetch = [&](int p0, T (&v)[kCols]) {
            const int p = p0 + i;
#pragma unroll
            for (int c = 0; c < kCols; ++c)
                v[c] = (p < npa && p <= R && colv[c] >= 0) ? *aptr(c, row_of(p)) : T(0);
        };
        auto put = [&](int p0, const T (&v)[kCols]) {
            const int p = p0 + i;
            if (p < npa) {
#pragma unroll
                for (int c = 0; c < kCols; ++c) lds_st(&in_ring[p % kZ][r][c], v[c]);
            }
        };
        // chunk k (positions 16k .. 16k+15) is fetched 16 steps before it is
        // written to the ring, 32 steps before it is read
        T chunk[kCols];
        {
            T v[kCols];
            fetch(0, v);
            put(0, v);
            fetch(kChunk, chunk);
        }

        // ---- input prefetch: positions 0 .. kDin-1
        T inbuf[kDin][kCols];
        auto ensure_in = [&](int p) -> bool {
            bool good = true;
            if (p >= npa && p <= R && p >= avail_seen) good = wait_ge(sh, &sh.avail[w], p + 1, avail_seen, spin);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            return good;
        };
        auto read_in = [&](int p, T (&v)[kCols]) {
#pragma unroll
            for (int c = 0; c < kCols; ++c) v[c] = lds_ld(&in_ring[p % kZ][r][c]);
        };
#pragma unroll
        for (int d = 0; d < kDin; ++d) {
            ok = ok && ensure_in(d);
            read_in(d, inbuf[d]);
        }

        // ---- table prefetch: steps 0 .. kD-1
#pragma unroll
        for (int d = 0; d < kD; d += kSPD) dma_block(sh, w, tab_band, d, lane16, T(0));

        T Y[S][kCols], E[S][kCols];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < kCols; ++c) Y[s][c] = E[s][c] = T(0);

        // entries of step 0 (read one step ahead from here on)
        SC scn;
        wait_block();
        read_sc(sh, w, 0, i, scn);

        // Steps run in blocks of kU.  The waits are made once per block, so
        // a step is only its DPP shifts, LDS reads, rotations and the
        // broadcast of its output.
        int written = 0;  // ring positions of the next band written (exclusive end)
        for (int t0 = 0; t0 < steps; t0 += kU) {
            if (t0 % kChunk == 0) {
                // positions t0+16 .. t0+31 to the ring, into the slots of
                // positions t0-16 .. t0-1 (all read); fetch the next 16
                const int p0 = t0 + kChunk;
                if (p0 < npa) {
                    put(p0, chunk);
                    fetch(p0 + kChunk, chunk);
                }
            }
            // every input position the block reads ahead is in the ring
            {
                const int pl = min(t0 + kU - 1 + kDin, R);
                if (pl >= npa && pl >= avail_seen && !wait_ge(sh, &sh.avail[w], pl + 1, avail_seen, spin)) {
                    ok = false;
                    break;
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
            }
            // the block's outputs, positions po0 .. po0+kU-1: those in
            // [0, R] go to the next band's ring up to position ring_hi,
            // beyond it (or from the last band) they leave the chain
            const int po0 = t0 - kOutDelay;
            const int ring_hi = last ? -1 : min(R, Rnext - npa_next);
            {
                const int hi = min(po0 + kU - 1, ring_hi);
                if (hi >= max(po0, 0)) {
                    const int P1 = hi + npa_next;
                    if (P1 - kZ >= freed_seen && !wait_ge(sh, &sh.freed[w + 1], P1 - kZ + 1, freed_seen, spin)) {
                        ok = false;
                        break;
                    }
                    written = P1 + 1;
                }
            }
            // the addresses of the block's LDS accesses: ring slots of the
            // reads (the last kDin steps may wrap to the ring's start), table
            // slots, DMA targets and sources
            const T *pin = &in_ring[t0 % kZ][r][0];
            const T *pin_hi = (t0 % kZ == kZ - kU) ? pin - kZ * kRows * kCols : pin;
            const int sb = t0 % kD;
            const Pair<T> *psc = &sh.scr[w][sb][i][0];
            const Pair<T> *psc_hi = (sb == kD - kU) ? psc - kD * kLanes * S : psc;
            const unsigned m0b = lds_addr(&sh.scr[w][0][0][0]) + (unsigned)sb * kStepBytes;
            const char *srcb = tab_band + (long long)(t0 + kD) * kStepBytes;  // step t0+kD+K: the DMA of step u, K = u+1-kSPD
            // step u's output (the row-end values of the last sequence, lane
            // 15 of each row) is broadcast along its row and kept by lane u;
            // the block's kU outputs are stored together after its last step
            T outv[kCols];
#pragma unroll
            for (int c = 0; c < kCols; ++c) outv[c] = T(0);
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                // keep each step's instructions together: the entries and
                // rows read for the next steps then have a whole step of
                // issue time to arrive before their first use
                __builtin_amdgcn_sched_barrier(0);
                T X0[kCols];
#pragma unroll
                for (int c = 0; c < kCols; ++c) X0[c] = shift_in(inbuf[u % kDin][c], E[S - 1][c]);
                {
                    const T *q = (u + kDin >= kU ? pin_hi : pin) + (u + kDin) * kRows * kCols;
#pragma unroll
                    for (int c = 0; c < kCols; ++c) inbuf[u % kDin][c] = lds_ld(q + c);
                }
                const SC sc = scn;
                if ((u + 1) % kSPD == 0) wait_block();
                {
                    const Pair<T> *q = (u + 1 >= kU ? psc_hi : psc) + (u + 1) * kLanes * S;
#pragma unroll
                    for (int s2 = 0; s2 < S; ++s2) scn.p[s2] = q[s2];
                }
                // the reads above are issued before this step's arithmetic
                __builtin_amdgcn_sched_barrier(0);
                rotate<MASK>(Y, E, X0, sc, act);
                if (u % kSPD == kSPD - 1) {
                    // u is a constant once the loop is unrolled
                    switch (u + 1 - kSPD) {
                        case 0: dma_block_at<0>(m0b, srcb, lane16, E[0][0]); break;
                        case 2: dma_block_at<2>(m0b, srcb, lane16, E[0][0]); break;
                        case 4: dma_block_at<4>(m0b, srcb, lane16, E[0][0]); break;
                        case 6: dma_block_at<6>(m0b, srcb, lane16, E[0][0]); break;
                        default: break;
                    }
                }
#pragma unroll
                for (int c = 0; c < kCols; ++c) {
                    const T b = row_end(E[S - 1][c]);
                    outv[c] = i == u ? b : outv[c];
                }
            }
            // lane i < kU of each row stores output position po0 + i: to the
            // next band's ring up to ring_hi, else (leaving the chain) to the
            // stage; positions outside [0, R] carry no row
            {
                const int po = po0 + i;
                if (i < kU && po >= 0 && po <= R) {
                    T *dst = po <= ring_hi ? &out_ring[(po + npa_next) % kZ][r][0] : &sh.stage[row_of(po)][r][0];
#pragma unroll
                    for (int c = 0; c < kCols; ++c) lds_st(dst + c, outv[c]);
                }
            }
            if (!ok) break;
            // publish what the block wrote and read
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (!last && lane == 0) lds_st(&sh.avail[w + 1], written);
            if (lane == 0) lds_st(&sh.freed[w], t0 + kU);
        }

        if (!ok) {
            fail(sh);
            atomicOr(a.err, FTAMD_ERR_TIMEOUT);
            return;
        }
#ifdef FTAMD_TRACE
        if (a.trace && lane == 0) {
            unsigned long long *rec = a.trace + ((long long)blockIdx.x * kWaves + w) * 8;
            rec[0] = tr_start;
            rec[1] = wall_clock64();
            rec[2] = __builtin_amdgcn_s_memtime() - tr_c0;
            rec[3] = tr_in();
            rec[4] = tr_out();
            rec[5] = steps;
            rec[6] = R;
            rec[7] = tr_dma();
        }
#endif
        // ---- rows that left the chain here, staged by class row, to A:
        // all rows 0 .. R from the last band; none from the other hi2lo bands
        // (every row they emit goes on to the next band); rows Rnext+1 .. R
        // from the other lo2hi bands
        const int lo = last ? 0 : (dir == 0 ? R + 1 : Rnext + 1);
        for (int e0 = lo; e0 <= R; e0 += kLanes) {
            const int e = e0 + i;
            if (e <= R) {
#pragma unroll
                for (int c = 0; c < kCols; ++c)
                    if (colv[c] >= 0) *aptr(c, e) = lds_ld(&sh.stage[e][r][c]);
            }
        }
    }

    __device__ static void run(const FtamdSysLaunch &a, Shared &sh) {
        {
            T *rp = &sh.ring[0][0][0][0];
            for (int x = threadIdx.x; x < kWaves * kZ * kRows * kCols; x += blockDim.x) rp[x] = T(0);
            if (threadIdx.x < kWaves) {
                sh.avail[threadIdx.x] = 0;
                sh.freed[threadIdx.x] = 0;
            }
            if (threadIdx.x == 0) sh.abort_flag = 0;
        }
        __syncthreads();
        const FtamdSysChain &ch = a.chains[blockIdx.x];
        const int lane = threadIdx.x & 63;
        const int w = uniform(threadIdx.x >> 6);
        if (w >= uniform(ch.nb)) return;
        switch (uniform(ch.prio)) {
            case 3: __builtin_amdgcn_s_setprio(3); break;
            case 2: __builtin_amdgcn_s_setprio(2); break;
            case 1: __builtin_amdgcn_s_setprio(1); break;
            default: break;
        }
        // a band needs masks if some column of the chain does not apply some
        // of its sequences; with ROWMASK one mask per lane row serves both
        // columns (sph and disk: the two columns of a row are the cosine and
        // sine columns of one order, so they apply the same sequences)
        const int gbase = uniform(a.bands[uniform(ch.band0) + w].gbase);
        const int glast = gbase + uniform(a.bands[uniform(ch.band0) + w].nseq) - 1;
        bool masked = false;
        for (int k = 0; k < FTAMD_SYS_COLS; ++k)
            masked |= ch.col[k] >= 0 && (ch.glo[k] > gbase || ch.ghi[k] < glast);
        if (!masked) run_band<0>(a, sh, ch, w, lane);
        else run_band<ROWMASK ? 1 : 2>(a, sh, ch, w, lane);
        __builtin_amdgcn_s_setprio(0);
    }
};

template <typename T, int S, bool ROWMASK>
__global__ void __launch_bounds__(kWaves * 64) sys_kernel(FtamdSysLaunch a) {
    using E = Engine<T, S, ROWMASK>;
    __shared__ typename E::Shared sh;
    E::run(a, sh);
}

}  // namespace ftsys
