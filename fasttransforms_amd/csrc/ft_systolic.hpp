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
//  * A chain runs its bands as waves of one workgroup; band b+1 reads band
//    b's output stream through an LDS ring as it is produced, so a chain's
//    bands overlap and the chain's latency is about one band's stream plus
//    (S+1)*LANES steps per band.  The first band reads A; rows that leave the
//    chain are stored to A once per 8-step block by the band that emits them.
//  * A long chain runs in parts of at most a set number of bands, each part
//    on its own workgroup: the last band of a part stores every row it emits
//    to A and publishes a progress counter, and the first band of the next
//    part reads those rows back from A as they are published.  This spreads a
//    long chain's waves over several CUs and lifts the bound of 16 bands (one
//    workgroup's waves) per chain.
//  * Workgroups are persistent (about one per CU).  The host deals the parts
//    to slots, per XCD, and a workgroup takes its XCD's slots in order; a
//    slot's parts stream through the workgroup's waves one after another.
//    The parts of one chain run in increasing slots of one XCD, so a waiting
//    part only ever waits on a slot that a running workgroup holds or has
//    finished, and the rows they hand over stay in that XCD's L2.
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
// ring positions between two bands (a multiple of kU, at least 2*kChunk):
// 64 lets a band run up to 64 rows ahead of the next (64 KB of LDS per
// workgroup; 3-9 % faster than 32 at n = 1024, 80 no better)
constexpr int kZ = 64;
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
    static_assert(kZ % kU == 0 && kZ >= 2 * kChunk, "ring of whole blocks, two fetch chunks deep");
    static_assert(kSPD >= 1 && kU % kSPD == 0 && kD <= FTAMD_SYS_PREFETCH,
                  "DMA blocks must tile the unroll and the tables' spare steps");
    // a block's table slots are at most one wrap of the LDS ring, and its
    // DMAs use the compile-time offsets 0, 2, 4, 6 (kSPD 2) or 0, 4 (kSPD 4)
    static_assert(kD >= kU && kD % kU == 0 && (kSPD == 2 || kSPD == 4),
                  "table staging supports 2 or 4 steps per DMA and a ring of whole blocks");


    struct Shared {
        // table entries of the next kD steps of each wave, written by LDS DMA
        // (first in the block: the DMA's LDS address is taken from M0)
        // step-major; within a step slot-major for FP64 ([slot][lane]: a
        // lane's 16-byte pair next to the next lane's, so a 128-bit read of
        // 8 lanes spans all banks once) and lane-major for FP32 ([lane][slot]:
        // a lane's S pairs in one 128-bit read); the tables are laid out
        // the same way (FTAMD_SYS_SLOT_MAJOR)
        Pair<T> scr[kWaves][kD][kLanes * S];
        T ring[kWaves][kZ][kRows][kCols];  // ring g: input stream of the band wave g runs
        int avail[kWaves];                 // positions of ring g written by wave g-1 (exclusive end)
        int freed[kWaves];                 // positions of ring g wave g has read (exclusive end)
        int abort_flag;
        int slot;                          // the workgroup's current slot (ticket)
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
            // far behind (a band waiting for the chain's fill): poll rarely,
            // so waiting waves leave the issue slots to the working ones
            if (need - seen > 2 * kU) __builtin_amdgcn_s_sleep(16);
            else __builtin_amdgcn_s_sleep(1);
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
        const unsigned lds = (unsigned)(uintptr_t)&sh.scr[w][slot][0];
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
    // (Every lane has a source, so the old value is irrelevant: one DPP move.)
    __device__ static T row_end(T v) { return __builtin_amdgcn_update_dpp(T(0), v, 0x15f, 0xf, 0xf, true); }
    __device__ static unsigned lds_addr(const void *p) { return (unsigned)(uintptr_t)p; }

    // position of (lane i, slot s) in a step's entries, and of lane i's first
    static constexpr bool kSlotMajor = sizeof(T) == 8;
    static constexpr int kSlotStride = kSlotMajor ? kLanes : 1;
    __device__ static int lane_entry(int i) { return kSlotMajor ? i : i * S; }

    __device__ static void read_sc(Shared &sh, int w, int step, int i, SC &sc) {
        const Pair<T> *p = &sh.scr[w][step % kD][lane_entry(i)];
#pragma unroll
        for (int s = 0; s < S; ++s) sc.p[s] = p[s * kSlotStride];
    }

    // Progress counters of split chains.  The producing band stores its rows
    // to A with device-scope stores and, once they are complete, publishes
    // how many it has stored with a device-scope store; the consuming band
    // polls the counter with a scalar load that bypasses the scalar cache
    // (GLC), so that polling never waits on the vector-memory counter the
    // table DMAs run on, and then reads the rows with device-scope loads.
    // Both parts run on one XCD, so they meet in one L2.
    __device__ static int poll_prog(const int *p) {
        int v;
        asm volatile("s_load_dword %0, %1, 0x0 glc\n\ts_waitcnt lgkmcnt(0)" : "=s"(v) : "s"(p) : "memory");
        return v;
    }
    __device__ static bool wait_prog(Shared &sh, const int *p, int need, int &seen, int limit) {
        for (int it = 0; seen < need; ++it) {
            if (it >= limit || uniform(lds_ld(&sh.abort_flag))) return false;
            if (it > 0) __builtin_amdgcn_s_sleep(2);
            seen = uniform(poll_prog(p));
        }
        return true;
    }

    // One step: rotate every (sequence, column) of the lane.  X0: the fresh
    // rows of the lane's first sequence (the last sequence of the lane before
    // emitted them at the previous step, or the stream at lane 0); Y: carried
    // rows; sequence s > 0 takes as its fresh row what sequence s-1 emits in
    // this step.
    // MASK 0: every column applies every sequence of the band; 1: the two
    // columns of a lane row apply the same sequences (mask [s][0]); 2: per
    // column.  A sequence a column does not apply is the identity (0, 1):
    // with m = 1 where the column applies the sequence and 0 where not, the
    // applied pair is (s*m, fma(c, m, 1-m)), exact for both values of m and
    // two FP64 operations where a select of two doubles takes four.
    template <int MASK>
    __device__ static void rotate(T (&Y)[S][kCols], T (&E)[S][kCols], const T (&X0)[kCols], const SC &sc,
                                  const T (&m)[S][kCols], const T (&m1)[S][kCols]) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const Pair<T> e = sc.pair(s);
            const T sn_row = MASK == 1 ? e.x * m[s][0] : e.x;
            const T cs_row = MASK == 1 ? fmaT<T>(e.y, m[s][0], m1[s][0]) : e.y;
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                const T x = s == 0 ? X0[c] : E[s - 1][c];
                const T y = Y[s][c];
                T sn = sn_row, cs = cs_row;
                if (MASK == 2) {
                    sn = e.x * m[s][c];
                    cs = fmaT<T>(e.y, m[s][c], m1[s][c]);
                }
                Y[s][c] = fmaT<T>(cs, x, sn * y);
                E[s][c] = fmaT<T>(cs, y, -(sn * x));
            }
        }
    }

    // w: the band's index in its chain; g: the wave (and ring) index in the
    // workgroup.  Ring positions are counted across the chains a workgroup
    // streams: this band's input stream starts at position in_base of ring g,
    // its output (the next band's input) at out_base of ring g+1, both
    // multiples of kZ.  LINK: the band hands rows to or takes them from
    // another part of its chain (through A).  Returns false if a wait expired
    // (abort raised).
    template <int MASK, bool LINK>
    __device__ static bool run_band(const FtamdSysLaunch &a, Shared &sh, const FtamdSysChain &ch, int w, int g,
                                    int lane, int in_base, int out_base) {
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
        // a part after the first of its chain: its positions from npa_own on
        // are the previous part's output position p - npa_own, published on
        // counter split_in; the last band of a part before the last stores
        // every row it emits and publishes them on counter split_out
        const int split_in = LINK && first ? uniform(ch.split_in) : -1;
        const int split_out = LINK && last ? uniform(ch.split_out) : -1;
        const int npa_own = split_in >= 0 && dir == 0 ? R - uniform(a.bands[b - 1].R) : 0;
        int prog_seen = 0, published = 0, pub = 0;
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

        // matrix element of class row e of column c of this lane's row
        auto aptr = [&](int c, int e) -> T * { return A + (long long)colv[c] * lda + (long long)rstep * e + rpar; };
        // class row of stream position p
        auto row_of = [&](int p) { return dir == 0 ? R - p : p; };

        // with MASK 1 column 0 of a row carries the row's mask (column 1's
        // when column 0 is absent)
        T m[S][kCols], m1[S][kCols];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                const int q = gbase + S * i + s;
                const int k = MASK == 1 && c == 0 && colv[0] < 0 ? 2 * r + 1 : 2 * r + c;
                const bool on = q >= ch.glo[k] && q <= ch.ghi[k];
                m[s][c] = on ? T(1) : T(0);
                m1[s][c] = on ? T(0) : T(1);
            }

        T(*in_ring)[kRows][kCols] = sh.ring[g];
        T(*out_ring)[kRows][kCols] = last ? sh.ring[g] : sh.ring[g + 1];
        int avail_seen = 0, freed_seen = 0;  // last counter values read (global positions)
        bool ok = true;
#ifdef FTAMD_TRACE
        if (lane == 0) tr_in() = tr_out() = tr_dma() = 0;
        const unsigned long long tr_start = wall_clock64(), tr_c0 = __builtin_amdgcn_s_memtime();
#endif

        // ---- stream positions from A: lane i of each row fetches position
        // p0 + i of a chunk of 16.  (Loads and stores of A only make the
        // waits on the table loads stricter: see the output stores.)
        auto fetch = [&](int p0, T (&v)[kCols]) -> bool {
            const int p = p0 + i;
            if (split_in >= 0) {
                const int need = min(p0 + kChunk - 1, R) - npa_own + 1;
                if (need > 0 && !wait_prog(sh, a.prog + split_in, need, prog_seen, spin)) return false;
#pragma unroll
                for (int c = 0; c < kCols; ++c)
                    v[c] = (p <= R && colv[c] >= 0)
                               ? __hip_atomic_load(aptr(c, row_of(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : T(0);
                return true;
            }
#pragma unroll
            for (int c = 0; c < kCols; ++c)
                v[c] = (p < npa && p <= R && colv[c] >= 0) ? *aptr(c, row_of(p)) : T(0);
            return true;
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
            ok = fetch(0, v);
            put(0, v);
            ok = ok && fetch(kChunk, chunk);
        }

        // ---- input prefetch: positions 0 .. kDin-1
        T inbuf[kDin][kCols];
        auto ensure_in = [&](int p) -> bool {
            bool good = true;
            if (p >= npa && p <= R && in_base + p >= avail_seen)
                good = wait_ge(sh, &sh.avail[g], in_base + p + 1, avail_seen, spin);
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
        for (int d = 0; d < kD; d += kSPD) dma_block(sh, g, tab_band, d, lane16, T(0));

        T Y[S][kCols], E[S][kCols];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < kCols; ++c) Y[s][c] = E[s][c] = T(0);

        // entries of step 0 (read one step ahead from here on)
        SC scn;
        wait_block();
        read_sc(sh, g, 0, i, scn);

        // Steps run in blocks of kU.  The waits are made once per block, so
        // a step is only its DPP shifts, LDS reads, rotations and the
        // broadcast of its output.
        int written = out_base;  // ring positions of the next band written (exclusive end, global)
        for (int t0 = 0; t0 < steps; t0 += kU) {
            if (t0 % kChunk == 0) {
                // positions t0+16 .. t0+31 to the ring, into the slots of
                // positions t0+16-kZ .. t0+31-kZ (all read: kZ >= 32);
                // fetch the next 16
                const int p0 = t0 + kChunk;
                if (p0 < npa) {
                    put(p0, chunk);
                    if (!fetch(p0 + kChunk, chunk)) {
                        ok = false;
                        break;
                    }
                }
            }
            // every input position the block reads ahead is in the ring
            {
                const int pl = min(t0 + kU - 1 + kDin, R);
                if (pl >= npa && in_base + pl >= avail_seen &&
                    !wait_ge(sh, &sh.avail[g], in_base + pl + 1, avail_seen, spin)) {
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
                    const int P1 = out_base + hi + npa_next;
                    if (P1 - kZ >= freed_seen && !wait_ge(sh, &sh.freed[g + 1], P1 - kZ + 1, freed_seen, spin)) {
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
            const Pair<T> *psc = &sh.scr[g][sb][lane_entry(i)];
            const Pair<T> *psc_hi = (sb == kD - kU) ? psc - kD * kLanes * S : psc;
            const unsigned m0b = lds_addr(&sh.scr[g][0][0]) + (unsigned)sb * kStepBytes;
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
                    for (int s2 = 0; s2 < S; ++s2) scn.p[s2] = q[s2 * kSlotStride];
                }
                // the reads above are issued before this step's arithmetic
                __builtin_amdgcn_sched_barrier(0);
                rotate<MASK>(Y, E, X0, sc, m, m1);
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
            // a split part's last band publishes the rows the previous block
            // stored: they are complete once at most this block's DMAs are
            // outstanding (vector memory operations complete in issue order)
            if (split_out >= 0 && pub > published) {
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kU / kSPD) : "memory");
                if (lane == 0) __hip_atomic_store(a.prog + split_out, pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                published = pub;
            }
            // lane i < kU of each row stores output position po0 + i: to the
            // next band's ring up to ring_hi, else (leaving the chain or the
            // part) to its row of A; positions outside [0, R] carry no row.
            // (The global stores only lengthen the table waits that follow
            // them: those wait for the DMAs issued before them, and loads
            // return in order.)
            {
                const int po = po0 + i;
                if (i < kU && po >= 0 && po <= R) {
                    if (po <= ring_hi) {
                        T *dst = &out_ring[(po + npa_next) % kZ][r][0];
#pragma unroll
                        for (int c = 0; c < kCols; ++c) lds_st(dst + c, outv[c]);
                    } else if (split_out >= 0) {
#pragma unroll
                        for (int c = 0; c < kCols; ++c)
                            if (colv[c] >= 0)
                                __hip_atomic_store(aptr(c, row_of(po)), outv[c], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    } else {
#pragma unroll
                        for (int c = 0; c < kCols; ++c)
                            if (colv[c] >= 0) *aptr(c, row_of(po)) = outv[c];
                    }
                }
            }
            pub = max(0, min(po0 + kU, R + 1));
            if (!ok) break;
            // publish what the block wrote and read
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (!last && lane == 0) lds_st(&sh.avail[g + 1], written);
            // positions read, never beyond this chain's part of the ring
            if (lane == 0) lds_st(&sh.freed[g], in_base + min(t0 + kU, span(R)));
        }

        if (!ok) {
            fail(sh);
            atomicOr(a.ctl + FTAMD_SYS_CTL_ERR, FTAMD_ERR_TIMEOUT);
            return false;
        }
        if (split_out >= 0) {
            // every row stored and complete
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            if (lane == 0) __hip_atomic_store(a.prog + split_out, R + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#ifdef FTAMD_TRACE
        if (a.trace && lane == 0) {
            // accumulated over the bands the wave runs
            unsigned long long *rec = a.trace + ((long long)blockIdx.x * kWaves + g) * 8;
            if (rec[0] == 0) rec[0] = tr_start;
            rec[1] = wall_clock64();
            rec[2] += __builtin_amdgcn_s_memtime() - tr_c0;
            rec[3] += tr_in();
            rec[4] += tr_out();
            rec[5] += steps;
            rec[6] += R;
            rec[7] += tr_dma();
        }
#endif
        return true;
    }

    // Ring positions one band's stream takes: its R + 1 positions rounded up
    // to whole rings, so that every stream starts at ring slot 0.
    __device__ static int span(int R) { return (R + kZ) / kZ * kZ; }

    // XCD (accelerator complex die, with its own L2) this workgroup runs on.
    __device__ static int xcd_id(int nxcd) {
        const int id = (int)__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));  // HW_REG_XCC_ID[3:0]
        return nxcd > 1 ? id % nxcd : 0;
    }

    // A workgroup takes the slots of its XCD in order (tickets), and streams
    // a slot's chain parts (items) one after another through its waves: item
    // k runs its part's bands on waves wave0 .. wave0+nb-1, and every wave
    // takes the items in list order, so a wave starts its band of the next
    // part as soon as it has finished its band of the previous one.  A part
    // that waits on another part waits on an earlier slot of the same XCD,
    // which a running workgroup holds or has finished: the launch completes
    // however many workgroups the hardware runs at once.
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
        const int lane = threadIdx.x & 63;
        const int g = uniform(threadIdx.x >> 6);
        const int x = uniform(xcd_id(a.nxcd));
        const int s0 = uniform(a.xcd_slot[x]), ns = uniform(a.xcd_slot[x + 1]) - s0;
        unsigned *ticket = (unsigned *)(a.ctl + FTAMD_SYS_CTL_TICKET) + x;
        int in_base = 0, out_base = 0;  // where the current part's streams start in rings g and g+1
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) sh.slot = (int)atomicAdd(ticket, 1u);
            __syncthreads();
            const int t = uniform(lds_ld(&sh.slot));
            if (t >= ns || uniform(lds_ld(&sh.abort_flag))) break;
            const int slot = s0 + t;
            const int it0 = uniform(a.slot_items[slot]), it1 = uniform(a.slot_items[slot + 1]);
            for (int it = it0; it < it1; ++it) {
                const FtamdSysChain &ch = a.chains[uniform(a.items[it].chain)];
                const int nb = uniform(ch.nb), b0 = uniform(ch.band0);
                const int w = g - uniform(a.items[it].wave0);  // this wave's band of the part, if any
                if (w >= 0 && w < nb) {
                    switch (uniform(ch.prio)) {
                        case 3: __builtin_amdgcn_s_setprio(3); break;
                        case 2: __builtin_amdgcn_s_setprio(2); break;
                        case 1: __builtin_amdgcn_s_setprio(1); break;
                        default: __builtin_amdgcn_s_setprio(0); break;
                    }
                    // a band needs masks if some column of the chain does
                    // not apply some of its sequences; with ROWMASK one mask
                    // per lane row serves both columns (sph and disk: the two
                    // columns of a row are the cosine and sine columns of one
                    // order, so they apply the same sequences)
                    const int gbase = uniform(a.bands[b0 + w].gbase);
                    const int glast = gbase + uniform(a.bands[b0 + w].nseq) - 1;
                    bool masked = false;
                    for (int k = 0; k < FTAMD_SYS_COLS; ++k)
                        masked |= ch.col[k] >= 0 && (ch.glo[k] > gbase || ch.ghi[k] < glast);
                    // bands that hand rows to or take rows from another
                    // part of the chain run their own instantiation, so the
                    // others carry none of that code
                    const bool link = (w == 0 && uniform(ch.split_in) >= 0) || (w == nb - 1 && uniform(ch.split_out) >= 0);
                    constexpr int MK = ROWMASK ? 1 : 2;
                    const bool ok = link ? (!masked ? run_band<0, true>(a, sh, ch, w, g, lane, in_base, out_base)
                                                    : run_band<MK, true>(a, sh, ch, w, g, lane, in_base, out_base))
                                         : (!masked ? run_band<0, false>(a, sh, ch, w, g, lane, in_base, out_base)
                                                    : run_band<MK, false>(a, sh, ch, w, g, lane, in_base, out_base));
                    if (!ok) break;
                    in_base += span(uniform(a.bands[b0 + w].R));
                }
                if (w + 1 >= 0 && w + 1 < nb) out_base += span(uniform(a.bands[b0 + w + 1].R));
            }
            __builtin_amdgcn_s_setprio(0);
            __syncthreads();
            if (threadIdx.x == 0 && !lds_ld(&sh.abort_flag)) atomicAdd((unsigned *)(a.ctl + FTAMD_SYS_CTL_COMPLETED), 1u);
        }
        // The last workgroup out checks that every slot ran and resets the
        // tickets, counters and progress words for the next launch (or graph
        // replay).
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned *done = (unsigned *)(a.ctl + FTAMD_SYS_CTL_DONE);
            unsigned *completed = (unsigned *)(a.ctl + FTAMD_SYS_CTL_COMPLETED);
            const unsigned d = __hip_atomic_fetch_add(done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (d == gridDim.x - 1) {
                if (__hip_atomic_load(completed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != (unsigned)a.nslots)
                    atomicOr(a.ctl + FTAMD_SYS_CTL_ERR, FTAMD_ERR_INCOMPLETE);
                for (int k = 0; k < a.nprog; ++k)
                    __hip_atomic_store(a.prog + k, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int k = 0; k < a.nxcd; ++k)
                    __hip_atomic_store((unsigned *)(a.ctl + FTAMD_SYS_CTL_TICKET) + k, 0u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(completed, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(done, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
};

template <typename T, int S, bool ROWMASK>
__global__ void __launch_bounds__(kWaves * 64) sys_kernel(FtamdSysLaunch a) {
    using E = Engine<T, S, ROWMASK>;
    __shared__ typename E::Shared sh;
    E::run(a, sh);
}

}  // namespace ftsys
