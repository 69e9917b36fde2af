// ft_k_spin.hip -- first stage of the spin-weighted sphere's rotations
// (rotations_default.c:87-117, the loops over s2, c2): for a complex column
// of order m and spin s (am = |m|, as = |s|, d = |am - as|), the sequences
// j = min(am, as)-1 .. 0 of rotations of adjacent rows (k, k+1),
// k = n-2-d-j .. 0, with s2(k, j, d), c2(k, j, d) -- the rotation where
// m*s >= 0 and its transpose otherwise (hi2lo); lo2hi runs the transposes in
// the opposite order.  The real and imaginary parts take the same real
// rotation.
//
// A lane owns a complex column.  It applies up to Q sequences in one sweep
// down (hi2lo) or up (lo2hi) the column, as a systolic array held in its
// registers: sequence q trails sequence q-1 by two rows, takes the row q-1
// emitted at the previous step and emits the row above (below) the one it
// takes, so the Q rotations of a step are independent of each other and only
// each sequence's own carried row chains from step to step.  Rows outside a
// sequence's range pass through it unchanged.  Fresh rows and the (s, c) of
// every sequence are loaded, and the last sequence's rows stored, in blocks of
// steps, the next block's loads issued before the current block's arithmetic.
// Columns with more than Q sequences take several sweeps.  A lone sequence
// (Q = 1) takes sweep1, which has none of the systolic bookkeeping.
//
// The transpose of the rotation (s, c) is the rotation (-s, c), bit for bit
// in the fused forms: fma(c, x, (-s)*y) = fma(c, x, -(s*y)) and
// fma(c, y, -((-s)*x)) = fma(c, y, s*x), since negation is exact.  So every
// rotation here is X' = fma(c, x, t*y), Y' = fma(c, y, -(t*x)) with t = +-s,
// on the real and the imaginary parts alike.
#include <hip/hip_runtime.h>

#include "ft_hip_abi.h"

namespace ftspin {

__device__ __forceinline__ double2 rot_x(double C, double S, double2 x, double2 y) {
    return make_double2(__builtin_fma(C, x.x, S * y.x), __builtin_fma(C, x.y, S * y.y));
}
__device__ __forceinline__ double2 rot_y(double C, double S, double2 x, double2 y) {
    return make_double2(__builtin_fma(C, y.x, -(S * x.x)), __builtin_fma(C, y.y, -(S * x.y)));
}

// One sweep of qa <= Q sequences over rows 0 .. R of the column: sequence q
// is j = j0 - q (hi2lo) or j0 + q (lo2hi), rotating the pairs (k, k+1),
// k <= n-2-d-j.  Rows and (s, c) are loaded a block of KB steps ahead, at
// indices clamped into range (the values a step does not use are discarded),
// so no load waits on a branch or on the step before it.
template <int Q, int DIR>
__device__ void sweep(double2 *col, long long rs, const double2 *__restrict__ t, int n, int d, int j0, int qa,
                      double sg) {
    constexpr int KB = Q >= 8 ? 2 : (Q >= 4 ? 4 : 8);
    int top[Q], topc[Q];
    const double2 *tq[Q];
    const int jmin = DIR == 0 ? j0 - qa + 1 : j0;
    const int R = n - 1 - d - jmin;  // the longest sequence's last pair is (R-1, R)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = DIR == 0 ? j0 - q : j0 + q;
        top[q] = q < qa ? n - 2 - d - j : -1;
        topc[q] = top[q] > 0 ? top[q] : 0;
        tq[q] = t + (long long)(q < qa ? j : j0) * n;  // a sequence past qa reads (and ignores) j0's
    }
    double2 cy[Q], em[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) cy[q] = em[q] = make_double2(0.0, 0.0);
    // row of stream position p (clamped into 0 .. R)
    auto row_of = [&](int p) {
        const int pc = p < R ? p : R;
        return DIR == 0 ? R - pc : pc;
    };
    // the pair sequence q rotates at step tt (valid or not)
    auto pair_of = [&](int tt, int q) {
        const int p = tt - 2 * q;
        return DIR == 0 ? R - p : p - 1;
    };
    auto load = [&](int t0, double2 (&x)[KB], double2 (&sc)[KB][Q]) {
#pragma unroll
        for (int i = 0; i < KB; ++i) x[i] = col[row_of(t0 + i) * rs];
#pragma unroll
        for (int i = 0; i < KB; ++i)
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                int r = pair_of(t0 + i, q);
                r = r < 0 ? 0 : (r > topc[q] ? topc[q] : r);
                sc[i][q] = tq[q][r];
            }
    };
    const int steps = R + 2 + 2 * (Q - 1);
    double2 xb[KB], xn[KB], scb[KB][Q], scn[KB][Q];
    load(0, xb, scb);
    for (int t0 = 0; t0 < steps; t0 += KB) {
        load(t0 + KB, xn, scn);
        double2 ob[KB];
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int tt = t0 + i;
#pragma unroll
            for (int q = Q - 1; q >= 0; --q) {
                const int p = tt - 2 * q;
                const double2 x = q == 0 ? xb[i] : em[q - 1];  // what q-1 emitted at the previous step
                // hi2lo: the pair (r, r+1), x row r, the carry row r+1; lo2hi: the pair (r, r+1), the
                // carry row r, x row r+1.  The rotation is computed at every step and kept by selects
                // (the lanes' ranges differ; branches here would make the compiler wait on the
                // prefetched loads): position 0 takes x as the carry, 1 .. R rotate where the pair
                // is in the sequence's range and pass x on otherwise, R + 1 emits the carry.
                const bool mid = p >= 1 && p <= R;
                const bool act = mid && pair_of(tt, q) <= top[q];
                const double S = sg * scb[i][q].x, C = scb[i][q].y;
                const double2 e1 = DIR == 0 ? rot_y(C, S, x, cy[q]) : rot_x(C, S, cy[q], x);
                const double2 c1 = DIR == 0 ? rot_x(C, S, x, cy[q]) : rot_y(C, S, cy[q], x);
                const double2 en = act ? e1 : ((mid || p == R + 1) ? cy[q] : em[q]);
                const double2 cn = act ? c1 : ((mid || p == 0) ? x : cy[q]);
                em[q] = en;
                cy[q] = cn;
            }
            ob[i] = em[Q - 1];
        }
        // the last sequence emitted, at its position pl = tt - 2(Q-1) in 1 .. R+1,
        // row R - pl + 1 (hi2lo) or pl - 1 (lo2hi)
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int pl = t0 + i - 2 * (Q - 1);
            if (pl >= 1 && pl <= R + 1) col[(DIR == 0 ? R - pl + 1 : pl - 1) * rs] = ob[i];
        }
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            xb[i] = xn[i];
#pragma unroll
            for (int q = 0; q < Q; ++q) scb[i][q] = scn[i][q];
        }
    }
}

// One sequence (|s| = 1, or the last of a column): the sweep without the
// systolic bookkeeping.  The pairs are r = R-1 .. 0 (hi2lo) or 0 .. R-1
// (lo2hi), every one of them in the sequence's range, so a step is one
// rotation of the fresh row against the carried one, emitting a finished
// row.  Full blocks of KB steps run without a branch, the next block's rows
// and (s, c) loaded at clamped indices while this one is computed; the
// remaining R mod KB steps and the last row follow.
template <int DIR>
__device__ void sweep1(double2 *col, long long rs, const double2 *__restrict__ tq, int R, double sg) {
    constexpr int KB = 12;
    // step p = 1 .. R: the fresh row and the pair, clamped into range past R
    auto frow = [&](int p) {
        const int pc = p > R ? R : p;
        return DIR == 0 ? R - pc : pc;
    };
    auto pair = [&](int p) {
        const int pc = p > R ? R : p;
        return DIR == 0 ? R - pc : pc - 1;
    };
    auto load = [&](int p0, double2 (&x)[KB], double2 (&sc)[KB]) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            x[k] = col[(long long)frow(p0 + k) * rs];
            sc[k] = tq[pair(p0 + k)];
        }
    };
    // hi2lo: x row r, the carry row r+1, emits row r+1; lo2hi: the carry row r, x row r+1, emits row r
    auto step = [&](double2 x, double2 sc, double2 &cy) {
        const double S = sg * sc.x, C = sc.y;
        double2 em;
        if (DIR == 0) {
            em = rot_y(C, S, x, cy);
            cy = rot_x(C, S, x, cy);
        } else {
            em = rot_x(C, S, cy, x);
            cy = rot_y(C, S, cy, x);
        }
        return em;
    };
    auto out_row = [&](int p) { return (long long)(DIR == 0 ? R - p + 1 : p - 1) * rs; };
    double2 cy = col[(long long)(DIR == 0 ? R : 0) * rs];
    const int nb = R / KB;
    double2 xb[KB], sb[KB], xn[KB], sn[KB];
    load(1, xb, sb);
    for (int b = 0; b < nb; ++b) {
        const int p0 = 1 + b * KB;
        load(p0 + KB, xn, sn);
#pragma unroll
        for (int k = 0; k < KB; ++k) col[out_row(p0 + k)] = step(xb[k], sb[k], cy);
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            xb[k] = xn[k];
            sb[k] = sn[k];
        }
    }
    const int p0 = 1 + nb * KB, rem = R - nb * KB;
#pragma unroll
    for (int k = 0; k < KB; ++k)
        if (k < rem) col[out_row(p0 + k)] = step(xb[k], sb[k], cy);
    col[(long long)(DIR == 0 ? 0 : R) * rs] = cy;
}

template <int Q>
__global__ void __launch_bounds__(64) stage1_kernel(const double2 *__restrict__ tab, int n, int s, double2 *A, int c0,
                                                    int count, long long col_stride, int row_stride, int use_fixed,
                                                    int fixed_m, int dir) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int c = c0 + i;
    // the executors' column order: column 0 is order 0, column 2m-1 order -m, column 2m order m
    const int m = use_fixed ? fixed_m : ((c & 1) ? -((c + 1) >> 1) : (c >> 1));
    const int as = s < 0 ? -s : s, am = m < 0 ? -m : m;
    const int nseq = am < as ? am : as;
    const int d = am > as ? am - as : as - am;
    if (nseq == 0 || n - 2 - d < 0) return;
    // sequences with pairs: j <= n-2-d
    const int nrot = nseq < n - 1 - d ? nseq : n - 1 - d;
    // rotation (+s) where hi2lo applies the rotation (m*s >= 0), transpose (-s) otherwise
    const bool rot = (m >= 0) == (s >= 0);
    const double sg = (dir == 0) == rot ? 1.0 : -1.0;
    double2 *col = A + col_stride * i;
    const double2 *t = tab + (long long)d * as * n;
    if (Q == 1) {  // one sequence a sweep: R = n-1-d-j0, (s, c) of sequence j0
        if (dir == 0) {
            for (int j0 = nrot - 1; j0 >= 0; --j0) sweep1<0>(col, row_stride, t + (long long)j0 * n, n - 1 - d - j0, sg);
        } else {
            for (int j0 = 0; j0 < nrot; ++j0) sweep1<1>(col, row_stride, t + (long long)j0 * n, n - 1 - d - j0, sg);
        }
    } else if (dir == 0) {
        // j = nseq-1 .. 0; those above nrot-1 have no pairs
        for (int j0 = nrot - 1; j0 >= 0; j0 -= Q) sweep<Q, 0>(col, row_stride, t, n, d, j0, j0 + 1 < Q ? j0 + 1 : Q, sg);
    } else {
        for (int j0 = 0; j0 < nrot; j0 += Q) sweep<Q, 1>(col, row_stride, t, n, d, j0, nrot - j0 < Q ? nrot - j0 : Q, sg);
    }
}

}  // namespace ftspin

extern "C" int ftamd_hip_spin_stage1(const double *tab2, int n, int s, double *A, int c0, int count,
                                     long long col_stride, int row_stride, int use_fixed, int fixed_m, int dir,
                                     void *stream) {
    if (!A || n < 1 || c0 < 0 || count < 0 || row_stride < 1 || col_stride < 0 || (dir != 0 && dir != 1))
        return FTAMD_EINVAL;
    if (count == 0 || s == 0) return FTAMD_OK;
    if (!tab2) return FTAMD_EINVAL;
    const int block = 64, as = s < 0 ? -s : s;
    const dim3 grid((count + block - 1) / block);
    const hipStream_t st = (hipStream_t)stream;
    const double2 *tab = (const double2 *)tab2;
    double2 *a = (double2 *)A;
    // one sequence a sweep (sweep1), except lo2hi with |s| = 3, 4, where the four-sequence systolic
    // sweep is faster (measured transforms at n = 1024, s = 4: lo2hi 0.72 ms with it against 0.78
    // ms; hi2lo 0.66 ms one sequence at a time against 0.69; s = 9 hi2lo: 1.1 ms against 2.1 ms in
    // sweeps of eight sequences)
    if (dir == 1 && as >= 3 && as <= 4)
        hipLaunchKernelGGL(ftspin::stage1_kernel<4>, grid, dim3(block), 0, st, tab, n, s, a, c0, count, col_stride,
                           row_stride, use_fixed, fixed_m, dir);
    else
        hipLaunchKernelGGL(ftspin::stage1_kernel<1>, grid, dim3(block), 0, st, tab, n, s, a, c0, count, col_stride,
                           row_stride, use_fixed, fixed_m, dir);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}
