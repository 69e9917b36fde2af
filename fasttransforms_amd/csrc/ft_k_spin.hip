// ft_k_spin.hip -- first stage of the spin-weighted sphere's rotations
// (rotations_default.c:87-117, the loops over s2, c2): for a complex column
// of order m and spin s (am = |m|, as = |s|, d = |am - as|), the sequences
// j = min(am, as)-1 .. 0 of rotations of adjacent rows (k, k+1),
// k = n-2-d-j .. 0, with s2(k, j, d), c2(k, j, d) -- the rotation where
// m*s >= 0 and its transpose otherwise (hi2lo); lo2hi runs the transposes in
// the opposite order.  The real and imaginary parts take the same real
// rotation.
//
// Two lanes own a complex column, one its real and one its imaginary part,
// and sweep its sequences one after another down (hi2lo) or up (lo2hi) the
// column: a step rotates the fresh row against the carried one and emits a
// finished row.  The launch has few waves (one per 32 parts of columns), so
// a step's cost is its instructions issued in order plus the carried row's
// FP64 latency; the sweep keeps a step to one rotation (no per-step range
// bookkeeping: every pair of a sequence is in its range) and loads the next
// block of rows and (s, c) while the current block is computed.  (A
// systolic sweep of up to 8 sequences per lane, complex columns per lane,
// measured slower at every spin.)
//
// The transpose of the rotation (s, c) is the rotation (-s, c), bit for bit
// in the fused forms: fma(c, x, (-s)*y) = fma(c, x, -(s*y)) and
// fma(c, y, -((-s)*x)) = fma(c, y, s*x), since negation is exact.  So every
// rotation here is X' = fma(c, x, t*y), Y' = fma(c, y, -(t*x)) with t = +-s.
#include <hip/hip_runtime.h>

#include "ft_hip_abi.h"

namespace ftspin {

__device__ __forceinline__ double rot_x(double C, double S, double x, double y) { return __builtin_fma(C, x, S * y); }
__device__ __forceinline__ double rot_y(double C, double S, double x, double y) { return __builtin_fma(C, y, -(S * x)); }

// One sequence over one part of a column: the pairs are r = R-1 .. 0 (hi2lo) or 0 .. R-1
// (lo2hi), every one of them in the sequence's range, so a step is one
// rotation of the fresh row against the carried one, emitting a finished
// row.  Full blocks of KB steps run without a branch, the next block's rows
// and (s, c) loaded at clamped indices while this one is computed; the
// remaining R mod KB steps and the last row follow.
template <int DIR>
__device__ void sweep1(double *col, long long rs, const double2 *__restrict__ tq, int R, double sg) {
    using V = double;
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
    auto load = [&](int p0, V (&x)[KB], double2 (&sc)[KB]) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            x[k] = col[(long long)frow(p0 + k) * rs];
            sc[k] = tq[pair(p0 + k)];
        }
    };
    // hi2lo: x row r, the carry row r+1, emits row r+1; lo2hi: the carry row r, x row r+1, emits row r
    auto step = [&](V x, double2 sc, V &cy) {
        const double S = sg * sc.x, C = sc.y;
        V em;
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
    V cy = col[(long long)(DIR == 0 ? R : 0) * rs];
    const int nb = R / KB;
    V xb[KB], xn[KB];
    double2 sb[KB], sn[KB];
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

// One sequence a sweep (sweep1), two lanes a complex column (lane 2i + h
// owns part h of column i): R = n-1-d-j0, (s, c) of sequence j0.
__global__ void __launch_bounds__(64) stage1_split_kernel(const double2 *__restrict__ tab, int n, int s, double2 *A,
                                                          int c0, int count, long long col_stride, int row_stride,
                                                          int use_fixed, int fixed_m, int dir) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = l >> 1, h = l & 1;
    if (i >= count) return;
    const int c = c0 + i;
    const int m = use_fixed ? fixed_m : ((c & 1) ? -((c + 1) >> 1) : (c >> 1));
    const int as = s < 0 ? -s : s, am = m < 0 ? -m : m;
    const int nseq = am < as ? am : as;
    const int d = am > as ? am - as : as - am;
    if (nseq == 0 || n - 2 - d < 0) return;
    const int nrot = nseq < n - 1 - d ? nseq : n - 1 - d;
    const bool rot = (m >= 0) == (s >= 0);
    const double sg = (dir == 0) == rot ? 1.0 : -1.0;
    double *col = (double *)(A + col_stride * i) + h;
    const long long rs = 2LL * row_stride;
    const double2 *t = tab + (long long)d * as * n;
    if (dir == 0) {
        for (int j0 = nrot - 1; j0 >= 0; --j0) sweep1<0>(col, rs, t + (long long)j0 * n, n - 1 - d - j0, sg);
    } else {
        for (int j0 = 0; j0 < nrot; ++j0) sweep1<1>(col, rs, t + (long long)j0 * n, n - 1 - d - j0, sg);
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
    if (count > (1 << 29)) return FTAMD_EINVAL;  // two lanes a column, lane indices in an int
    const int block = 64;
    const dim3 grid((2 * count + block - 1) / block);
    const hipStream_t st = (hipStream_t)stream;
    const double2 *tab = (const double2 *)tab2;
    double2 *a = (double2 *)A;
    hipLaunchKernelGGL(ftspin::stage1_split_kernel, grid, dim3(block), 0, st, tab, n, s, a, c0, count, col_stride,
                       row_stride, use_fixed, fixed_m, dir);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}
