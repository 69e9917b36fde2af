// ft_trmm.hip -- the strided upper-triangular products of the harmonic
// transforms (drivers.c:833-864: four cblas_dtrmm calls per transform, one
// per class of columns j = c mod 4) as one launch on the FP64 matrix cores.
//
// Y[:, j] = op(U_c) X[:, j] for every column j of every class c, out of
// place.  A workgroup (4 waves) computes a 64 x 32 tile of Y: wave w the rows
// 16w .. 16w+15 as two 16 x 16 blocks, accumulated by
// v_mfma_f64_16x16x4_f64 over k in stages of 16 staged through LDS (op(U)
// 64 x 16, X 16 x 32), the next stage's global loads issued before the
// current stage's products.  Lane layout of the 16x16x4 FP64 MFMA: A
// operand lane i = A[i % 16][i / 16], B operand lane i = B[i / 16][i % 16],
// result register j of lane i = D[i / 16 + 4 j][i % 16].  Only the
// triangle's k range is visited: k >= first row of the tile for U, k <= last
// row for U^T.  The products sum in a different order from the reference's
// BLAS, so results agree to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include "ft_hip_abi.h"

namespace fttrmm {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TR = FTAMD_TRMM_ROWS;  // output rows per workgroup
constexpr int TC = FTAMD_TRMM_COLS;  // output columns per workgroup
constexpr int TK = 16;               // k per stage
constexpr int PAD = 2;               // LDS row padding (doubles)
constexpr int NU = TR * TK / 256;    // op(U) elements each thread stages
constexpr int NX = TK * TC / 256;    // X elements each thread stages

__global__ void __launch_bounds__(256) trmm_kernel(FtamdTrmmLaunch a) {
    __shared__ double sU[TK][TR + PAD];  // sU[k][r] = op(U)[r0 + r][k0 + k]
    __shared__ double sX[TK][TC + PAD];  // sX[k][q] = X[k0 + k][column q of the tile]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = a.n;
    // class and column tile of this workgroup
    int c = 0, t = blockIdx.x;
    while (c + 1 < a.ncls && t >= a.ctiles[c]) t -= a.ctiles[c++];
    const double *U = a.U[c];
    if (U == nullptr || t >= a.ctiles[c]) return;
    const int count = (a.M - c + a.cstride - 1) / a.cstride;  // columns of class c
    const int q0 = t * TC;                                    // first class-local column of the tile
    const int r0 = blockIdx.y * TR;
    const bool trans = a.trans != 0;
    const int kbeg = trans ? 0 : r0;
    const int kend = trans ? min(n, r0 + TR) : n;

    // staging: element e of op(U)'s 64 x 16 block and of X's 16 x 32 block
    auto load_u = [&](int k0, double (&v)[NU]) {
#pragma unroll
        for (int s = 0; s < NU; ++s) {
            const int e = tid + 256 * s;
            // U: consecutive threads take consecutive rows r (a column of U);
            // U^T: consecutive k (a column of U, read as a row of U^T)
            const int r = trans ? e / TK : e % TR, k = trans ? e % TK : e / TR;
            const int gr = r0 + r, gk = k0 + k;
            v[s] = (gr < n && gk < n) ? (trans ? U[gk + (long long)gr * n] : U[gr + (long long)gk * n]) : 0.0;
        }
    };
    auto load_x = [&](int k0, double (&v)[NX]) {
#pragma unroll
        for (int s = 0; s < NX; ++s) {
            const int e = tid + 256 * s;
            const int k = e % TK, q = e / TK;
            const int gk = k0 + k, gq = q0 + q;
            v[s] = (gk < n && gq < count) ? a.X[gk + (long long)(c + gq * a.cstride) * n] : 0.0;
        }
    };
    auto store = [&](const double (&u)[NU], const double (&x)[NX]) {
#pragma unroll
        for (int s = 0; s < NU; ++s) {
            const int e = tid + 256 * s;
            const int r = trans ? e / TK : e % TR, k = trans ? e % TK : e / TR;
            sU[k][r] = u[s];
        }
#pragma unroll
        for (int s = 0; s < NX; ++s) {
            const int e = tid + 256 * s;
            sX[e % TK][e / TK] = x[s];
        }
    };

    v4d acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    double u[NU], x[NX];
    const int k_first = kbeg / TK * TK;
    if (k_first < kend) {
        load_u(k_first, u);
        load_x(k_first, x);
    }
    for (int k0 = k_first; k0 < kend; k0 += TK) {
        store(u, x);
        __syncthreads();
        // the next stage's loads are in flight during this stage's products
        if (k0 + TK < kend) {
            load_u(k0 + TK, u);
            load_x(k0 + TK, x);
        }
#pragma unroll
        for (int ks = 0; ks < TK; ks += 4) {
            const int kk = ks + (lane >> 4), ii = lane & 15;
            const double av = sU[kk][16 * w + ii];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sX[kk][ii], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sX[kk][16 + ii], acc1, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int r = r0 + 16 * w + (lane >> 4) + 4 * j;
        if (r >= n) continue;
        const int qa = q0 + (lane & 15), qb = qa + 16;
        if (qa < count) a.Y[r + (long long)(c + qa * a.cstride) * n] = acc0[j];
        if (qb < count) a.Y[r + (long long)(c + qb * a.cstride) * n] = acc1[j];
    }
}

}  // namespace fttrmm

extern "C" int ftamd_hip_trmm_launch(const FtamdTrmmLaunch *a, void *stream) {
    if (a->n < 1 || a->M < 0 || a->cstride < 1 || a->ncls < 1 || a->ncls > 4 || !a->X || !a->Y) return FTAMD_EINVAL;
    int tiles = 0;
    for (int c = 0; c < a->ncls; ++c) {
        const int count = a->M > c ? (a->M - c + a->cstride - 1) / a->cstride : 0;
        if (a->ctiles[c] != (count + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS) return FTAMD_EINVAL;
        tiles += a->ctiles[c];
    }
    // element offsets of Y and X fit the kernel's 64-bit index math; the
    // grid's second dimension covers the rows
    const int rows = (a->n + FTAMD_TRMM_ROWS - 1) / FTAMD_TRMM_ROWS;
    if (tiles == 0) return FTAMD_OK;
    hipLaunchKernelGGL(fttrmm::trmm_kernel, dim3(tiles, rows), dim3(256), 0, (hipStream_t)stream, *a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : (e == hipErrorOutOfMemory ? FTAMD_ENOMEM : FTAMD_EDEVICE);
}
