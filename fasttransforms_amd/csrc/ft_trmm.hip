// ft_trmm.hip -- the upper-triangular products of the harmonic transforms
// (drivers.c:833-1021: cblas_dtrmm from the left on the column classes
// j = c mod 4, or on every column, and from the right on every row) as one
// launch each on the FP64 matrix cores.
//
// Left: Y[:, j] = op(U_c) X[:, j] for every column j of every class c.
// Right: Y = X op(U) with U's leading M x M block.  Out of place.  A
// workgroup (4 waves) computes a 64 x 32 tile of Y: wave w the rows 16w ..
// 16w+15 as two 16 x 16 blocks, accumulated by v_mfma_f64_16x16x4_f64 over k
// in stages of 16 staged through LDS (the A operand 64 x 16, the B operand
// 16 x 32), the next stage's global loads issued before the current stage's
// products.  Lane layout of the 16x16x4 FP64 MFMA: A operand lane i =
// A[i % 16][i / 16], B operand lane i = B[i / 16][i % 16], result register
// j of lane i = D[i / 16 + 4 j][i % 16].  Only the triangle's k range is
// visited.  The products sum in another order than the reference's BLAS, so
// results agree to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include "ft_hip_abi.h"

namespace fttrmm {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TR = FTAMD_TRMM_ROWS;  // output rows per workgroup
constexpr int TC = FTAMD_TRMM_COLS;  // output columns per workgroup
constexpr int TK = 16;               // k per stage
constexpr int PAD = 2;               // LDS row padding (doubles)
constexpr int NA = TR * TK / 256;    // A-operand elements each thread stages
constexpr int NB = TK * TC / 256;    // B-operand elements each thread stages

template <bool RIGHT>
__global__ void __launch_bounds__(256) trmm_kernel(FtamdTrmmLaunch a) {
    __shared__ double sA[TK][TR + PAD];  // sA[k][r] = A operand (r0 + r, k0 + k)
    __shared__ double sB[TK][TC + PAD];  // sB[k][q] = B operand (k0 + k, column q of the tile)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = a.n;
    // class and column tile of this workgroup
    int c = 0, t = blockIdx.x;
    while (c + 1 < a.ncls && t >= a.ctiles[c]) t -= a.ctiles[c++];
    const double *U = a.U[c];
    if (U == nullptr || t >= a.ctiles[c]) return;
    const int count = (a.M - c + a.cstride - 1) / a.cstride;  // output columns of the class
    const int q0 = t * TC;                                    // first class-local column of the tile
    const int r0 = blockIdx.y * TR;
    const bool trans = a.trans != 0;
    // k ranges of the triangle: left, rows of op(U) from r0 (U) or up to the
    // tile's last row (U^T); right, columns of op(U) up to the tile's last
    // column (U) or from q0 (U^T)
    const int K = RIGHT ? a.M : n;
    int kbeg, kend;
    if (!RIGHT) {
        kbeg = trans ? 0 : r0;
        kend = trans ? min(n, r0 + TR) : n;
    } else {
        kbeg = trans ? q0 : 0;
        kend = trans ? K : min(K, q0 + TC);
    }
    const long long xcol = a.cstride;  // X column of class-local column q: c + q * cstride

    // element e of the A operand's 64 x 16 block: consecutive threads along
    // the contiguous direction of the source
    auto a_rk = [&](int e, int &r, int &k) {
        const bool along_k = !RIGHT && trans;  // U^T: a column of U is a row of the operand
        r = along_k ? e / TK : e % TR;
        k = along_k ? e % TK : e / TR;
    };
    auto load_a = [&](int k0, double (&v)[NA]) {
#pragma unroll
        for (int s = 0; s < NA; ++s) {
            int r, k;
            a_rk(tid + 256 * s, r, k);
            const int gr = r0 + r, gk = k0 + k;
            double x = 0.0;
            if (gr < n && gk < K) {
                if (RIGHT) x = a.X[gr + (long long)gk * n];                  // X[gr][gk]
                else x = trans ? U[gk + (long long)gr * n] : U[gr + (long long)gk * n];  // op(U)[gr][gk]
            }
            v[s] = x;
        }
    };
    // element e of the B operand's 16 x 32 block
    auto b_kq = [&](int e, int &k, int &q) {
        const bool along_q = RIGHT && trans;  // U^T from the right: op(U)[k][q] = U[q][k], contiguous in q
        k = along_q ? e / TC : e % TK;
        q = along_q ? e % TC : e / TK;
    };
    auto load_b = [&](int k0, double (&v)[NB]) {
#pragma unroll
        for (int s = 0; s < NB; ++s) {
            int k, q;
            b_kq(tid + 256 * s, k, q);
            const int gk = k0 + k, gq = q0 + q;
            double x = 0.0;
            if (gk < K && gq < count) {
                if (RIGHT) x = trans ? U[gq + (long long)gk * n] : U[gk + (long long)gq * n];  // op(U)[gk][gq]
                else x = a.X[gk + (c + gq * xcol) * n];                                        // X[gk][column]
            }
            v[s] = x;
        }
    };
    auto store = [&](const double (&va)[NA], const double (&vb)[NB]) {
#pragma unroll
        for (int s = 0; s < NA; ++s) {
            int r, k;
            a_rk(tid + 256 * s, r, k);
            sA[k][r] = va[s];
        }
#pragma unroll
        for (int s = 0; s < NB; ++s) {
            int k, q;
            b_kq(tid + 256 * s, k, q);
            sB[k][q] = vb[s];
        }
    };

    v4d acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    double va[NA], vb[NB];
    const int k_first = kbeg / TK * TK;
    if (k_first < kend) {
        load_a(k_first, va);
        load_b(k_first, vb);
    }
    for (int k0 = k_first; k0 < kend; k0 += TK) {
        store(va, vb);
        __syncthreads();
        // the next stage's loads are in flight during this stage's products
        if (k0 + TK < kend) {
            load_a(k0 + TK, va);
            load_b(k0 + TK, vb);
        }
#pragma unroll
        for (int ks = 0; ks < TK; ks += 4) {
            const int kk = ks + (lane >> 4), ii = lane & 15;
            const double av = sA[kk][16 * w + ii];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sB[kk][ii], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sB[kk][16 + ii], acc1, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int r = r0 + 16 * w + (lane >> 4) + 4 * j;
        if (r >= n) continue;
        const int qa = q0 + (lane & 15), qb = qa + 16;
        if (qa < count) a.Y[r + (c + qa * xcol) * n] = acc0[j];
        if (qb < count) a.Y[r + (c + qb * xcol) * n] = acc1[j];
    }
}

__global__ void cheb_scale_kernel(double *A, int n, int M, double s1, double s2) {
    const long long total = (long long)n * M;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        double v = A[e];
        if (j == 0) v *= s1;
        if (i == 0) v *= s1;
        A[e] = v * s2;
    }
}

// The annulus's semi-classical <-> classical conversion on the columns j = 1,
// 2 mod 4 of the n x M matrix A (drivers.c:1073-1121 with banded_source.c:
// 381-422 for the bidiagonal factor): x <- R^{-1} x, R^{-T} x (solves) or
// R x, R^T x (products), R upper bidiagonal with diagonal d and
// superdiagonal u, products and differences rounded one at a time in the
// reference's order.  Columns q = 0 .. count-1 of the two classes are
// column 4 (q / 2) + 1 + q % 2.
__device__ inline double *bidiag_column(double *A, int n, int q) { return A + (4LL * (q >> 1) + 1 + (q & 1)) * n; }

// The solves are a recurrence: one thread per column walks it (R^{-T} from row
// 0 up, R^{-1} from row n-1 down).  The column streams through registers in
// chunks of CH rows (one 128-byte line) with the next chunk's loads in flight
// during this one's arithmetic, so the division chain is the critical path.
template <bool TRANS>
__global__ void __launch_bounds__(64) bidiag_solve_kernel(double *__restrict__ A, int n, int count,
                                                          const double *__restrict__ d, const double *__restrict__ u) {
#pragma clang fp contract(off)
    constexpr int CH = 16;
    constexpr bool ASC = TRANS;
    const int q = blockIdx.x * 64 + threadIdx.x;
    if (q >= count) return;
    double *x = bidiag_column(A, n, q);
    // row i's superdiagonal entry is u[i - 1] (R^T) or u[i] (R); the walk's first row has none.
    // u[-1] and u[n - 1] are in the caller's buffer (d[n - 1] and the padding entry).
    auto has_u = [&](int i) { return TRANS ? i > 0 : i + 1 < n; };
    // full chunks: rows b .. b + CH - 1 with b = c CH (ascending) or n - (c + 1) CH (descending)
    const int nfull = n / CH;
    auto base = [&](int c) { return ASC ? c * CH : n - (c + 1) * CH; };
    double cx[CH], cd[CH], cu[CH], nx[CH], nd[CH], nu[CH];
    auto load = [&](int b, double *vx, double *vd, double *vu) {
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            vx[k] = x[b + k];
            vd[k] = d[b + k];
            vu[k] = u[TRANS ? b + k - 1 : b + k];
        }
    };
    if (nfull > 0) load(base(0), cx, cd, cu);
    double carry = 0.0;  // the last new entry
    for (int c = 0; c < nfull; ++c) {
        const int b = base(c);
        load(c + 1 < nfull ? base(c + 1) : b, nx, nd, nu);
#pragma unroll
        for (int kk = 0; kk < CH; ++kk) {
            const int k = ASC ? kk : CH - 1 - kk;
            const int i = b + k;
            carry = (cx[k] - (has_u(i) ? cu[k] * carry : 0.0)) / cd[k];
            x[i] = carry;
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            cx[k] = nx[k];
            cd[k] = nd[k];
            cu[k] = nu[k];
        }
    }
    // the remaining n mod CH rows (ascending: the last ones, descending: the first ones)
    const int rem = n - nfull * CH;
    if (rem == 0) return;
    const int b = ASC ? nfull * CH : 0;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const bool in = k < rem;
        cx[k] = in ? x[b + k] : 0.0;
        cd[k] = in ? d[b + k] : 1.0;
        cu[k] = in && has_u(b + k) ? u[TRANS ? b + k - 1 : b + k] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
        const int k = ASC ? kk : CH - 1 - kk;
        if (k >= rem) continue;
        const int i = b + k;
        carry = (cx[k] - (has_u(i) ? cu[k] * carry : 0.0)) / cd[k];
        x[i] = carry;
    }
}

// The products have no recurrence: one wave per column, lane l on row r0 + l
// of each 64-row segment (coalesced 512-byte accesses), its neighbour from
// the next lane (R: x_{i+1}) or the previous one (R^T: x_{i-1}), the segment
// edge's from memory.  Segments run in the order that reads every neighbour
// before its own segment overwrites it: upward for R, downward for R^T.
template <bool TRANS>
__global__ void __launch_bounds__(256) bidiag_mul_kernel(double *A, int n, int count, const double *d,
                                                        const double *u) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= count) return;  // whole waves
    double *x = bidiag_column(A, n, q);
    const int nseg = (n + 63) / 64;
    for (int s = 0; s < nseg; ++s) {
        const int r0 = (TRANS ? nseg - 1 - s : s) * 64, i = r0 + lane;
        const bool in = i < n;
        const double xi = in ? x[i] : 0.0;
        double y;
        if (!TRANS) {  // x_i = d_i x_i + u_i x_{i+1}
            const double edge = r0 + 64 < n ? x[r0 + 64] : 0.0;
            const double nb = __shfl_down(xi, 1, 64);
            const double xn = lane == 63 ? edge : nb;
            y = in ? (i + 1 < n ? d[i] * xi + u[i] * xn : d[i] * xi) : 0.0;
        } else {  // x_i = u_{i-1} x_{i-1} + d_i x_i
            const double edge = r0 > 0 ? x[r0 - 1] : 0.0;
            const double nb = __shfl_up(xi, 1, 64);
            const double xp = lane == 0 ? edge : nb;
            y = in ? (i > 0 ? u[i - 1] * xp + d[i] * xi : d[i] * xi) : 0.0;
        }
        if (in) x[i] = y;
    }
}

}  // namespace fttrmm

extern "C" int ftamd_hip_bidiag(double *A, int n, int M, const double *R, int solve, int trans, void *stream) {
    if (n < 1 || M < 0 || !A || !R) return FTAMD_EINVAL;
    const int count = M / 4 * 2 + (M % 4 >= 2 ? 1 : 0) + (M % 4 >= 3 ? 1 : 0);  // columns j < M, j = 1, 2 mod 4
    if (count == 0) return FTAMD_OK;
    hipStream_t st = (hipStream_t)stream;
    const double *d = R, *u = R + n;
    const dim3 sgrid((count + 63) / 64), sblock(64);  // a thread per column
    const dim3 mgrid((count + 3) / 4), mblock(256);   // a wave per column
    if (solve && trans) hipLaunchKernelGGL(fttrmm::bidiag_solve_kernel<true>, sgrid, sblock, 0, st, A, n, count, d, u);
    else if (solve) hipLaunchKernelGGL(fttrmm::bidiag_solve_kernel<false>, sgrid, sblock, 0, st, A, n, count, d, u);
    else if (trans) hipLaunchKernelGGL(fttrmm::bidiag_mul_kernel<true>, mgrid, mblock, 0, st, A, n, count, d, u);
    else hipLaunchKernelGGL(fttrmm::bidiag_mul_kernel<false>, mgrid, mblock, 0, st, A, n, count, d, u);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}

extern "C" int ftamd_hip_trmm_launch(const FtamdTrmmLaunch *a, void *stream) {
    if (a->n < 1 || a->M < 0 || a->cstride < 1 || a->ncls < 1 || a->ncls > 4 || !a->X || !a->Y) return FTAMD_EINVAL;
    if (a->right && (a->ncls != 1 || a->cstride != 1 || a->M > a->n)) return FTAMD_EINVAL;
    int tiles = 0;
    for (int c = 0; c < a->ncls; ++c) {
        const int count = a->M > c ? (a->M - c + a->cstride - 1) / a->cstride : 0;
        if (a->ctiles[c] != (count + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS) return FTAMD_EINVAL;
        tiles += a->ctiles[c];
    }
    const int rows = (a->n + FTAMD_TRMM_ROWS - 1) / FTAMD_TRMM_ROWS;
    if (tiles == 0) return FTAMD_OK;
    if (a->right) hipLaunchKernelGGL(fttrmm::trmm_kernel<true>, dim3(tiles, rows), dim3(256), 0, (hipStream_t)stream, *a);
    else hipLaunchKernelGGL(fttrmm::trmm_kernel<false>, dim3(tiles, rows), dim3(256), 0, (hipStream_t)stream, *a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : (e == hipErrorOutOfMemory ? FTAMD_ENOMEM : FTAMD_EDEVICE);
}

extern "C" int ftamd_hip_cheb_scale(double *A, int n, int M, double s1, double s2, void *stream) {
    if (n < 1 || M < 0 || !A) return FTAMD_EINVAL;
    if (M == 0) return FTAMD_OK;
    const long long total = (long long)n * M;
    const long long want = (total + 255) / 256;
    const int blocks = (int)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(fttrmm::cheb_scale_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, A, n, M, s1, s2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}
