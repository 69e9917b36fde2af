// ft_grid.hip -- synthesis and analysis on the sphere's tensor-product grid
// (fftw.c:62-270): per column of the n x M coefficient array, a DCT or DST
// of length N along theta (FFTW's REDFT01/RODFT01 for synthesis plans,
// REDFT10/RODFT10 for analysis plans), and per row a real DFT of length M
// along phi (c2r / r2c), with the reference's scalings and the column layout
// of its data_r2c / data_c2r (fftw.c:40-60).
//
// The DCTs run as complex FFTs of length N (hipFFT; the synthesis plans put
// two columns in one FFT) between two HIP kernels, by Makhoul's reordering: DCT-II(x)_k = 2 Re(e^{-i pi k/2N} FFT(v)_k) with
// v_m = x_{2m}, v_{N-1-m} = x_{2m+1}; DCT-III(X) is the inverse map,
// v = IFFT(e^{i pi k/2N} (X_k - i X_{N-k})) (unnormalized, X_N = 0),
// y_{2m} = Re v_m, y_{2m+1} = Re v_{N-1-m}.  The DSTs are DCTs of the
// reversed input with alternating output signs (DST-III) or of the
// sign-alternated input with reversed output (DST-II).  The phi DFTs are
// hipFFT's strided transforms (row i is the transform at element i, stride
// N); with N even, rows 2p and 2p+1 share one complex DFT over X read as
// N/2 x M complex.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <math.h>
#include <stdlib.h>

#include "ft_hip_abi.h"

struct FtamdGrid {
    int N, M, type3;
    int plan_sphv;  // column classes of the plan: 0 the sphere's, 1 the vector field's, 3 all cosine (annulus)
    int disk;     // the disk's plans: DCT-IV (REDFT11) on the sine class
    int nsin;     // sine-class columns (disk)
    int pair;     // phi DFTs of two real rows as one complex DFT (N even)
    int tpair;    // theta FFTs of two columns as one (synthesis, not the disk): column j in part j & 1 of line j / 2
    hipfftHandle theta, phi, iv;
    hipfftHandle phi_r;  // the real phi DFTs of a paired plan, made on first use (X not 16-byte aligned)
    double2 *tw;  // e^{i pi k / 2N}, k < N
    double2 *tw4; // e^{i pi (2k+1) / 4N}, k < N (disk)
    double2 *Z;   // N x M complex: the columns' FFTs
    double2 *Z4;  // 2N x nsin complex: the disk's sine-class columns, zero-padded (disk)
    double2 *Y;   // N x (M/2+1) complex: the rows' half spectra (paired, not the disk: N/2 x M full spectra)
    double2 *Yp;  // the disk's paired plans: N/2 x M full spectra of the row pairs
    double *w;    // the annulus's radial weights (N), NULL otherwise
};

namespace ftgrid {

// Columns of the cosine class: j mod 4 in {0, 3} (fftw.c:84-96: plans 1 and
// 4 take kind[0], plans 2 and 3 kind[1]); the vector-field plans (mode 1)
// swap the kinds; the annulus's (mode 3, fftw.c:560-575) have one kind.
__device__ __forceinline__ bool cos_class(int j, int mode) {
    if (mode == 3) return true;
    const int r = j & 3;
    return ((r == 0) || (r == 3)) != (mode == 1);
}

// The annulus's radial weights on columns 1 and 2 mod 4 (fftw.c:588-592):
// wmode 1 multiplies, 2 divides, 0 leaves the entry.
__device__ __forceinline__ double weigh(double v, int j, int n, const double *w, int wmode) {
    const int r = j & 3;
    if (wmode == 0 || (r != 1 && r != 2)) return v;
    return wmode == 1 ? v * w[n] : v / w[n];
}

// Pipeline with DCT-III first (synthesis plans): fix-ups, DCT-III / DST-III
// of every column into Z (pre-twiddled for the inverse FFT).  skip_sin: the
// sine class is the disk's DCT-IV, run by the iv kernels.
__device__ __forceinline__ double2 pre3_value(const double *__restrict__ X, const double2 *__restrict__ tw, int N,
                                               int k, int j, int plan_sphv, int exec_sphv, double f_row0_cos,
                                               double f_last_sin) {
    const double *x = X + (long long)N * j;
    const bool tcos = cos_class(j, plan_sphv), fcos = cos_class(j, exec_sphv);
    // entry e of the column after the executor's fix-ups
    auto val = [&](int e) {
        double v = x[e];
        if (fcos && e == 0) v *= f_row0_cos;
        if (!fcos && e == N - 1) v *= f_last_sin;
        return v;
    };
    // input of the DCT-III: the column itself (cosine class) or reversed (sine class)
    auto in = [&](int e) { return tcos ? val(e) : val(N - 1 - e); };
    const double re = in(k), im = k == 0 ? 0.0 : -in(N - k);
    const double2 t = tw[k];
    return make_double2(t.x * re - t.y * im, t.y * re + t.x * im);
}

__global__ void pre3(const double *__restrict__ X, double2 *__restrict__ Z, const double2 *__restrict__ tw, int N,
                     int M, int plan_sphv, int exec_sphv, double f_row0_cos, double f_last_sin, int skip_sin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= N || j >= M) return;
    if (skip_sin && !cos_class(j, plan_sphv)) return;
    Z[(long long)N * j + k] = pre3_value(X, tw, N, k, j, plan_sphv, exec_sphv, f_row0_cos, f_last_sin);
}

// Columns 2c and 2c+1 share one inverse FFT: their inputs Z0, Z1 (whose
// transforms are real) go in as Z0 + i Z1, so the transform holds column 2c
// in its real and column 2c+1 in its imaginary part.  A thread per (k, c).
__global__ void pre3_pair(const double *__restrict__ X, double2 *__restrict__ Z, const double2 *__restrict__ tw,
                          int N, int M, int plan_sphv, int exec_sphv, double f_row0_cos, double f_last_sin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= N || 2 * c >= M) return;
    const double2 z0 = pre3_value(X, tw, N, k, 2 * c, plan_sphv, exec_sphv, f_row0_cos, f_last_sin);
    const double2 z1 = 2 * c + 1 < M ? pre3_value(X, tw, N, k, 2 * c + 1, plan_sphv, exec_sphv, f_row0_cos, f_last_sin)
                                     : make_double2(0.0, 0.0);
    Z[(long long)N * c + k] = make_double2(z0.x - z1.y, z0.y + z1.x);
}

// After the inverse FFT: y_n of every column, signs of the sine class, the
// scalings a (all) and b (column 0 again), into Y as data_r2c lays it out.
__device__ __forceinline__ double post3_value(const double2 *__restrict__ Z, int N, int n, int j, int plan_sphv,
                                              double a, double b, const double *__restrict__ w, int wmode, int tp) {
    const int src = (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
    double y = tp ? ((const double *)Z)[2 * ((long long)N * (j >> 1) + src) + (j & 1)] : Z[(long long)N * j + src].x;
    if (!cos_class(j, plan_sphv) && (n & 1)) y = -y;
    y = weigh(y, j, n, w, wmode);
    y *= a;
    if (j == 0) y *= b;
    return y;
}

__global__ void post3(const double2 *__restrict__ Z, double2 *__restrict__ Y, int N, int M, int plan_sphv, double a,
                      double b, int skip_sin, const double *__restrict__ w, int wmode, int tp) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (n >= N || j >= M) return;
    if (skip_sin && !cos_class(j, plan_sphv)) return;
    const double y = post3_value(Z, N, n, j, plan_sphv, a, b, w, wmode, tp);
    double *Yd = (double *)Y;
    if (j == 0) {
        Y[n] = make_double2(y, 0.0);
    } else {
        const int q = (j + 1) >> 1;
        Yd[2 * ((long long)N * q + n) + ((j & 1) ? 1 : 0)] = (j & 1) ? -y : y;
    }
}

// The same for rows paired into one complex DFT (N even): rows 2p and 2p+1
// have half spectra Y0, Y1; the pair's full spectrum is Zp[q] = Y0[q] + i
// Y1[q], Zp[M-q] = conj(Y0[q]) + i conj(Y1[q]), at Yp[p + q N/2], whose
// inverse DFT is row 2p in its real and row 2p+1 in its imaginary parts: X
// itself read as N/2 x M complex.  A thread per (p, q <= M/2).
__global__ void post3_pair(const double2 *__restrict__ Z, double2 *__restrict__ Yp, int N, int M, int plan_sphv,
                           double a, double b, const double *__restrict__ w, int wmode, int tp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    const int P = N >> 1;
    if (p >= P || 2 * q > M) return;
    const int e0 = 2 * p, e1 = 2 * p + 1;
    // real parts from column 2q, imaginary parts (data_r2c: minus the value) from column 2q - 1
    const double r0 = post3_value(Z, N, e0, 2 * q, plan_sphv, a, b, w, wmode, tp);
    const double r1 = post3_value(Z, N, e1, 2 * q, plan_sphv, a, b, w, wmode, tp);
    const double i0 = q > 0 ? -post3_value(Z, N, e0, 2 * q - 1, plan_sphv, a, b, w, wmode, tp) : 0.0;
    const double i1 = q > 0 ? -post3_value(Z, N, e1, 2 * q - 1, plan_sphv, a, b, w, wmode, tp) : 0.0;
    Yp[p + (long long)q * P] = make_double2(r0 - i1, i0 + r1);
    if (q > 0) Yp[p + (long long)(M - q) * P] = make_double2(r0 + i1, r1 - i0);
}

// The disk's paired plans, whose sine class reaches Y through the DCT-IV
// kernels: the half spectra Y of rows 2p and 2p+1 into the pair's full
// spectrum (as post3_pair; the imaginary part of each row's mean is taken
// as 0, as a c2r DFT takes it), and back (as pre2_pair).
__global__ void pack_pairs(const double2 *__restrict__ Y, double2 *__restrict__ Yp, int N, int M) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    const int P = N >> 1;
    if (p >= P || 2 * q > M) return;
    const double2 y0 = Y[2 * p + (long long)q * N], y1 = Y[2 * p + 1 + (long long)q * N];
    const double i0 = q > 0 ? y0.y : 0.0, i1 = q > 0 ? y1.y : 0.0;
    Yp[p + (long long)q * P] = make_double2(y0.x - i1, i0 + y1.x);
    if (q > 0) Yp[p + (long long)(M - q) * P] = make_double2(y0.x + i1, y1.x - i0);
}

__global__ void unpack_pairs(const double2 *__restrict__ Yp, double2 *__restrict__ Y, int N, int M) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (n >= N || 2 * q > M) return;
    const int P = N >> 1, p = n >> 1;
    const double2 zk = Yp[p + (long long)q * P], zm = Yp[p + (long long)(q == 0 ? 0 : M - q) * P];
    Y[n + (long long)q * N] = (n & 1) == 0 ? make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y))
                                           : make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x));
}

// Pipeline with the row DFT first (analysis plans): data_c2r, the scalings,
// the sine class's alternating signs, Makhoul's reordering into Z.
__device__ __forceinline__ void pre2_store(double x, double2 *__restrict__ Z, int N, int n, int j, int plan_sphv,
                                           double a, double b, const double *__restrict__ w, int wmode) {
    x *= a;
    if (j == 0) x *= b;
    x = weigh(x, j, n, w, wmode);
    if (!cos_class(j, plan_sphv) && (n & 1)) x = -x;
    const int dst = (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
    Z[(long long)N * j + dst] = make_double2(x, 0.0);
}

__global__ void pre2(const double2 *__restrict__ Y, double2 *__restrict__ Z, int N, int M, int plan_sphv, double a,
                     double b, int skip_sin, const double *__restrict__ w, int wmode) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (n >= N || j >= M) return;
    if (skip_sin && !cos_class(j, plan_sphv)) return;
    const double *Yd = (const double *)Y;
    double x;
    if (j == 0) {
        x = Yd[2 * n];
    } else {
        const int q = (j + 1) >> 1;
        const double v = Yd[2 * ((long long)N * q + n) + ((j & 1) ? 1 : 0)];
        x = (j & 1) ? -v : v;
    }
    pre2_store(x, Z, N, n, j, plan_sphv, a, b, w, wmode);
}

// pre2 from the paired rows' spectrum Zp (the forward DFT of X read as N/2 x
// M complex): Y0[q] = (Zp[q] + conj(Zp[M-q])) / 2, Y1[q] = (Zp[q] -
// conj(Zp[M-q])) / 2i.
__global__ void pre2_pair(const double2 *__restrict__ Yp, double2 *__restrict__ Z, int N, int M, int plan_sphv,
                          double a, double b, const double *__restrict__ w, int wmode) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (n >= N || j >= M) return;
    const int P = N >> 1, p = n >> 1, q = (j + 1) >> 1;
    const double2 zk = Yp[p + (long long)q * P], zm = Yp[p + (long long)(q == 0 ? 0 : M - q) * P];
    double re, im;
    if ((n & 1) == 0) {
        re = 0.5 * (zk.x + zm.x);
        im = 0.5 * (zk.y - zm.y);
    } else {
        re = 0.5 * (zk.y + zm.y);
        im = 0.5 * (zm.x - zk.x);
    }
    // data_c2r: column 2q the real part, column 2q - 1 minus the imaginary part
    const double x = (j & 1) ? -im : re;
    pre2_store(x, Z, N, n, j, plan_sphv, a, b, w, wmode);
}

// After the forward FFT: the DCT-II (cosine class) or DST-II (sine class) of
// every column, then the executor's fix-ups, into X.
__global__ void post2(const double2 *__restrict__ Z, double *__restrict__ X, const double2 *__restrict__ tw, int N,
                      int M, int plan_sphv, int exec_sphv, double g_row0_cos, double g_last_sin, int skip_sin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= N || j >= M) return;
    if (skip_sin && !cos_class(j, plan_sphv)) return;
    const bool tcos = cos_class(j, plan_sphv), fcos = cos_class(j, exec_sphv);
    const int src = tcos ? k : N - 1 - k;
    const double2 V = Z[(long long)N * j + src];
    const double2 t = tw[src];
    double y = 2.0 * (t.x * V.x + t.y * V.y);
    if (fcos && k == 0) y *= g_row0_cos;
    if (!fcos && k == N - 1) y *= g_last_sin;
    X[(long long)N * j + k] = y;
}

// ---- the disk's DCT-IV (REDFT11) on its sine class (fftw.c:450-458):
// y_k = 2 sum_n x_n cos(pi (2n+1)(2k+1) / 4N) = 2 Re(e^{-i pi (2k+1)/4N} U_k),
// U the 2N-point DFT of u_n = x_n e^{-i pi n / 2N} (n < N, zero above).
// Sine-class column s (of nsin) is column 4(s/2) + 1 + s%2.

__device__ __forceinline__ int sin_col(int s) { return 4 * (s >> 1) + 1 + (s & 1); }

// u from the columns of X (synthesis plans; the executor's last-row fix-up f1)
__global__ void pre4_x(const double *__restrict__ X, double2 *__restrict__ Z4, const double2 *__restrict__ tw, int N,
                       int nsin, double f_last_sin) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (n >= 2 * N || s >= nsin) return;
    double2 u = make_double2(0.0, 0.0);
    if (n < N) {
        double x = X[(long long)N * sin_col(s) + n];
        if (n == N - 1) x *= f_last_sin;
        const double2 t = tw[n];
        u = make_double2(x * t.x, -x * t.y);
    }
    Z4[(long long)2 * N * s + n] = u;
}

// u from the half spectra Y after the phi DFT (analysis plans): data_c2r and the scaling a
__global__ void pre4_y(const double2 *__restrict__ Y, double2 *__restrict__ Z4, const double2 *__restrict__ tw, int N,
                       int nsin, double a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (n >= 2 * N || s >= nsin) return;
    double2 u = make_double2(0.0, 0.0);
    if (n < N) {
        const int j = sin_col(s), q = (j + 1) >> 1;
        const double v = ((const double *)Y)[2 * ((long long)N * q + n) + ((j & 1) ? 1 : 0)];
        double x = (j & 1) ? -v : v;
        x *= a;
        const double2 t = tw[n];
        u = make_double2(x * t.x, -x * t.y);
    }
    Z4[(long long)2 * N * s + n] = u;
}

// y into Y (synthesis plans): the scaling a, data_r2c
__global__ void post4_y(const double2 *__restrict__ Z4, double2 *__restrict__ Y, const double2 *__restrict__ tw4,
                        int N, int nsin, double a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= N || s >= nsin) return;
    const double2 U = Z4[(long long)2 * N * s + k], t = tw4[k];
    double y = 2.0 * (t.x * U.x + t.y * U.y);
    y *= a;
    const int j = sin_col(s), q = (j + 1) >> 1;
    ((double *)Y)[2 * ((long long)N * q + k) + ((j & 1) ? 1 : 0)] = (j & 1) ? -y : y;
}

// y into X (analysis plans), the executor's last-row fix-up g1
__global__ void post4_x(const double2 *__restrict__ Z4, double *__restrict__ X, const double2 *__restrict__ tw4, int N,
                        int nsin, double g_last_sin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= N || s >= nsin) return;
    const double2 U = Z4[(long long)2 * N * s + k], t = tw4[k];
    double y = 2.0 * (t.x * U.x + t.y * U.y);
    if (k == N - 1) y *= g_last_sin;
    X[(long long)N * sin_col(s) + k] = y;
}

// ---- the triangle's and the tetrahedron's DCTs (fftw.c:280-400): FFTW's
// r2r_2d(N, M) / r2r_3d(N, L, M) over the buffer as a row-major array, one
// pass per axis.  Line b along an axis of length len and stride st (the
// product of the later axes' lengths) holds the elements
// e = (b / st) * len * st + b % st + t * st, t < len.  The executors' face
// scalings act where a coordinate of e in the reference's column-major view
// (e = i + N (j + L k)) is 0.

struct AxisScale {
    int pre_fix;     // x2 per zero coordinate (before the transform)
    double pre_mul;  // then x pre_mul
    double pre_div;  // then / pre_div (0: none)
    int post_fix;    // x0.5 per zero coordinate (after the transform)
    double post_mul; // then x post_mul
    double post_div; // then / post_div (0: none)
    int rank;        // 2 (triangle) or 3 (tetrahedron)
    int cN, cL;      // the column-major view's first two sizes (cL unused for rank 2)
    // the rectangle-disk's (fftw.c:657-740): a factor on one set of entries of
    // the N x M column-major array before (pre_mask) and after (post_mask)
    // the transform -- 1: row 0 of the even columns, 2: the last row of the
    // odd columns, 3: column 0 -- and DSTs on the odd lines (sin_odd)
    int pre_mask, post_mask;
    double pre_mask_f, post_mask_f;
    int sin_odd;
};

__device__ __forceinline__ double masked(double v, long long e, int mask, double f, int cN) {
    if (mask == 0) return v;
    const long long i = e % cN, j = e / cN;
    const bool on = mask == 1 ? (i == 0 && (j & 1) == 0) : mask == 2 ? (i == cN - 1 && (j & 1)) : j == 0;
    return on ? v * f : v;
}

__device__ __forceinline__ int zero_coords(long long e, const AxisScale &sc) {
    if (sc.rank == 2) return (e % sc.cN == 0) + (e < sc.cN);
    const long long i = e % sc.cN, jk = e / sc.cN;
    return (i == 0) + (jk % sc.cL == 0) + (jk < sc.cL);
}

__device__ __forceinline__ double scale_pre(double v, long long e, const AxisScale &sc) {
    v = masked(v, e, sc.pre_mask, sc.pre_mask_f, sc.cN);
    if (sc.pre_fix)
        for (int z = zero_coords(e, sc); z > 0; --z) v *= 2.0;
    v *= sc.pre_mul;
    if (sc.pre_div != 0.0) v /= sc.pre_div;
    return v;
}

__device__ __forceinline__ double scale_post(double v, long long e, const AxisScale &sc) {
    v = masked(v, e, sc.post_mask, sc.post_mask_f, sc.cN);
    if (sc.post_fix)
        for (int z = zero_coords(e, sc); z > 0; --z) v *= 0.5;
    v *= sc.post_mul;
    if (sc.post_div != 0.0) v /= sc.post_div;
    return v;
}

__device__ __forceinline__ long long line_elem(int b, int t, int len, long long st) {
    return (b / st) * len * st + b % st + t * st;
}

__global__ void axis_pre3(const double *__restrict__ X, double2 *__restrict__ Z, const double2 *__restrict__ tw,
                          int len, int batch, long long st, AxisScale sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y + blockIdx.z * 65535;
    if (k >= len || b >= batch) return;
    const bool sn = sc.sin_odd && (b & 1);
    auto val = [&](int t) {
        if (sn) t = len - 1 - t;  // DST-III: the DCT-III of the reversed line
        const long long e = line_elem(b, t, len, st);
        return scale_pre(X[e], e, sc);
    };
    const double re = val(k), im = k == 0 ? 0.0 : -val(len - k);
    const double2 t = tw[k];
    Z[(long long)len * b + k] = make_double2(t.x * re - t.y * im, t.y * re + t.x * im);
}

__global__ void axis_post3(const double2 *__restrict__ Z, double *__restrict__ X, int len, int batch, long long st,
                           AxisScale sc) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y + blockIdx.z * 65535;
    if (n >= len || b >= batch) return;
    const int src = (n & 1) ? len - 1 - (n >> 1) : (n >> 1);
    const long long e = line_elem(b, n, len, st);
    double y = Z[(long long)len * b + src].x;
    if (sc.sin_odd && (b & 1) && (n & 1)) y = -y;  // DST-III: alternating signs
    X[e] = scale_post(y, e, sc);
}

__global__ void axis_pre2(const double *__restrict__ X, double2 *__restrict__ Z, int len, int batch, long long st,
                          AxisScale sc) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y + blockIdx.z * 65535;
    if (n >= len || b >= batch) return;
    const long long e = line_elem(b, n, len, st);
    const int dst = (n & 1) ? len - 1 - (n >> 1) : (n >> 1);
    double x = scale_pre(X[e], e, sc);
    if (sc.sin_odd && (b & 1) && (n & 1)) x = -x;  // DST-II: the DCT-II of the sign-alternated line
    Z[(long long)len * b + dst] = make_double2(x, 0.0);
}

__global__ void axis_post2(const double2 *__restrict__ Z, double *__restrict__ X, const double2 *__restrict__ tw,
                           int len, int batch, long long st, AxisScale sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y + blockIdx.z * 65535;
    if (k >= len || b >= batch) return;
    const int src = sc.sin_odd && (b & 1) ? len - 1 - k : k;  // DST-II: reversed output
    const double2 V = Z[(long long)len * b + src], t = tw[src];
    const long long e = line_elem(b, k, len, st);
    X[e] = scale_post(2.0 * (t.x * V.x + t.y * V.y), e, sc);
}

}  // namespace ftgrid

static int fft_ok(hipfftResult r) { return r == HIPFFT_SUCCESS ? FTAMD_OK : FTAMD_EDEVICE; }

static int phi_plan_real(FtamdGrid *g, hipfftHandle *h) {
    const int N = g->N, M = g->M;
    int n2[1] = {M}, nc[1] = {M / 2 + 1};
    if (g->type3) return fft_ok(hipfftPlanMany(h, 1, n2, nc, N, 1, n2, N, 1, HIPFFT_Z2D, N));
    return fft_ok(hipfftPlanMany(h, 1, n2, n2, N, 1, nc, N, 1, HIPFFT_D2Z, N));
}

static int phi_plan_pair(FtamdGrid *g) {
    const int P = g->N / 2;
    int n2[1] = {g->M};
    return fft_ok(hipfftPlanMany(&g->phi, 1, n2, n2, P, 1, n2, P, 1, HIPFFT_Z2Z, P));
}

extern "C" int ftamd_grid_create(int N, int M, int type3, int plan_mode, const double *w, FtamdGrid **out) {
    *out = NULL;
    /* columns are the launch grid's y dimension (at most 65535) */
    if (N < 1 || M < 1 || M % 2 == 0 || M > 65535 || plan_mode < 0 || plan_mode > 3 || (plan_mode == 3 && !w))
        return FTAMD_EINVAL;
    FtamdGrid *g = (FtamdGrid *)calloc(1, sizeof(FtamdGrid));
    if (!g) return FTAMD_ENOMEM;
    g->N = N;
    g->M = M;
    g->type3 = type3;
    g->plan_sphv = plan_mode == 1 ? 1 : plan_mode == 3 ? 3 : 0;
    g->disk = plan_mode == 2;
    g->nsin = g->disk ? (M + 2) / 4 + (M + 1) / 4 : 0;  /* columns 1 and 2 mod 4 (fftw.c:427-431) */
    g->pair = N % 2 == 0;
    /* the synthesis plans only: splitting a paired forward FFT (analysis) costs the post kernel a
     * second read per entry and measured slower (n = 1024: 0.108 against 0.104 ms) */
    g->tpair = !g->disk && type3;
    g->theta = g->phi = g->iv = g->phi_r = 0;
    int rc = FTAMD_OK;
    const size_t zbytes = sizeof(double2) * (size_t)N * M, ybytes = sizeof(double2) * (size_t)N * (M / 2 + 1);
    double2 *tw = (double2 *)malloc(sizeof(double2) * (size_t)N);
    if (!tw) rc = FTAMD_ENOMEM;
    for (int k = 0; rc == FTAMD_OK && k < N; k++) {
        const long double th = 3.141592653589793238462643383279502884L * k / (2.0L * N);
        tw[k] = make_double2((double)cosl(th), (double)sinl(th));
    }
    if (rc == FTAMD_OK && hipMalloc((void **)&g->tw, sizeof(double2) * (size_t)N) != hipSuccess) rc = FTAMD_ENOMEM;
    if (rc == FTAMD_OK && hipMalloc((void **)&g->Z, zbytes) != hipSuccess) rc = FTAMD_ENOMEM;
    if (rc == FTAMD_OK && hipMalloc((void **)&g->Y, ybytes) != hipSuccess) rc = FTAMD_ENOMEM;
    if (rc == FTAMD_OK && g->pair && g->disk &&
        hipMalloc((void **)&g->Yp, sizeof(double2) * (size_t)(N / 2) * M) != hipSuccess)
        rc = FTAMD_ENOMEM;
    if (rc == FTAMD_OK && hipMemcpy(g->tw, tw, sizeof(double2) * (size_t)N, hipMemcpyHostToDevice) != hipSuccess)
        rc = FTAMD_EDEVICE;
    free(tw);
    if (rc == FTAMD_OK && g->disk && g->nsin > 0) {
        double2 *h = (double2 *)malloc(sizeof(double2) * (size_t)N);
        if (!h) rc = FTAMD_ENOMEM;
        for (int k = 0; rc == FTAMD_OK && k < N; k++) {
            const long double th = 3.141592653589793238462643383279502884L * (2 * k + 1) / (4.0L * N);
            h[k] = make_double2((double)cosl(th), (double)sinl(th));
        }
        if (rc == FTAMD_OK && hipMalloc((void **)&g->tw4, sizeof(double2) * (size_t)N) != hipSuccess) rc = FTAMD_ENOMEM;
        if (rc == FTAMD_OK && hipMemcpy(g->tw4, h, sizeof(double2) * (size_t)N, hipMemcpyHostToDevice) != hipSuccess)
            rc = FTAMD_EDEVICE;
        free(h);
        if (rc == FTAMD_OK && hipMalloc((void **)&g->Z4, sizeof(double2) * 2 * (size_t)N * g->nsin) != hipSuccess)
            rc = FTAMD_ENOMEM;
        int n4[1] = {2 * N};
        if (rc == FTAMD_OK) rc = fft_ok(hipfftPlanMany(&g->iv, 1, n4, n4, 1, 2 * N, n4, 1, 2 * N, HIPFFT_Z2Z, g->nsin));
    }
    if (rc == FTAMD_OK && plan_mode == 3) {
        if (hipMalloc((void **)&g->w, sizeof(double) * (size_t)N) != hipSuccess) rc = FTAMD_ENOMEM;
        else if (hipMemcpy(g->w, w, sizeof(double) * (size_t)N, hipMemcpyHostToDevice) != hipSuccess) rc = FTAMD_EDEVICE;
    }
    /* the sine-class slots of Z stay zero in the disk's plans (never written) */
    if (rc == FTAMD_OK && g->disk && hipMemset(g->Z, 0, zbytes) != hipSuccess) rc = FTAMD_EDEVICE;
    // theta: M contiguous complex transforms of length N
    int n1[1] = {N};
    if (rc == FTAMD_OK)
        rc = fft_ok(hipfftPlanMany(&g->theta, 1, n1, n1, 1, N, n1, 1, N, HIPFFT_Z2Z, g->tpair ? (M + 1) / 2 : M));
    // phi: N real transforms of length M, element i of row-transform e at e + i*N; paired, N/2
    // complex ones, element i of pair p at p + i*N/2
    if (rc == FTAMD_OK) rc = g->pair ? phi_plan_pair(g) : phi_plan_real(g, &g->phi);
    if (rc != FTAMD_OK) {
        ftamd_grid_destroy(g);
        return rc;
    }
    *out = g;
    return FTAMD_OK;
}

extern "C" void ftamd_grid_destroy(FtamdGrid *g) {
    if (!g) return;
    hipDeviceSynchronize();
    if (g->theta) hipfftDestroy(g->theta);
    if (g->phi) hipfftDestroy(g->phi);
    if (g->phi_r) hipfftDestroy(g->phi_r);
    if (g->iv) hipfftDestroy(g->iv);
    hipFree(g->tw);
    hipFree(g->tw4);
    hipFree(g->Z);
    hipFree(g->Z4);
    hipFree(g->w);
    hipFree(g->Y);
    hipFree(g->Yp);
    free(g);
}

extern "C" int ftamd_grid_run(FtamdGrid *g, double *X, int exec_sphv, double f0, double f1, double a, double b,
                              int wmode, void *stream) {
    if (!g || !X) return FTAMD_EINVAL;
    const int N = g->N, M = g->M;
    const hipStream_t st = (hipStream_t)stream;
    const dim3 block(256), grid((N + 255) / 256, M);
    // the paired phi DFTs read or write X as complex: 16-byte aligned X only, else the real DFTs
    const int pair = g->pair && ((uintptr_t)X & 15) == 0;
    if (g->pair && !pair && !g->phi_r) {
        const int r = phi_plan_real(g, &g->phi_r);
        if (r != FTAMD_OK) return r;
    }
    const hipfftHandle phi = g->pair && !pair ? g->phi_r : g->phi;
    const dim3 gridp((N / 2 + 255) / 256, M / 2 + 1);
    const int tp = g->tpair;
    int rc = fft_ok(hipfftSetStream(g->theta, st));
    if (rc == FTAMD_OK) rc = fft_ok(hipfftSetStream(phi, st));
    if (rc == FTAMD_OK && g->iv) rc = fft_ok(hipfftSetStream(g->iv, st));
    if (rc != FTAMD_OK) return rc;
    const int iv = g->disk && g->nsin > 0;
    const dim3 grid4((2 * N + 255) / 256, g->nsin > 0 ? g->nsin : 1), grid4o((N + 255) / 256, g->nsin > 0 ? g->nsin : 1);
    // the disk's sine class: DCT-IV of the columns (synthesis plans, before the phi DFT)
    auto iv_from_x = [&]() -> int {
        hipLaunchKernelGGL(ftgrid::pre4_x, grid4, block, 0, st, X, g->Z4, g->tw, N, g->nsin, f1);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        int r = fft_ok(hipfftExecZ2Z(g->iv, (hipfftDoubleComplex *)g->Z4, (hipfftDoubleComplex *)g->Z4, HIPFFT_FORWARD));
        if (r != FTAMD_OK) return r;
        hipLaunchKernelGGL(ftgrid::post4_y, grid4o, block, 0, st, g->Z4, g->Y, g->tw4, N, g->nsin, a);
        return hipGetLastError() == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
    };
    // ... and from the half spectra (analysis plans, after the phi DFT)
    auto iv_from_y = [&]() -> int {
        hipLaunchKernelGGL(ftgrid::pre4_y, grid4, block, 0, st, g->Y, g->Z4, g->tw, N, g->nsin, a);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        int r = fft_ok(hipfftExecZ2Z(g->iv, (hipfftDoubleComplex *)g->Z4, (hipfftDoubleComplex *)g->Z4, HIPFFT_FORWARD));
        if (r != FTAMD_OK) return r;
        hipLaunchKernelGGL(ftgrid::post4_x, grid4o, block, 0, st, g->Z4, X, g->tw4, N, g->nsin, f1);
        return hipGetLastError() == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
    };
    if (g->type3) {
        // fix-ups f0 (row 0, cosine columns) and f1 (last row, sine columns),
        // DCT-III / DST-III (the disk: DCT-IV on the sine class), scalings,
        // data_r2c, c2r
        if (tp)
            hipLaunchKernelGGL(ftgrid::pre3_pair, dim3((N + 255) / 256, (M + 1) / 2), block, 0, st, X, g->Z, g->tw, N,
                               M, g->plan_sphv, exec_sphv, f0, f1);
        else
            hipLaunchKernelGGL(ftgrid::pre3, grid, block, 0, st, X, g->Z, g->tw, N, M, g->plan_sphv, exec_sphv, f0, f1,
                               iv);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        rc = fft_ok(hipfftExecZ2Z(g->theta, (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_BACKWARD));
        if (rc != FTAMD_OK) return rc;
        if (pair && !g->disk) {
            hipLaunchKernelGGL(ftgrid::post3_pair, gridp, block, 0, st, g->Z, g->Y, N, M, g->plan_sphv, a, b, g->w,
                               g->w ? wmode : 0, tp);
            if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
            return fft_ok(hipfftExecZ2Z(phi, (hipfftDoubleComplex *)g->Y, (hipfftDoubleComplex *)X, HIPFFT_BACKWARD));
        }
        hipLaunchKernelGGL(ftgrid::post3, grid, block, 0, st, g->Z, g->Y, N, M, g->plan_sphv, a, b, iv, g->w,
                           g->w ? wmode : 0, tp);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        if (iv && (rc = iv_from_x()) != FTAMD_OK) return rc;
        if (pair) {  // the disk
            hipLaunchKernelGGL(ftgrid::pack_pairs, gridp, block, 0, st, g->Y, g->Yp, N, M);
            if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
            return fft_ok(hipfftExecZ2Z(phi, (hipfftDoubleComplex *)g->Yp, (hipfftDoubleComplex *)X, HIPFFT_BACKWARD));
        }
        return fft_ok(hipfftExecZ2D(phi, (hipfftDoubleComplex *)g->Y, X));
    }
    // r2c, data_c2r, scalings, DCT-II / DST-II (the disk: DCT-IV on the sine
    // class), fix-ups f0 and f1
    if (pair && !g->disk) {
        rc = fft_ok(hipfftExecZ2Z(phi, (hipfftDoubleComplex *)X, (hipfftDoubleComplex *)g->Y, HIPFFT_FORWARD));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::pre2_pair, grid, block, 0, st, g->Y, g->Z, N, M, g->plan_sphv, a, b, g->w,
                           g->w ? wmode : 0);
    } else if (pair) {  // the disk: unpacked into half spectra for pre2 and the DCT-IV
        rc = fft_ok(hipfftExecZ2Z(phi, (hipfftDoubleComplex *)X, (hipfftDoubleComplex *)g->Yp, HIPFFT_FORWARD));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::unpack_pairs, dim3((N + 255) / 256, M / 2 + 1), block, 0, st, g->Yp, g->Y, N, M);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        hipLaunchKernelGGL(ftgrid::pre2, grid, block, 0, st, g->Y, g->Z, N, M, g->plan_sphv, a, b, iv, g->w,
                           g->w ? wmode : 0);
    } else {
        rc = fft_ok(hipfftExecD2Z(phi, X, (hipfftDoubleComplex *)g->Y));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::pre2, grid, block, 0, st, g->Y, g->Z, N, M, g->plan_sphv, a, b, iv, g->w,
                           g->w ? wmode : 0);
    }
    if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
    if (iv && (rc = iv_from_y()) != FTAMD_OK) return rc;
    rc = fft_ok(hipfftExecZ2Z(g->theta, (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_FORWARD));
    if (rc != FTAMD_OK) return rc;
    hipLaunchKernelGGL(ftgrid::post2, grid, block, 0, st, g->Z, X, g->tw, N, M, g->plan_sphv, exec_sphv, f0, f1, iv);
    return hipGetLastError() == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}

// ---- triangle and tetrahedron grids

struct FtamdTriGrid {
    int rank, type3;
    int dims[3];         // FFTW's order: (N, M) or (N, L, M)
    int cN, cL;          // the column-major view's first sizes
    long long total;
    hipfftHandle ax[3];  // one batched FFT per axis
    double2 *tw[3];      // e^{i pi k / 2 len} per axis
    double2 *Z;          // total complex
};

static int twiddles(int L, double2 **out) {
    double2 *h = (double2 *)malloc(sizeof(double2) * (size_t)L);
    if (!h) return FTAMD_ENOMEM;
    for (int k = 0; k < L; k++) {
        const long double th = 3.141592653589793238462643383279502884L * k / (2.0L * L);
        h[k] = make_double2((double)cosl(th), (double)sinl(th));
    }
    int rc = hipMalloc((void **)out, sizeof(double2) * (size_t)L) == hipSuccess ? FTAMD_OK : FTAMD_ENOMEM;
    if (rc == FTAMD_OK && hipMemcpy(*out, h, sizeof(double2) * (size_t)L, hipMemcpyHostToDevice) != hipSuccess)
        rc = FTAMD_EDEVICE;
    free(h);
    return rc;
}

extern "C" int ftamd_boxgrid_create(int rank, const int *dims, int type3, FtamdTriGrid **out) {
    *out = NULL;
    if (rank != 2 && rank != 3) return FTAMD_EINVAL;
    long long total = 1;
    for (int a = 0; a < rank; a++) {
        if (dims[a] < 1) return FTAMD_EINVAL;
        total *= dims[a];
    }
    /* lines are the launch grid's y and z dimensions: at most 65535^2 per axis */
    if (total > 0x7fffffffLL) return FTAMD_EINVAL;
    FtamdTriGrid *g = (FtamdTriGrid *)calloc(1, sizeof(FtamdTriGrid));
    if (!g) return FTAMD_ENOMEM;
    g->rank = rank;
    g->type3 = type3;
    for (int a = 0; a < rank; a++) g->dims[a] = dims[a];
    g->cN = dims[0];
    g->cL = rank == 3 ? dims[1] : 1;
    g->total = total;
    int rc = hipMalloc((void **)&g->Z, sizeof(double2) * (size_t)total) == hipSuccess ? FTAMD_OK : FTAMD_ENOMEM;
    for (int a = 0; rc == FTAMD_OK && a < rank; a++) {
        const int len = dims[a];
        const int batch = (int)(total / len);
        rc = twiddles(len, &g->tw[a]);
        int l1[1] = {len};
        if (rc == FTAMD_OK) rc = fft_ok(hipfftPlanMany(&g->ax[a], 1, l1, l1, 1, len, l1, 1, len, HIPFFT_Z2Z, batch));
    }
    if (rc != FTAMD_OK) {
        ftamd_trigrid_destroy(g);
        return rc;
    }
    *out = g;
    return FTAMD_OK;
}

extern "C" int ftamd_trigrid_create(int N, int M, int type3, FtamdTriGrid **out) {
    const int dims[2] = {N, M};
    return ftamd_boxgrid_create(2, dims, type3, out);
}

extern "C" void ftamd_trigrid_destroy(FtamdTriGrid *g) {
    if (!g) return;
    hipDeviceSynchronize();
    for (int a = 0; a < 3; a++) {
        if (g->ax[a]) hipfftDestroy(g->ax[a]);
        hipFree(g->tw[a]);
    }
    hipFree(g->Z);
    free(g);
}

extern "C" int ftamd_boxgrid_run(FtamdTriGrid *g, double *X, int pre_fix, double pre_mul, double pre_div, int post_fix,
                                 double post_mul, double post_div, void *stream) {
    if (!g || !X) return FTAMD_EINVAL;
    const hipStream_t st = (hipStream_t)stream;
    ftgrid::AxisScale none = {0, 1.0, 0.0, 0, 1.0, 0.0, g->rank, g->cN, g->cL, 0, 0, 1.0, 1.0, 0};
    for (int a = g->rank - 1; a >= 0; a--) {
        // the last axis first: contiguous lines
        const int first = a == g->rank - 1, last = a == 0;
        const int len = g->dims[a];
        const int batch = (int)(g->total / len);
        long long stride = 1;
        for (int b = a + 1; b < g->rank; b++) stride *= g->dims[b];
        ftgrid::AxisScale sc = none;
        if (first) sc.pre_fix = pre_fix, sc.pre_mul = pre_mul, sc.pre_div = pre_div;
        if (last) sc.post_fix = post_fix, sc.post_mul = post_mul, sc.post_div = post_div;
        const dim3 block(256), grid((len + 255) / 256, batch < 65535 ? batch : 65535, (batch + 65534) / 65535);
        int rc = fft_ok(hipfftSetStream(g->ax[a], st));
        if (rc != FTAMD_OK) return rc;
        if (g->type3) {
            hipLaunchKernelGGL(ftgrid::axis_pre3, grid, block, 0, st, X, g->Z, g->tw[a], len, batch, stride, sc);
            if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
            rc = fft_ok(hipfftExecZ2Z(g->ax[a], (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z,
                                      HIPFFT_BACKWARD));
            if (rc != FTAMD_OK) return rc;
            hipLaunchKernelGGL(ftgrid::axis_post3, grid, block, 0, st, g->Z, X, len, batch, stride, sc);
        } else {
            hipLaunchKernelGGL(ftgrid::axis_pre2, grid, block, 0, st, X, g->Z, len, batch, stride, sc);
            if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
            rc = fft_ok(hipfftExecZ2Z(g->ax[a], (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z,
                                      HIPFFT_FORWARD));
            if (rc != FTAMD_OK) return rc;
            hipLaunchKernelGGL(ftgrid::axis_post2, grid, block, 0, st, g->Z, X, g->tw[a], len, batch, stride, sc);
        }
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
    }
    return FTAMD_OK;
}

extern "C" int ftamd_trigrid_run(FtamdTriGrid *g, double *X, int pre_fix, double pre_mul, double pre_div, int post_fix,
                                 double post_div, void *stream) {
    return ftamd_boxgrid_run(g, X, pre_fix, pre_mul, pre_div, post_fix, 1.0, post_div, stream);
}

// ---- rectangle-disk grids (fftw.c:657-740): the column transforms (x:
// DCT on the even columns, DST on the odd ones) and the row transforms (y:
// DCT along rows), each a pass of the axis kernels over the N x M
// column-major array, in the executor's order with its scalings.

struct FtamdRectGrid {
    int N, M, type3;
    hipfftHandle ax[2];  // [0] columns: length N, M lines; [1] rows: length M, N lines
    double2 *tw[2];
    double2 *Z;
};

extern "C" int ftamd_rectgrid_create(int N, int M, int type3, FtamdRectGrid **out) {
    *out = NULL;
    if (N < 1 || M < 1 || (long long)N * M > 0x7fffffffLL) return FTAMD_EINVAL;
    FtamdRectGrid *g = (FtamdRectGrid *)calloc(1, sizeof(FtamdRectGrid));
    if (!g) return FTAMD_ENOMEM;
    g->N = N;
    g->M = M;
    g->type3 = type3;
    int rc = hipMalloc((void **)&g->Z, sizeof(double2) * (size_t)N * M) == hipSuccess ? FTAMD_OK : FTAMD_ENOMEM;
    const int lens[2] = {N, M}, batches[2] = {M, N};
    for (int a = 0; rc == FTAMD_OK && a < 2; a++) {
        rc = twiddles(lens[a], &g->tw[a]);
        int l1[1] = {lens[a]};
        if (rc == FTAMD_OK)
            rc = fft_ok(hipfftPlanMany(&g->ax[a], 1, l1, l1, 1, lens[a], l1, 1, lens[a], HIPFFT_Z2Z, batches[a]));
    }
    if (rc != FTAMD_OK) {
        ftamd_rectgrid_destroy(g);
        return rc;
    }
    *out = g;
    return FTAMD_OK;
}

extern "C" void ftamd_rectgrid_destroy(FtamdRectGrid *g) {
    if (!g) return;
    hipDeviceSynchronize();
    for (int a = 0; a < 2; a++) {
        if (g->ax[a]) hipfftDestroy(g->ax[a]);
        hipFree(g->tw[a]);
    }
    hipFree(g->Z);
    free(g);
}

// One pass: axis 0 the columns (DST on the odd ones), 1 the rows; the masks
// and factors before and after it.
extern "C" int ftamd_rectgrid_pass(FtamdRectGrid *g, double *X, int axis, int pre_mask, double pre_f, int post_mask,
                                   double post_f, double post_mul, double post_div, void *stream) {
    if (!g || !X || (axis != 0 && axis != 1)) return FTAMD_EINVAL;
    const hipStream_t st = (hipStream_t)stream;
    const int N = g->N, M = g->M;
    const int len = axis == 0 ? N : M, batch = axis == 0 ? M : N;
    const long long stride = axis == 0 ? 1 : N;
    ftgrid::AxisScale sc = {0, 1.0, 0.0, 0, post_mul, post_div, 2, N, 1, pre_mask, post_mask, pre_f, post_f, axis == 0};
    // line b of the rows: e = (b / N) * M * N + b % N + t * N = b + t * N (b < N)
    const dim3 block(256), grid((len + 255) / 256, batch < 65535 ? batch : 65535, (batch + 65534) / 65535);
    int rc = fft_ok(hipfftSetStream(g->ax[axis], st));
    if (rc != FTAMD_OK) return rc;
    if (g->type3) {
        hipLaunchKernelGGL(ftgrid::axis_pre3, grid, block, 0, st, X, g->Z, g->tw[axis], len, batch, stride, sc);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        rc = fft_ok(hipfftExecZ2Z(g->ax[axis], (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_BACKWARD));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::axis_post3, grid, block, 0, st, g->Z, X, len, batch, stride, sc);
    } else {
        hipLaunchKernelGGL(ftgrid::axis_pre2, grid, block, 0, st, X, g->Z, len, batch, stride, sc);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        rc = fft_ok(hipfftExecZ2Z(g->ax[axis], (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_FORWARD));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::axis_post2, grid, block, 0, st, g->Z, X, g->tw[axis], len, batch, stride, sc);
    }
    return hipGetLastError() == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}

// ---- spin-weighted sphere grids (fftw.c:742-937): the sphere's theta
// transforms on the real and the imaginary part of every complex column
// (the column classes of the sphere for even spin, of the vector field for
// odd), the scaling a, the column permutation colswap (fftw.c:6-21: column 0,
// then 2j to j and 2j-1 to M-j) and a complex DFT of length M along the rows
// (backward for synthesis plans, forward for analysis plans).  Line 2j + p of
// Z is part p (0 re, 1 im) of column j.

namespace ftgrid {

__device__ __forceinline__ int swap_dst(int j, int M) { return j == 0 ? 0 : ((j & 1) ? M - ((j + 1) >> 1) : (j >> 1)); }

// synthesis plans: fix-ups, DCT-III / DST-III of both parts into Z
__global__ void cpre3(const double2 *__restrict__ X, double2 *__restrict__ Z, const double2 *__restrict__ tw, int N,
                      int M, int cls, double f0, double f1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int line = blockIdx.y;  // 2j + p
    if (k >= N || line >= 2 * M) return;
    const int j = line >> 1, p = line & 1;
    const double2 *x = X + (long long)N * j;
    const bool c = cos_class(j, cls);
    auto val = [&](int e) {
        double v = p ? x[e].y : x[e].x;
        if (c && e == 0) v *= f0;
        if (!c && e == N - 1) v *= f1;
        return v;
    };
    auto in = [&](int e) { return c ? val(e) : val(N - 1 - e); };
    const double re = in(k), im = k == 0 ? 0.0 : -in(N - k);
    const double2 t = tw[k];
    Z[(long long)N * line + k] = make_double2(t.x * re - t.y * im, t.y * re + t.x * im);
}

// after the inverse FFT: y, the sine class's signs, x a, into Yc at the colswap position
__global__ void cpost3(const double2 *__restrict__ Z, double2 *__restrict__ Yc, int N, int M, int cls, double a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int line = blockIdx.y;
    if (n >= N || line >= 2 * M) return;
    const int j = line >> 1, p = line & 1;
    const int src = (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
    double y = Z[(long long)N * line + src].x;
    if (!cos_class(j, cls) && (n & 1)) y = -y;
    y *= a;
    ((double *)Yc)[2 * ((long long)N * swap_dst(j, M) + n) + p] = y;
}

// analysis plans, after the row DFT: colswap_t, x a, the sine class's signs, Makhoul's reordering
__global__ void cpre2(const double2 *__restrict__ Yc, double2 *__restrict__ Z, int N, int M, int cls, double a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int line = blockIdx.y;
    if (n >= N || line >= 2 * M) return;
    const int j = line >> 1, p = line & 1;
    double x = ((const double *)Yc)[2 * ((long long)N * swap_dst(j, M) + n) + p];
    x *= a;
    if (!cos_class(j, cls) && (n & 1)) x = -x;
    const int dst = (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
    Z[(long long)N * line + dst] = make_double2(x, 0.0);
}

// after the forward FFT: DCT-II / DST-II of both parts, the fix-ups, into X
__global__ void cpost2(const double2 *__restrict__ Z, double2 *__restrict__ X, const double2 *__restrict__ tw, int N,
                       int M, int cls, double g0, double g1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int line = blockIdx.y;
    if (k >= N || line >= 2 * M) return;
    const int j = line >> 1, p = line & 1;
    const bool c = cos_class(j, cls);
    const int src = c ? k : N - 1 - k;
    const double2 V = Z[(long long)N * line + src], t = tw[src];
    double y = 2.0 * (t.x * V.x + t.y * V.y);
    if (c && k == 0) y *= g0;
    if (!c && k == N - 1) y *= g1;
    ((double *)X)[2 * ((long long)N * j + k) + p] = y;
}

}  // namespace ftgrid

struct FtamdSpinGrid {
    int N, M, type3, cls;
    hipfftHandle theta, phi;
    double2 *tw, *Z, *Yc;
};

extern "C" int ftamd_spingrid_create(int N, int M, int type3, int odd_spin, FtamdSpinGrid **out) {
    *out = NULL;
    /* lines (two per column) are the launch grid's y dimension */
    if (N < 1 || M < 1 || M % 2 == 0 || 2 * M > 65535) return FTAMD_EINVAL;
    FtamdSpinGrid *g = (FtamdSpinGrid *)calloc(1, sizeof(FtamdSpinGrid));
    if (!g) return FTAMD_ENOMEM;
    g->N = N;
    g->M = M;
    g->type3 = type3;
    g->cls = odd_spin ? 1 : 0;
    int rc = twiddles(N, &g->tw);
    if (rc == FTAMD_OK && hipMalloc((void **)&g->Z, sizeof(double2) * 2 * (size_t)N * M) != hipSuccess) rc = FTAMD_ENOMEM;
    if (rc == FTAMD_OK && hipMalloc((void **)&g->Yc, sizeof(double2) * (size_t)N * M) != hipSuccess) rc = FTAMD_ENOMEM;
    int n1[1] = {N}, n2[1] = {M};
    if (rc == FTAMD_OK) rc = fft_ok(hipfftPlanMany(&g->theta, 1, n1, n1, 1, N, n1, 1, N, HIPFFT_Z2Z, 2 * M));
    if (rc == FTAMD_OK) rc = fft_ok(hipfftPlanMany(&g->phi, 1, n2, n2, N, 1, n2, N, 1, HIPFFT_Z2Z, N));
    if (rc != FTAMD_OK) {
        ftamd_spingrid_destroy(g);
        return rc;
    }
    *out = g;
    return FTAMD_OK;
}

extern "C" void ftamd_spingrid_destroy(FtamdSpinGrid *g) {
    if (!g) return;
    hipDeviceSynchronize();
    if (g->theta) hipfftDestroy(g->theta);
    if (g->phi) hipfftDestroy(g->phi);
    hipFree(g->tw);
    hipFree(g->Z);
    hipFree(g->Yc);
    free(g);
}

extern "C" int ftamd_spingrid_run(FtamdSpinGrid *g, double *X, double f0, double f1, double a, void *stream) {
    if (!g || !X) return FTAMD_EINVAL;
    const int N = g->N, M = g->M;
    const hipStream_t st = (hipStream_t)stream;
    const dim3 block(256), grid((N + 255) / 256, 2 * M);
    double2 *Xc = (double2 *)X;
    int rc = fft_ok(hipfftSetStream(g->theta, st));
    if (rc == FTAMD_OK) rc = fft_ok(hipfftSetStream(g->phi, st));
    if (rc != FTAMD_OK) return rc;
    if (g->type3) {
        hipLaunchKernelGGL(ftgrid::cpre3, grid, block, 0, st, Xc, g->Z, g->tw, N, M, g->cls, f0, f1);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        rc = fft_ok(hipfftExecZ2Z(g->theta, (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_BACKWARD));
        if (rc != FTAMD_OK) return rc;
        hipLaunchKernelGGL(ftgrid::cpost3, grid, block, 0, st, g->Z, g->Yc, N, M, g->cls, a);
        if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
        return fft_ok(hipfftExecZ2Z(g->phi, (hipfftDoubleComplex *)g->Yc, (hipfftDoubleComplex *)Xc, HIPFFT_BACKWARD));
    }
    rc = fft_ok(hipfftExecZ2Z(g->phi, (hipfftDoubleComplex *)Xc, (hipfftDoubleComplex *)g->Yc, HIPFFT_FORWARD));
    if (rc != FTAMD_OK) return rc;
    hipLaunchKernelGGL(ftgrid::cpre2, grid, block, 0, st, g->Yc, g->Z, N, M, g->cls, a);
    if (hipGetLastError() != hipSuccess) return FTAMD_EDEVICE;
    rc = fft_ok(hipfftExecZ2Z(g->theta, (hipfftDoubleComplex *)g->Z, (hipfftDoubleComplex *)g->Z, HIPFFT_FORWARD));
    if (rc != FTAMD_OK) return rc;
    hipLaunchKernelGGL(ftgrid::cpost2, grid, block, 0, st, g->Z, Xc, g->tw, N, M, g->cls, f0, f1);
    return hipGetLastError() == hipSuccess ? FTAMD_OK : FTAMD_EDEVICE;
}
