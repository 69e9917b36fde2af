/*
 * fasttransforms_amd.h -- C ABI of the MI355X (gfx950) Givens-rotation stage.
 *
 * A drop-in for the rotation path of FastTransforms' src/fasttransforms.h:
 * the declarations below keep the reference's names, argument meaning and
 * column-major layouts (A[i + n*j]).  Every computation runs on the GPU
 * through hand-written HIP kernels; A may be a host pointer (staged through
 * device memory for the call) or a device pointer (updated in place).
 *
 * The reference functions return void and validate nothing.  These keep the
 * void signatures; invalid arguments leave A untouched and record an error
 * that ft_gpu_last_error() returns (valid arguments never change results).
 */
#ifndef FASTTRANSFORMS_AMD_H
#define FASTTRANSFORMS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* fasttransforms.h:298.  The first three members are the reference's
 * {double *s; double *c; int n;} (ftinternal.h:370-374), so RP->s, RP->c and
 * RP->n are the host FP64 tables; device copies follow them. */
typedef struct ft_rotation_plan_s ft_rotation_plan;

/* fasttransforms.h:11 -- a complex number as two doubles (re, im). */
typedef double ft_complex[2];

/* fasttransforms.h:336.  The first members are the reference's
 * {double *s1, *c1, *s2, *c2; int n, s;} (ftinternal.h:425-432): the host
 * FP64 tables in its layout; device state follows them. */
typedef struct ft_spin_rotation_plan_s ft_spin_rotation_plan;

/* fasttransforms.h:295 -- kept for link compatibility.  No computation here
 * uses host threads, so the value is recorded and otherwise ignored. */
void ft_set_num_threads(const int n);

/* fasttransforms.h:301 (rotations.c:6-10) -- also frees the device copies. */
void ft_destroy_rotation_plan(ft_rotation_plan *RP);

/* fasttransforms.h:303 (rotations.c:116-133) -- spherical harmonics; tables
 * bit-identical to the reference's. */
ft_rotation_plan *ft_plan_rotsphere(const int n);

/* fasttransforms.h:306, 308 (rotations.c:135-141) -- one vector with stride S,
 * from order m2 down to m1 / from m1 up to m2. */
void ft_kernel_sph_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);
void ft_kernel_sph_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);

/* fasttransforms.h:310 (rotations.c:143-160) -- Proriol (triangle). */
ft_rotation_plan *ft_plan_rottriangle(const int n, const double alpha, const double beta, const double gamma);

/* fasttransforms.h:313, 315 (rotations.c:162-168). */
void ft_kernel_tri_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);
void ft_kernel_tri_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);

/* fasttransforms.h:317 (rotations.c:173-190) -- Zernike (disk). */
ft_rotation_plan *ft_plan_rotdisk(const int n, const double alpha, const double beta);

/* fasttransforms.h:321, 323 (rotations.c:262-268). */
void ft_kernel_disk_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);
void ft_kernel_disk_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);

/* fasttransforms.h:318 (rotations.c:194-260) -- annulus with inner radius
 * rho (0 <= rho < 1): the Givens rotations of successive QR factorizations
 * of the shifted Jacobi operator of P^{(beta, alpha)}, built on the host in
 * FP64, in the disk's layout; run with the disk kernels and executors.  The
 * reference implements gamma = 0 only (and warns otherwise); here any other
 * gamma, or rho outside [0, 1), returns NULL with FT_GPU_EINVAL. */
ft_rotation_plan *ft_plan_rotannulus(const int n, const double alpha, const double beta, const double gamma,
                                     const double rho);

/* fasttransforms.h:325 (rotations.c:270-287) -- Dunkl-Xu (rectangle and
 * disk) with parameter beta; tables bit-identical to the reference's, in the
 * sphere's layout (n(n+1)/2 entries, s(l, m) at l + m(2n+1-m)/2). */
ft_rotation_plan *ft_plan_rotrectdisk(const int n, const double beta);

/* fasttransforms.h:328, 330 (rotations.c:289-295) -- the sphere's kernel
 * (rows k and k+2, k <= n-3-j) on a rectangle-disk plan. */
void ft_kernel_rectdisk_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);
void ft_kernel_rectdisk_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S);

/* fasttransforms.h:339, 341 (rotations.c:431-475) -- spin-weighted spherical
 * harmonics of spin s; tables bit-identical to the reference's. */
ft_spin_rotation_plan *ft_plan_rotspinsphere(const int n, const int s);
void ft_destroy_spin_rotation_plan(ft_spin_rotation_plan *SRP);

/* fasttransforms.h:344, 346 (rotations.c:477-483, rotations_default.c:87-117)
 * -- one complex vector of order m (any sign) with stride S (complex
 * entries): hi2lo the min(|m|,|s|) sequences of adjacent-row rotations
 * (s2, c2; their transposes where m*s < 0), then the sphere's sequences of
 * order |m|-|s| ... (s1, c1); lo2hi the transposes in reverse order. */
void ft_kernel_spinsph_hi2lo(const ft_spin_rotation_plan *SRP, const int m, ft_complex *A, const int S);
void ft_kernel_spinsph_lo2hi(const ft_spin_rotation_plan *SRP, const int m, ft_complex *A, const int S);

/* fasttransforms.h:332-333 (rotations.c:297-321) -- the tetrahedron's block
 * rotations of slab m: A is n x L (n = RP->n, L <= n), its columns l and l+1
 * rotated for j = m-1 .. 0 and l = L-2-j .. 0 with RP's s(l, j), c(l, j).
 * RP is a triangle plan. */
void ft_kernel_tet_hi2lo(const ft_rotation_plan *RP, const int L, const int m, double *A);
void ft_kernel_tet_lo2hi(const ft_rotation_plan *RP, const int L, const int m, double *A);

/* fasttransforms.h:349-350 (drivers.c:58-92) -- spherical harmonics.
 * A is n x M (M odd, M <= 2n-1), updated in place; B is the caller's n x M
 * scratch, unused here (the GPU path needs no host scratch). */
void ft_execute_sph_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M);
void ft_execute_sph_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M);

/* fasttransforms.h:352-353 (drivers.c:168-220) -- spherical vector fields
 * on a sphere plan: order m (2 <= m <= M/2-1) in columns 2m+1 and 2m+2.  A is
 * n x M, M odd, M <= 2n+1. */
void ft_execute_sphv_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M);
void ft_execute_sphv_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M);

/* fasttransforms.h:355-356 (drivers.c:278-312) -- A is n x M, M <= n. */
void ft_execute_tri_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M);
void ft_execute_tri_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M);

/* fasttransforms.h:358-359 (drivers.c:378-412) -- A is n x M, M odd, M <= 4n-3. */
void ft_execute_disk_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M);
void ft_execute_disk_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M);

/* fasttransforms.h:361-362 (drivers.c:488-508) -- rectangle and disk on a
 * rectangle-disk plan: order m (2 <= m <= M-1) in column m.  A is n x M,
 * M <= n. */
void ft_execute_rectdisk_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M);
void ft_execute_rectdisk_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M);

/* fasttransforms.h:367-368 (drivers.c:645-705) -- spin-weighted spherical
 * harmonics: A is n x M complex (column 0 order 0, columns 2m-1 and 2m
 * orders -m and m), M odd, M <= 2n-1; B is unused. */
void ft_execute_spinsph_hi2lo(const ft_spin_rotation_plan *SRP, ft_complex *A, ft_complex *B, const int M);
void ft_execute_spinsph_lo2hi(const ft_spin_rotation_plan *SRP, ft_complex *A, ft_complex *B, const int M);

/* fasttransforms.h:364-365 (drivers.c:510-533) -- Proriol^3 (tetrahedron):
 * A is N x L x M (slab m = the N x L block at N*L*m), RP1 and RP2 triangle
 * plans of size N (the reference's are rottriangle(N, alpha, beta,
 * gamma+delta+1) and rottriangle(N, beta, gamma, delta)).  hi2lo: in every
 * slab m, the triangle's rotations of order l+m on its column l (l < L-m)
 * with RP1, then the block rotations of ft_kernel_tet_hi2lo with RP2; lo2hi
 * the transposes in reverse order.  L <= N; other shapes set FT_GPU_EINVAL. */
void ft_execute_tet_hi2lo(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, double *A, const int L,
                          const int M);
void ft_execute_tet_lo2hi(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, double *A, const int L,
                          const int M);

/* fasttransforms.h:371-385 -- the harmonic plan: rotation plans, the 1D
 * connection matrices P and their inverses Pinv (host FP64, n x n
 * column-major, upper triangular) and parameters; the reference's members
 * first, then this library's device state.  ft_modified_plan is opaque here
 * (only ft_plan_ann2cxf builds them; every other plan has MP NULL, NMP 0),
 * and B, the reference's host scratch, is NULL: the device path needs
 * none. */
typedef struct ft_modified_plan_s ft_modified_plan;
typedef struct {
    ft_rotation_plan **RP;
    ft_modified_plan **MP;
    double *B;
    double **P;
    double **Pinv;
    double alpha;
    double beta;
    double gamma;
    double delta;
    double rho;
    int NRP;
    int NMP;
    int NP;
    struct ftamd_harmonic_dev *dev; /* extension: device copies of P and Pinv, work buffers */
} ft_harmonic_plan;

/* fasttransforms.h:388 (drivers.c:793-814) -- also frees the device state. */
void ft_destroy_harmonic_plan(ft_harmonic_plan *P);

/* fasttransforms.h:391 (drivers.c:816-831) -- spherical harmonics <->
 * bivariate Fourier: the sphere's rotation plan, P = {normalized Legendre ->
 * Chebyshev, normalized ultraspherical(3/2) -> ultraspherical(1)} and their
 * inverses (transforms.c:50-93, built on the host in long double as the
 * reference does). */
ft_harmonic_plan *ft_plan_sph2fourier(const int n);

/* fasttransforms.h:394-399 (drivers.c:833-899) -- A is N x M (N the plan's
 * n, M odd, M <= 2N-1; M <= 2N+1 for the vector-field pair), updated in
 * place, host or device memory.  TRANS 'N' applies the transform, 'T' its
 * transpose.  sph2fourier 'N' is the sphere's hi2lo rotations then, on the
 * columns j = 0, 3 mod 4, P[0] and on j = 1, 2 mod 4, P[1] (upper triangular
 * products); fourier2sph inverts it with Pinv and lo2hi.  The vector-field
 * pair swaps the two matrices' column classes and runs the sphv
 * rotations. */
void ft_execute_sph2fourier(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_fourier2sph(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_sphv2fourier(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_fourier2sphv(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);

/* fasttransforms.h:402-407 (drivers.c:901-957) -- Proriol <-> bivariate
 * Chebyshev: the triangle's rotation plan, P = {P^{(beta+gamma+1, alpha)},
 * P^{(gamma, beta)} -> Chebyshev} (normalized Jacobi connections) and their
 * inverses.  A is N x M (M <= N); tri2cheb 'N' is hi2lo, P[0] from the left
 * on every column, P[1]^T from the right on every row (each skipped where it
 * is the identity, as the reference does) and the 2D Chebyshev
 * normalization. */
ft_harmonic_plan *ft_plan_tri2cheb(const int n, const double alpha, const double beta, const double gamma);
void ft_execute_tri2cheb(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_cheb2tri(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);

/* fasttransforms.h:426-431 (drivers.c:1123-1171) -- Dunkl-Xu (rectangle and
 * disk) <-> bivariate Chebyshev: the rectangle-disk rotation plan, P =
 * {ultraspherical(beta+1) -> Chebyshev, ultraspherical(beta+2) ->
 * ultraspherical(1), ultraspherical(beta+1/2) -> Chebyshev} and their
 * inverses.  A is N x M (M <= N); rectdisk2cheb 'N' is hi2lo, P[0] and P[1]
 * from the left on the even and odd columns, P[2]^T from the right. */
ft_harmonic_plan *ft_plan_rectdisk2cheb(const int n, const double beta);
void ft_execute_rectdisk2cheb(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_cheb2rectdisk(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);

/* fasttransforms.h:410-415 (drivers.c:959-1021) -- generalized Zernike <->
 * Chebyshev x Fourier: the disk's rotation plan, P[0] = Jacobi(beta, alpha)
 * -> Chebyshev and P[1] = Jacobi(beta, alpha+1) -> Jacobi(-1/2, 1/2), scaled
 * as the reference does, and their inverses; the column classes as for
 * sph2fourier.  A is N x M (M odd, M <= 4N-3). */
ft_harmonic_plan *ft_plan_disk2cxf(const int n, const double alpha, const double beta);
void ft_execute_disk2cxf(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_cxf2disk(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);

/* fasttransforms.h:417-423 (drivers.c:1023-1121) -- annulus polynomials
 * (inner radius rho, gamma = 0 only, as the reference) <-> Chebyshev x
 * Fourier: the annulus rotation plan; MP[0] the identity (columns j = 0, 3
 * mod 4) and MP[1] the upper bidiagonal Cholesky factor R of t I + J, J the
 * normalized Jacobi(beta, alpha) multiplication operator and t = 1 +
 * 2 rho^2 / (1 - rho^2) (columns j = 1, 2 mod 4); P[0] = Jacobi(beta, alpha)
 * -> Chebyshev and Pinv[0] its inverse, scaled by sqrt(2) (2 / (1 -
 * rho^2))^((alpha + beta + 1) / 2) and its reciprocal.  A is N x M (M odd,
 * M <= 4N-3).  ann2cxf 'N' is hi2lo, R^{-1} on the columns j = 1, 2 mod 4,
 * then P[0] from the left on every column; 'T' the transposes in reverse
 * order.  cxf2ann 'N' is Pinv[0], R, then lo2hi.  gamma != 0 or rho outside
 * [0, 1) sets FT_GPU_EINVAL (the reference warns and builds an unusable
 * plan). */
ft_harmonic_plan *ft_plan_ann2cxf(const int n, const double alpha, const double beta, const double gamma,
                                  const double rho);
void ft_execute_ann2cxf(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);
void ft_execute_cxf2ann(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M);

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference).
 * ---------------------------------------------------------------------- */

/* Error side channel.  Codes are FT_GPU_*; reading clears the calling
 * thread's last error. */
enum {
    FT_GPU_OK = 0,
    FT_GPU_EINVAL = 1,     /* bad plan, n, M, m1/m2 or stride */
    FT_GPU_ENOMEM = 2,     /* device allocation failed */
    FT_GPU_EDEVICE = 3,    /* no usable HIP device / HIP runtime error */
    FT_GPU_ETIMEOUT = 4,   /* a kernel's bounded wait expired (results invalid) */
    FT_GPU_EPTR = 5,       /* pointer not host or device memory of the current device */
    FT_GPU_EINCOMPLETE = 6 /* a launch ended with work not done (results invalid) */
};
int ft_gpu_last_error(void);
const char *ft_gpu_error_string(int code);

/* Precision of the device path: 0 = FP64 (default), 1 = FP32 storage and
 * arithmetic with s/c rounded from the FP64 plan. */
enum { FT_GPU_FP64 = 0, FT_GPU_FP32 = 1 };

/* Executor kinds of the device-resident calls: the whole-matrix executors
 * ft_execute_{sph,tri,disk,sphv,rectdisk}_*.  A sphv call takes a sphere
 * plan, a rectdisk call a rectangle-disk plan (any plan of the sphere's table
 * layout runs, as in the reference). */
enum { FT_GPU_SPH = 0, FT_GPU_TRI = 1, FT_GPU_DISK = 2, FT_GPU_SPHV = 3, FT_GPU_RECTDISK = 4 };

/* Device-resident execution on a HIP stream (stream = NULL: the null stream).
 * kind: FT_GPU_SPH .. FT_GPU_RECTDISK; direction: 0 hi2lo, 1 lo2hi.
 * dA is n x M column-major in device memory (double* for FP64, float* for
 * FP32).  Asynchronous: returns after the launch; 0 on success or an FT_GPU_*
 * code.  Launches with the same (plan, M, direction, precision) share a
 * schedule and its device workspace: order them on one stream (or wait for
 * one before the next).  The first call for such a combination builds and
 * uploads the schedule (hipMalloc, synchronous copies), so make one call
 * outside any HIP-graph capture before capturing. */
int ft_gpu_execute(const ft_rotation_plan *RP, int kind, int direction, int precision, void *dA, int M,
                   void *stream);

/* ft_execute_tet_{hi2lo,lo2hi} (direction 0, 1) on a device-resident FP64
 * N x L x M array dA, asynchronously on a HIP stream, as ft_gpu_execute (the
 * same rules for shared schedules and the first call).  Two launches of the
 * band-systolic engine: the slab columns' rotations and the block rotations. */
int ft_gpu_execute_tet(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, int direction, double *dA, int L,
                       int M, void *stream);

/* fasttransforms.h:466-493 (fftw.c:62-270) -- synthesis (coefficients of
 * the bivariate Fourier series sph2fourier produces -> values on the
 * N x M grid theta_i = (i+1/2) pi / N, phi_j = 2 pi j / M) and analysis (the
 * inverse) on the sphere, and their vector-field variants, on the GPU
 * (hipFFT plus this library's kernels in place of FFTW).  The plan holds
 * device work buffers (about 24 N M bytes); flags are ignored.  X is N x M
 * column-major, host or device, M odd, overwritten in place.  Synthesis 'N'
 * and analysis 'T' take a synthesis plan, synthesis 'T' and analysis 'N' an
 * analysis plan (as the reference's own tests pair them); another plan, or
 * N, M other than the plan's, sets FT_GPU_EINVAL.  Results agree with
 * FFTW's to rounding (the FFTs sum in another order). */
typedef struct ft_sphere_fftw_plan_s ft_sphere_fftw_plan;
void ft_destroy_sphere_fftw_plan(ft_sphere_fftw_plan *P);
ft_sphere_fftw_plan *ft_plan_sph_synthesis(const int N, const int M, const unsigned flags);
ft_sphere_fftw_plan *ft_plan_sph_analysis(const int N, const int M, const unsigned flags);
ft_sphere_fftw_plan *ft_plan_sphv_synthesis(const int N, const int M, const unsigned flags);
ft_sphere_fftw_plan *ft_plan_sphv_analysis(const int N, const int M, const unsigned flags);
void ft_execute_sph_synthesis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M);
void ft_execute_sph_analysis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M);
void ft_execute_sphv_synthesis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M);
void ft_execute_sphv_analysis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M);

/* fasttransforms.h:497-512 (fftw.c:275-330) -- synthesis and analysis on the
 * triangle's grid: FFTW's two-dimensional REDFT01 (synthesis plans) or
 * REDFT10 (analysis plans) over X read, as the reference's r2r_2d(N, M) reads
 * it, as an N x M row-major buffer (for N = M, as the reference's tests use,
 * the column-major array's columns and rows), with the executors' scalings;
 * nothing happens unless N > 1 and M > 1.  Any plan of size N x M runs with
 * either executor and TRANS (the plan fixes the kinds). */
typedef struct ft_triangle_fftw_plan_s ft_triangle_fftw_plan;
void ft_destroy_triangle_fftw_plan(ft_triangle_fftw_plan *P);
ft_triangle_fftw_plan *ft_plan_tri_synthesis(const int N, const int M, const unsigned flags);
ft_triangle_fftw_plan *ft_plan_tri_analysis(const int N, const int M, const unsigned flags);
void ft_execute_tri_synthesis(const char TRANS, const ft_triangle_fftw_plan *P, double *X, const int N, const int M);
void ft_execute_tri_analysis(const char TRANS, const ft_triangle_fftw_plan *P, double *X, const int N, const int M);
/* The triangle's executors on a device-resident X on a HIP stream
 * (analysis 0: synthesis, 1: analysis). */
int ft_gpu_execute_tri_grid(int analysis, char TRANS, const ft_triangle_fftw_plan *P, double *dX, int N, int M,
                            void *stream);

/* fasttransforms.h:542-556 (fftw.c:400-520) -- synthesis and analysis on
 * the disk's grid (coefficients of the Chebyshev x Fourier series disk2cxf
 * produces): REDFT01 / REDFT10 on columns 0 and 3 mod 4, REDFT11 on columns
 * 1 and 2 mod 4, the phi DFT as on the sphere; the same rules for plans,
 * TRANS and M as the sphere's executors.  ft_disk_fftw_plan is the sphere's
 * plan type; a disk executor takes a disk plan only, and a sphere executor
 * a sphere plan only. */
typedef ft_sphere_fftw_plan ft_disk_fftw_plan;
void ft_destroy_disk_fftw_plan(ft_disk_fftw_plan *P);
ft_disk_fftw_plan *ft_plan_disk_synthesis(const int N, const int M, const unsigned flags);
ft_disk_fftw_plan *ft_plan_disk_analysis(const int N, const int M, const unsigned flags);
void ft_execute_disk_synthesis(const char TRANS, const ft_disk_fftw_plan *P, double *X, const int N, const int M);
void ft_execute_disk_analysis(const char TRANS, const ft_disk_fftw_plan *P, double *X, const int N, const int M);

/* fasttransforms.h:514-527 (fftw.c:333-395) -- synthesis and analysis on the
 * tetrahedron's grid: FFTW's three-dimensional REDFT01 (synthesis plans) or
 * REDFT10 (analysis plans) over X read, as the reference's r2r_3d(N, L, M)
 * reads it, as a row-major N x L x M buffer, with the executors' scalings on
 * the faces of its column-major view; nothing unless N, L and M exceed 1. */
typedef struct ft_tetrahedron_fftw_plan_s ft_tetrahedron_fftw_plan;
void ft_destroy_tetrahedron_fftw_plan(ft_tetrahedron_fftw_plan *P);
ft_tetrahedron_fftw_plan *ft_plan_tet_synthesis(const int N, const int L, const int M, const unsigned flags);
ft_tetrahedron_fftw_plan *ft_plan_tet_analysis(const int N, const int L, const int M, const unsigned flags);
void ft_execute_tet_synthesis(const char TRANS, const ft_tetrahedron_fftw_plan *P, double *X, const int N, const int L,
                              const int M);
void ft_execute_tet_analysis(const char TRANS, const ft_tetrahedron_fftw_plan *P, double *X, const int N, const int L,
                             const int M);
/* The tetrahedron's executors on a device-resident X on a HIP stream
 * (analysis 0: synthesis, 1: analysis). */
int ft_gpu_execute_tet_grid(int analysis, char TRANS, const ft_tetrahedron_fftw_plan *P, double *dX, int N, int L,
                            int M, void *stream);

/* fasttransforms.h:558-575 (fftw.c:530-655) -- synthesis and analysis on the
 * annulus's grid (inner radius rho in [0, 1)): REDFT01 / REDFT10 on every
 * column, the radial weights sqrt(2 (sin^2 t + rho^2 cos^2 t) / (1 - rho^2)),
 * t = (N - n - 1/2) pi / 2N, on columns 1 and 2 mod 4 (multiplied by
 * synthesis, divided by analysis), the phi DFT as on the sphere.
 * ft_annulus_fftw_plan is the sphere's plan type; an annulus executor takes
 * an annulus plan only. */
typedef ft_sphere_fftw_plan ft_annulus_fftw_plan;
void ft_destroy_annulus_fftw_plan(ft_annulus_fftw_plan *P);
double ft_get_rho_annulus_fftw_plan(ft_annulus_fftw_plan *P);
ft_annulus_fftw_plan *ft_plan_annulus_synthesis(const int N, const int M, const double rho, const unsigned flags);
ft_annulus_fftw_plan *ft_plan_annulus_analysis(const int N, const int M, const double rho, const unsigned flags);
void ft_execute_annulus_synthesis(const char TRANS, const ft_annulus_fftw_plan *P, double *X, const int N,
                                  const int M);
void ft_execute_annulus_analysis(const char TRANS, const ft_annulus_fftw_plan *P, double *X, const int N, const int M);

/* fasttransforms.h:577-590 (fftw.c:657-740) -- synthesis and analysis on the
 * rectangle-disk's grid: along the columns REDFT01 (even columns) and
 * RODFT01 (odd columns), along the rows REDFT01 (synthesis plans; REDFT10,
 * RODFT10, REDFT10 for analysis plans), in the executors' order with their
 * scalings.  Any plan of size N x M runs with either executor and TRANS. */
typedef struct ft_rectdisk_fftw_plan_s ft_rectdisk_fftw_plan;
void ft_destroy_rectdisk_fftw_plan(ft_rectdisk_fftw_plan *P);
ft_rectdisk_fftw_plan *ft_plan_rectdisk_synthesis(const int N, const int M, const unsigned flags);
ft_rectdisk_fftw_plan *ft_plan_rectdisk_analysis(const int N, const int M, const unsigned flags);
void ft_execute_rectdisk_synthesis(const char TRANS, const ft_rectdisk_fftw_plan *P, double *X, const int N,
                                   const int M);
void ft_execute_rectdisk_analysis(const char TRANS, const ft_rectdisk_fftw_plan *P, double *X, const int N,
                                  const int M);
/* The rectangle-disk's executors on a device-resident X on a HIP stream
 * (analysis 0: synthesis, 1: analysis). */
int ft_gpu_execute_rectdisk_grid(int analysis, char TRANS, const ft_rectdisk_fftw_plan *P, double *dX, int N, int M,
                                 void *stream);

/* fasttransforms.h:592-617 (fftw.c:742-937) -- synthesis and analysis of
 * spin-weighted functions on the sphere's grid: X is N x M complex (M odd);
 * the theta transforms of the sphere (even spin S) or of the vector field
 * (odd S) on the real and imaginary parts, the column permutation of the
 * reference's colswap and a complex DFT along the rows (synthesis backward,
 * analysis forward).  TRANS is 'N' or 'C'; synthesis 'N' and analysis 'C'
 * take a synthesis plan, synthesis 'C' and analysis 'N' an analysis plan. */
typedef struct ft_spinsphere_fftw_plan_s ft_spinsphere_fftw_plan;
void ft_destroy_spinsphere_fftw_plan(ft_spinsphere_fftw_plan *P);
int ft_get_spin_spinsphere_fftw_plan(const ft_spinsphere_fftw_plan *P);
ft_spinsphere_fftw_plan *ft_plan_spinsph_synthesis(const int N, const int M, const int S, const unsigned flags);
ft_spinsphere_fftw_plan *ft_plan_spinsph_analysis(const int N, const int M, const int S, const unsigned flags);
void ft_execute_spinsph_synthesis(const char TRANS, const ft_spinsphere_fftw_plan *P, ft_complex *X, const int N,
                                  const int M);
void ft_execute_spinsph_analysis(const char TRANS, const ft_spinsphere_fftw_plan *P, ft_complex *X, const int N,
                                 const int M);
/* The spin-weighted sphere's grid executors on a device-resident complex X
 * on a HIP stream (analysis 0: synthesis, 1: analysis). */
int ft_gpu_execute_spinsph_grid(int analysis, char TRANS, const ft_spinsphere_fftw_plan *P, void *dX, int N, int M,
                                void *stream);

/* The grid transforms on a device-resident X asynchronously on a HIP stream:
 * which = 0 sph synthesis, 1 sph analysis, 2 sphv synthesis, 3 sphv
 * analysis, 4 disk synthesis, 5 disk analysis, 6 annulus synthesis, 7
 * annulus analysis, TRANS and the plan as the executors above.  Calls on one plan share its work buffers: order them on
 * one stream. */
int ft_gpu_execute_sph_grid(int which, char TRANS, const ft_sphere_fftw_plan *P, double *dX, int N, int M,
                            void *stream);

/* ft_execute_spinsph_{hi2lo,lo2hi} (direction 0, 1) on a device-resident
 * n x M complex array dA (interleaved FP64), asynchronously on a HIP stream,
 * as ft_gpu_execute: the first stage's kernel and one band-systolic launch. */
int ft_gpu_execute_spinsph(const ft_spin_rotation_plan *SRP, int direction, void *dA, int M, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP-graph-capture discipline, tight serving loops,
 * harnesses that forbid hipMalloc): ft_gpu_execute on the band-systolic
 * engine with its work list and error word in a fixed device scratch array of
 * the library (256 KiB per device, held until the process ends) instead of a
 * schedule it allocates; it calls no hipMalloc or hipFree.  The plan's
 * band-systolic tables for this direction and precision must already be
 * resident on the current device: one earlier ft_gpu_execute (or reference
 * executor) call on the plan with the same direction and precision that ran
 * on that engine leaves them so.  Returns FT_GPU_EINVAL where they are not,
 * where the engine does not apply (beyond sph n = 1024, tri and disk n = 512
 * in FP64; twice that in FP32), or where the work list would exceed the
 * scratch.  Asynchronous like ft_gpu_execute (errors through
 * ft_gpu_synchronize).  All calls on one device share its scratch, which
 * keeps the last call's contents: calls that overlap on one device (several
 * streams or threads) race. */
int ft_gpu_execute_noalloc(const ft_rotation_plan *RP, int kind, int direction, int precision, void *dA, int M,
                           void *stream);

/* The harmonic transforms on a device-resident A (dA, N x M FP64) on a HIP
 * stream, asynchronously: transform FT_GPU_SPH2FOURIER .. FT_GPU_CXF2ANN,
 * TRANS 'N' or 'T', as the executors above.
 * Errors of the launches come back through ft_gpu_synchronize.  Calls on
 * one harmonic plan share its device work buffer: order them on one
 * stream. */
enum {
    FT_GPU_SPH2FOURIER = 0,
    FT_GPU_FOURIER2SPH = 1,
    FT_GPU_SPHV2FOURIER = 2,
    FT_GPU_FOURIER2SPHV = 3,
    FT_GPU_DISK2CXF = 4,
    FT_GPU_CXF2DISK = 5,
    FT_GPU_TRI2CHEB = 6,
    FT_GPU_CHEB2TRI = 7,
    FT_GPU_RECTDISK2CHEB = 8,
    FT_GPU_CHEB2RECTDISK = 9,
    FT_GPU_ANN2CXF = 10,
    FT_GPU_CXF2ANN = 11
};
int ft_gpu_execute_harmonic(int transform, char TRANS, const ft_harmonic_plan *P, void *dA, int N, int M,
                            void *stream);

/* Waits for `stream` and reads the device error words of every asynchronous
 * launch (ft_gpu_execute, ft_gpu_execute_batch, ft_gpu_execute_batch_plan)
 * made on it since the last such check, clearing them: FT_GPU_OK, or
 * FT_GPU_ETIMEOUT / FT_GPU_EINCOMPLETE if one of those launches ended with
 * its results invalid (or a HIP error code).  No reference counterpart: the
 * reference has no asynchronous calls. */
int ft_gpu_synchronize(void *stream);

/* Batched transforms of uneven degree in one launch.  plans[i] may differ in
 * n (plans of one constructor and one set of parameters); dA[i] is
 * plans[i]->n x M[i] on the device.
 * Every transform runs as one persistent, cost-ordered launch.  The schedule
 * of the last argument list is cached; a call with another list replaces it,
 * so a HIP graph that captured this call is valid only until then (use a
 * batch plan below for graphs). */
int ft_gpu_execute_batch(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                         void *const *dA, const int *M, int count, void *stream);

/* A batch plan: the schedule of one argument list of ft_gpu_execute_batch,
 * owned by the caller (for example to launch it inside a captured HIP graph,
 * which stays valid until ft_gpu_destroy_batch).  NULL on error
 * (ft_gpu_last_error tells which).  It uses the plans' device tables: destroy
 * it before any of its plans. */
typedef struct ft_gpu_batch_s ft_gpu_batch;
ft_gpu_batch *ft_gpu_plan_batch(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                                void *const *dA, const int *M, int count);
int ft_gpu_execute_batch_plan(const ft_gpu_batch *B, void *stream);
void ft_gpu_destroy_batch(ft_gpu_batch *B);

/* Tuning knobs of the pipelined kernel: sequences per wave (L), stage waves
 * per workgroup (W), workgroups per CU.  0 keeps the tuned (or built-in)
 * choice. */
int ft_gpu_set_tuning(int L, int W, int wg_per_cu);

/* Autotuning table (no reference counterpart; the reference picks its SIMD
 * kernel by CPU feature alone, src/drivers.c:58-110).  One row of
 * FT_GPU_TUNE_COLS ints per (kind, precision, direction, batch, degree
 * range): kind, precision, direction, batch (0 single transform, 1 batch),
 * nlo, nhi, engine (FT_GPU_ENGINE_PIPELINE or _SYSTOLIC), band-systolic
 * bands per chain part (0: built-in), pipelined L and W (0: built-in),
 * pipelined workgroups per CU (0: built-in).  The library carries the table
 * fasttransforms_amd/tune.py measured on gfx950; knobs set with
 * ft_gpu_set_engine, ft_gpu_set_systolic and ft_gpu_set_tuning override it.
 * ft_gpu_set_tuning_table installs `count` rows (count 0: the compiled
 * table again, -1: no table, the built-in choices); FT_GPU_EINVAL for an
 * invalid row.  ft_gpu_tuning_table copies up to `cap` rows of the active
 * table and returns how many it has.  ft_gpu_tuned_choice gives the engine,
 * bands per part, L, W and workgroups per CU a launch of that shape uses
 * now (knobs, then table, then built-in). */
#define FT_GPU_TUNE_COLS 11
int ft_gpu_set_tuning_table(const int *rows, int count);
int ft_gpu_tuning_table(int *rows, int cap);
int ft_gpu_tuned_choice(int kind, int precision, int direction, int batch, int n, int *choice);

/* Engine of the whole-matrix executors (no reference counterpart; for tests
 * and measurements): FT_GPU_ENGINE_AUTO (default) runs the band-systolic
 * engine wherever the build has it for the precision, the pipelined engine
 * elsewhere; FT_GPU_ENGINE_PIPELINE and FT_GPU_ENGINE_SYSTOLIC force one (the
 * latter fails with FT_GPU_EINVAL where it does not apply).  Batches run the
 * pipelined engine unless the systolic one is forced; single-vector kernels
 * always run the pipelined engine. */
enum { FT_GPU_ENGINE_AUTO = 0, FT_GPU_ENGINE_PIPELINE = 1, FT_GPU_ENGINE_SYSTOLIC = 2 };
int ft_gpu_set_engine(int engine);

/* Band-systolic tuning knob (no reference counterpart): most bands of a
 * chain part (1..16).  A chain of more bands runs in parts on several
 * workgroups of one XCD.  0 keeps the built-in choice.  Schedules are cached
 * per value; results are bitwise the same for every value. */
int ft_gpu_set_systolic(int maxpart);

/* Polls after which a band-systolic wait between bands gives up and the
 * launch reports FT_GPU_ETIMEOUT (no reference counterpart; for tests of the
 * failure path).  0 restores the default (2^22 polls, seconds). */
int ft_gpu_set_spin_limit(int limit);

/* Engine (FT_GPU_ENGINE_PIPELINE or _SYSTOLIC) of the calling thread's last
 * whole-matrix or batch launch, 0 before any (no reference counterpart;
 * diagnostics). */
int ft_gpu_last_engine(void);

#ifdef __cplusplus
}
#endif

#endif /* FASTTRANSFORMS_AMD_H */
