/*
 * ref_exec.c -- harness around the reference's own compiled kernels.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Built by oracle/Makefile into
 * oracle/_ref/libftref.so together with the reference files compiled where
 * they lie (src/rotations/rotations_{default,SSE2,AVX,AVX_FMA,AVX512F}.c,
 * src/permute.c, src/permute/permute_{default,SSE,SSE2,AVX,AVX512F}.c).
 *
 * src/drivers.c and src/rotations.c cannot be compiled here: both include
 * fasttransforms.h, which includes <cblas.h> and <fftw3.h>, and this image has
 * neither.  So this file restates only the executor glue of drivers.c -- the
 * ISA dispatch (drivers.c:58-92) and the EXECUTE_SPH / EXECUTE_TRI /
 * EXECUTE_DISK sequences (drivers.c:112-126, 328-336, 432-446): warp, permute
 * into B, scalar head, OpenMP loop over S-column blocks, permute back, unwarp.
 * Every arithmetic operation happens inside the reference's compiled kernels
 * (kernel_*_{default,SSE2,AVX,AVX_FMA,AVX512F}); the tables come from the
 * caller (the oracle's restatement of rotations.c:113-190).
 */
#include <omp.h>
#include "ftinternal.h"

typedef void (*ref_kernel)(const ft_rotation_plan *, const int, const int, double *, const int);

enum { REF_DEFAULT = 0, REF_SSE2 = 1, REF_AVX = 2, REF_AVX_FMA = 3, REF_AVX512F = 4 };

/* The dispatch order of drivers.c:58-74. */
int ref_select_isa(void) {
    ft_simd simd = get_simd();
    if (simd.avx512f) return REF_AVX512F;
    if (simd.avx) return simd.fma ? REF_AVX_FMA : REF_AVX;
    if (simd.sse2) return REF_SSE2;
    return REF_DEFAULT;
}

static int isa_width(int isa) {
    switch (isa) {
        case REF_SSE2: return 2;
        case REF_AVX: case REF_AVX_FMA: return 4;
        case REF_AVX512F: return 8;
        default: return 1;
    }
}

/* kind: 0 sphere, 1 triangle, 2 disk; dir: 0 hi2lo, 1 lo2hi. */
static ref_kernel pick(int kind, int dir, int isa) {
    static const ref_kernel table[3][2][5] = {
        {{kernel_sph_hi2lo_default, kernel_sph_hi2lo_SSE2, kernel_sph_hi2lo_AVX, kernel_sph_hi2lo_AVX_FMA, kernel_sph_hi2lo_AVX512F},
         {kernel_sph_lo2hi_default, kernel_sph_lo2hi_SSE2, kernel_sph_lo2hi_AVX, kernel_sph_lo2hi_AVX_FMA, kernel_sph_lo2hi_AVX512F}},
        {{kernel_tri_hi2lo_default, kernel_tri_hi2lo_SSE2, kernel_tri_hi2lo_AVX, kernel_tri_hi2lo_AVX_FMA, kernel_tri_hi2lo_AVX512F},
         {kernel_tri_lo2hi_default, kernel_tri_lo2hi_SSE2, kernel_tri_lo2hi_AVX, kernel_tri_lo2hi_AVX_FMA, kernel_tri_lo2hi_AVX512F}},
        {{kernel_disk_hi2lo_default, kernel_disk_hi2lo_SSE2, kernel_disk_hi2lo_AVX, kernel_disk_hi2lo_AVX_FMA, kernel_disk_hi2lo_AVX512F},
         {kernel_disk_lo2hi_default, kernel_disk_lo2hi_SSE2, kernel_disk_lo2hi_AVX, kernel_disk_lo2hi_AVX_FMA, kernel_disk_lo2hi_AVX512F}},
    };
    return table[kind][dir][isa];
}

/* Scalar executors (drivers.c:94-110, 314-326, 414-430). */
static void run_default(int kind, int dir, const ft_rotation_plan *RP, double *A, int M) {
    int N = RP->n;
    ref_kernel k = pick(kind, dir, REF_DEFAULT);
    if (kind == 1) {
        #pragma omp parallel for schedule(static, 1)
        for (int m = 1; m < M; m++)
            k(RP, 0, m, A + (long)N * m, 1);
    } else {
        #pragma omp parallel for schedule(static, 1)
        for (int m = 2; m <= M / 2; m++) {
            k(RP, m % 2, m, A + (long)N * (2 * m - 1), 1);
            k(RP, m % 2, m, A + (long)N * (2 * m), 1);
        }
    }
}

/*
 * The SIMD executors.  Sphere and disk (drivers.c:112-126, 432-446): regroup
 * equal-parity orders (warp), interleave S columns row-major into B, run the
 * M % 2S leading columns with the scalar kernel, then every thread takes
 * blocks of 2S columns (S of order parity m%2, S of parity (m+1)%2).
 * Triangle (drivers.c:328-336): no warp, blocks of S consecutive orders.
 */
static void run_simd(int kind, int dir, int isa, const ft_rotation_plan *RP, double *A, double *B, int M) {
    int N = RP->n, S = isa_width(isa);
    ref_kernel kd = pick(kind, dir, REF_DEFAULT), ks = pick(kind, dir, isa);
    if (kind == 1) {
        permute_tri(A, B, N, M, S);
        for (int m = 1; m < M % S; m++)
            kd(RP, 0, m, B + (long)N * m, 1);
        #pragma omp parallel
        {
            int tid = omp_get_thread_num(), nt = omp_get_num_threads();
            for (int m = M % S + S * tid; m < M; m += S * nt)
                ks(RP, 0, m, B + (long)N * m, S);
        }
        permute_t_tri(A, B, N, M, S);
        return;
    }
    warp(A, N, M, S / 2);
    permute_sph(A, B, N, M, S);
    for (int m = 2; m <= (M % (2 * S)) / 2; m++) {
        kd(RP, m % 2, m, B + (long)N * (2 * m - 1), 1);
        kd(RP, m % 2, m, B + (long)N * (2 * m), 1);
    }
    #pragma omp parallel
    {
        int tid = omp_get_thread_num(), nt = omp_get_num_threads();
        for (int m = (M % (2 * S) + 1) / 2 + S * tid; m <= M / 2; m += S * nt) {
            ks(RP, m % 2, m, B + (long)N * (2 * m - 1), S);
            ks(RP, (m + 1) % 2, m + 1, B + (long)N * (2 * m - 1 + S), S);
        }
    }
    permute_t_sph(A, B, N, M, S);
    warp_t(A, N, M, S / 2);
}

/*
 * ft_execute_{sph,tri,disk}_{hi2lo,lo2hi} as the reference dispatches them.
 * s, c, n: the plan tables (ftinternal.h:370-374 layout).  isa < 0 selects
 * what the host supports, as drivers.c:58-74 does; B needs n*M doubles.
 */
void ref_execute(int kind, int dir, int isa, double *s, double *c, int n, double *A, double *B, int M) {
    ft_rotation_plan RP = {s, c, n};
    if (isa < 0) isa = ref_select_isa();
    if (isa == REF_DEFAULT) run_default(kind, dir, &RP, A, M);
    else run_simd(kind, dir, isa, &RP, A, B, M);
}

/* ft_kernel_{sph,tri,disk}_{hi2lo,lo2hi} (rotations.c:135-141, 162-168, 262-268):
 * always the scalar default kernel. */
void ref_kernel_column(int kind, int dir, double *s, double *c, int n, int m1, int m2, double *A, int S) {
    ft_rotation_plan RP = {s, c, n};
    pick(kind, dir, REF_DEFAULT)(&RP, m1, m2, A, S);
}

/* ft_execute_sphv_{hi2lo,lo2hi}'s scalar path (drivers.c:204-220) and
 * ft_execute_rectdisk_{hi2lo,lo2hi} (drivers.c:488-508, scalar only) around
 * the reference's compiled kernels (kernel_sph_*_default,
 * kernel_rectdisk_*_default). */
void ref_execute_sphv(int dir, double *s, double *c, int n, double *A, int M) {
    ft_rotation_plan RP = {s, c, n};
    ref_kernel k = pick(0, dir, REF_DEFAULT);
    #pragma omp parallel for schedule(static, 1)
    for (int m = 2; m <= M / 2 - 1; m++) {
        k(&RP, m % 2, m, A + (long)n * (2 * m + 1), 1);
        k(&RP, m % 2, m, A + (long)n * (2 * m + 2), 1);
    }
}

void ref_execute_rectdisk(int dir, double *s, double *c, int n, double *A, int M) {
    ft_rotation_plan RP = {s, c, n};
    ref_kernel k = dir == 0 ? kernel_rectdisk_hi2lo_default : kernel_rectdisk_lo2hi_default;
    #pragma omp parallel for schedule(static, 1)
    for (int m = 2; m < M; m++)
        k(&RP, m % 2, m, A + (long)n * m, 1);
}

/* ft_execute_tet_{hi2lo,lo2hi} (drivers.c:510-533) around the reference's
 * compiled triangle kernels (kernel_tri_*_default).  The slab columns run
 * through them as drivers.c:514 does (plan 1, size n).  The block rotations
 * (rotations.c:297-321, which cannot be compiled here) run as the same
 * compiled kernel on every row k of the slab at stride n, with the table of
 * the triangle plan of size L (s2, c2): rotations.c:297-321 applies, on each
 * row, j = m-1 .. 0 and l = L-2-j .. 0 with s(l, j), the loop nest of
 * rotations_default.c:45-50 for a plan of size L, whose entries are those of
 * the size-n plan (rotations.c:143-160 does not read the size). */
void ref_execute_tet(int dir, double *s1, double *c1, double *s2, double *c2, int n, int L, int M, double *A) {
    ft_rotation_plan RP1 = {s1, c1, n}, RP2 = {s2, c2, L};
    ref_kernel tri = pick(1, dir, REF_DEFAULT);
    #pragma omp parallel for schedule(static, 1)
    for (int m = 0; m < M; m++) {
        double *slab = A + (long)n * L * m;
        if (dir == 1)
            for (int k = 0; k < n; k++) tri(&RP2, 0, m, slab + k, n);
        for (int l = 0; l < L - m; l++) tri(&RP1, 0, l + m, slab + (long)n * l, 1);
        if (dir == 0)
            for (int k = 0; k < n; k++) tri(&RP2, 0, m, slab + k, n);
    }
}

/* ft_execute_spinsph_{hi2lo,lo2hi} (drivers.c:645-773) around the
 * reference's compiled spin kernels: isa REF_DEFAULT the scalar executor
 * (drivers.c:687-705, kernel_spinsph_*_default), REF_AVX_FMA the dispatched
 * one (drivers.c:747-773: orders 0, -1, 1 by the scalar kernel, the rest as
 * column pairs permuted into B by permute_sph and run by
 * kernel_spinsph_*_AVX_FMA).  B holds 2n x M doubles. */
void ref_execute_spinsph(int dir, int isa, double *s1, double *c1, double *s2, double *c2, int n, int s, double *A,
                         double *B, int M) {
    ft_spin_rotation_plan SRP = {s1, c1, s2, c2, n, s};
    ft_complex *AC = (ft_complex *)A, *BC = (ft_complex *)B;
    void (*kd)(const ft_spin_rotation_plan *, const int, ft_complex *, const int) =
        dir == 0 ? kernel_spinsph_hi2lo_default : kernel_spinsph_lo2hi_default;
    if (isa == REF_DEFAULT) {
        kd(&SRP, 0, AC, 1);
        #pragma omp parallel for schedule(static, 1)
        for (int m = 1; m <= M / 2; m++) {
            kd(&SRP, -m, AC + (long)n * (2 * m - 1), 1);
            kd(&SRP, m, AC + (long)n * (2 * m), 1);
        }
        return;
    }
    void (*kf)(const ft_spin_rotation_plan *, const int, ft_complex *, const int) =
        dir == 0 ? kernel_spinsph_hi2lo_AVX_FMA : kernel_spinsph_lo2hi_AVX_FMA;
    kd(&SRP, 0, AC, 1);
    if (M > 1) kd(&SRP, -1, AC + n, 1);
    if (M > 2) kd(&SRP, 1, AC + 2 * n, 1);
    permute_sph(A, B, 2 * n, M, 2);
    #pragma omp parallel for schedule(static, 1)
    for (int m = 2; m <= M / 2; m++) kf(&SRP, m, BC + (long)n * (2 * m - 1), 2);
    permute_t_sph(A, B, 2 * n, M, 2);
}

void ref_set_num_threads(int t) { omp_set_num_threads(t); }
int ref_max_threads(void) { return omp_get_max_threads(); }
