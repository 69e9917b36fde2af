/*
 * A C caller of the drop-in boundary, built against include/fasttransforms_amd.h
 * and linked with -lfasttransforms_amd the way the reference's own tests link
 * libfasttransforms.  It runs the reference's accuracy checks for this path:
 *
 *   test_drivers.c:37-57 (and the tri/disk blocks after it): hi2lo then
 *   lo2hi of an n x M matrix returns it within 4*sqrt(N)*eps (relative
 *   2-norm) and 2*N*eps (relative inf-norm), N = 64*2^i + J;
 *   test_rotations.c:25-39: the single-column kernels ft_kernel_*_hi2lo then
 *   _lo2hi for every order m of a column of ones, summed squared relative
 *   errors within 2*n*eps;
 *   test_drivers.c:226-248, 700-720, 937-957, 1128-1156, 1234-1254: the
 *   harmonic transforms and their inverses, 'N' then 'T', within the bounds
 *   each block states (sph2fourier and rectdisk2cheb 4 sqrt(N), 2N; disk2cxf
 *   8 sqrt(N), 4N; ann2cxf 12 sqrt(N), 6N) except tri2cheb's: its 32 sqrt(N),
 *   8N are exceeded at N = 257 on uniform [-1, 1] data by the reference's own
 *   FP64 arithmetic (its composite with FP64 products, as its BLAS computes
 *   them, gives 33 sqrt(N) and 12 N eps; P[0] has condition number 2.4e3
 *   there), so tri2cheb is held to 64 sqrt(N), 32 N here and to the oracle
 *   composite (1e-14) by the Python GPU tests.
 *
 * The matrices are filled from a fixed 64-bit LCG (the reference's tests use
 * all-ones or rand(); any fixed fill serves a round trip).  Exit status 0 when
 * every check passes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fasttransforms_amd.h"

static const double EPSD = 0x1p-52;  /* the reference's eps() (ftinternal.h:62, 91) */
static int failures = 0;

static void check(const char *what, int n, double err, double cst) {
    int ok = fabs(err) < cst * EPSD;
    printf("%-44s n = %5d: %10.3e (bound %.3e) %s\n", what, n, err, cst * EPSD, ok ? "ok" : "FAIL");
    if (!ok) failures++;
}

static double norm2(const double *a, const double *b, long len) {
    double d = 0, r = 0;
    for (long i = 0; i < len; i++) {
        double x = b ? a[i] - b[i] : a[i];
        d += x * x;
    }
    (void)r;
    return sqrt(d);
}

static double norminf(const double *a, const double *b, long len) {
    double d = 0;
    for (long i = 0; i < len; i++) {
        double x = fabs(b ? a[i] - b[i] : a[i]);
        if (x > d) d = x;
    }
    return d;
}

static unsigned long long lcg = 0x9E3779B97F4A7C15ULL;
static double next_uniform(void) {
    lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
    return 2.0 * ((double)(lcg >> 11) * 0x1p-53) - 1.0;
}

/* valid rows of column j of each layout (ftutilities.c:75-113) */
static int valid_rows(int kind, int n, int j) {
    if (kind == 0) return n - (j + 1) / 2;
    if (kind == 1) return n - j;
    return n - (j + 3) / 4;
}

static void roundtrip(int kind, int N) {
    static const char *names[] = {"sph", "tri", "disk"};
    int M = kind == 0 ? 2 * N - 1 : kind == 1 ? N : 4 * N - 3;
    long len = (long)N * M;
    double *A = malloc(sizeof(double) * len), *B = malloc(sizeof(double) * len), *W = malloc(sizeof(double) * len);
    for (int j = 0; j < M; j++)
        for (int i = 0; i < N; i++) A[i + (long)N * j] = i < valid_rows(kind, N, j) ? next_uniform() : 0.0;
    memcpy(B, A, sizeof(double) * len);
    ft_rotation_plan *RP = kind == 0 ? ft_plan_rotsphere(N) : kind == 1 ? ft_plan_rottriangle(N, 0.0, -0.5, -0.5)
                                                                        : ft_plan_rotdisk(N, 0.0, 0.0);
    if (kind == 0) {
        ft_execute_sph_hi2lo(RP, A, W, M);
        ft_execute_sph_lo2hi(RP, A, W, M);
    } else if (kind == 1) {
        ft_execute_tri_hi2lo(RP, A, W, M);
        ft_execute_tri_lo2hi(RP, A, W, M);
    } else {
        ft_execute_disk_hi2lo(RP, A, W, M);
        ft_execute_disk_lo2hi(RP, A, W, M);
    }
    int e = ft_gpu_last_error();
    if (e) {
        printf("%s: %s\n", names[kind], ft_gpu_error_string(e));
        failures++;
    }
    char what[64];
    snprintf(what, sizeof what, "%s hi2lo->lo2hi eps_2", names[kind]);
    check(what, N, norm2(A, B, len) / norm2(B, NULL, len), 4 * sqrt(N));
    snprintf(what, sizeof what, "%s hi2lo->lo2hi eps_inf", names[kind]);
    check(what, N, norminf(A, B, len) / norminf(B, NULL, len), 2 * N);
    ft_destroy_rotation_plan(RP);
    free(A);
    free(B);
    free(W);
}

static void columns(int kind, int n) {
    static const char *names[] = {"sph", "tri", "disk"};
    ft_rotation_plan *RP = kind == 0 ? ft_plan_rotsphere(n) : kind == 1 ? ft_plan_rottriangle(n, 0.0, -0.5, -0.5)
                                                                        : ft_plan_rotdisk(n, 0.0, 0.0);
    double *A = calloc(n, sizeof(double)), *B = calloc(n, sizeof(double));
    double err = 0;
    int mmax = kind == 2 ? 2 * n - 1 : n;
    for (int m = 2; m < mmax; m++) {
        int rows = kind == 2 ? n - (m + 1) / 2 : n - m;
        for (int i = 0; i < n; i++) A[i] = B[i] = i < rows ? 1.0 : 0.0;
        int m1 = kind == 1 ? 0 : m % 2;
        if (kind == 0) {
            ft_kernel_sph_hi2lo(RP, m1, m, A, 1);
            ft_kernel_sph_lo2hi(RP, m1, m, A, 1);
        } else if (kind == 1) {
            ft_kernel_tri_hi2lo(RP, m1, m, A, 1);
            ft_kernel_tri_lo2hi(RP, m1, m, A, 1);
        } else {
            ft_kernel_disk_hi2lo(RP, m1, m, A, 1);
            ft_kernel_disk_lo2hi(RP, m1, m, A, 1);
        }
        double r = norm2(A, B, n) / norm2(B, NULL, n);
        err += r * r;
    }
    int e = ft_gpu_last_error();
    if (e) {
        printf("%s columns: %s\n", names[kind], ft_gpu_error_string(e));
        failures++;
    }
    char what[64];
    snprintf(what, sizeof what, "%s single columns, all orders", names[kind]);
    check(what, n, sqrt(err), 2 * n);
    ft_destroy_rotation_plan(RP);
    free(A);
    free(B);
}

/* kind 0 sph2fourier, 1 tri2cheb, 2 disk2cxf, 3 rectdisk2cheb, 4 ann2cxf */
static void harmonic(int kind, int N) {
    static const char *names[] = {"sph2fourier", "tri2cheb", "disk2cxf", "rectdisk2cheb", "ann2cxf"};
    static const double c2[] = {4, 64, 8, 4, 12}, cinf[] = {2, 32, 4, 2, 6};
    const int M = kind == 0 ? 2 * N - 1 : (kind == 2 || kind == 4) ? 4 * N - 3 : N;
    const int shape = kind == 0 ? 0 : (kind == 2 || kind == 4) ? 2 : 1; /* valid rows of the sph, tri, disk layouts */
    long len = (long)N * M;
    double *A = malloc(sizeof(double) * len), *B = malloc(sizeof(double) * len);
    for (int j = 0; j < M; j++)
        for (int i = 0; i < N; i++) A[i + (long)N * j] = i < valid_rows(shape, N, j) ? next_uniform() : 0.0;
    memcpy(B, A, sizeof(double) * len);
    ft_harmonic_plan *P = kind == 0 ? ft_plan_sph2fourier(N)
                          : kind == 1 ? ft_plan_tri2cheb(N, 0.0, 0.0, 0.0)
                          : kind == 2 ? ft_plan_disk2cxf(N, 0.0, 0.0)
                          : kind == 3 ? ft_plan_rectdisk2cheb(N, 0.0)
                                      : ft_plan_ann2cxf(N, 0.0, 0.0, 0.0, sqrt(1.0 / 3.0));
    if (!P) {
        printf("%s: plan failed: %s\n", names[kind], ft_gpu_error_string(ft_gpu_last_error()));
        failures++;
        free(A);
        free(B);
        return;
    }
    for (int t = 0; t < 2; t++) {
        const char T = t ? 'T' : 'N';
        switch (kind) {
        case 0: ft_execute_sph2fourier(T, P, A, N, M); ft_execute_fourier2sph(T, P, A, N, M); break;
        case 1: ft_execute_tri2cheb(T, P, A, N, M); ft_execute_cheb2tri(T, P, A, N, M); break;
        case 2: ft_execute_disk2cxf(T, P, A, N, M); ft_execute_cxf2disk(T, P, A, N, M); break;
        case 3: ft_execute_rectdisk2cheb(T, P, A, N, M); ft_execute_cheb2rectdisk(T, P, A, N, M); break;
        default: ft_execute_ann2cxf(T, P, A, N, M); ft_execute_cxf2ann(T, P, A, N, M); break;
        }
    }
    int e = ft_gpu_last_error();
    if (e) {
        printf("%s: %s\n", names[kind], ft_gpu_error_string(e));
        failures++;
    }
    char what[64];
    snprintf(what, sizeof what, "%s and inverse, N then T, eps_2", names[kind]);
    check(what, N, norm2(A, B, len) / norm2(B, NULL, len), c2[kind] * sqrt(N));
    snprintf(what, sizeof what, "%s and inverse, N then T, eps_inf", names[kind]);
    check(what, N, norminf(A, B, len) / norminf(B, NULL, len), cinf[kind] * N);
    ft_destroy_harmonic_plan(P);
    free(A);
    free(B);
}

int main(void) {
    ft_set_num_threads(1);
    for (int i = 0; i < 3; i++)
        for (int kind = 0; kind < 3; kind++) roundtrip(kind, 64 * (1 << i) + 1);
    for (int kind = 0; kind < 3; kind++) columns(kind, 100);
    for (int i = 0; i < 3; i++)
        for (int kind = 0; kind < 5; kind++) harmonic(kind, 64 * (1 << i) + 1);
    printf("%s (%d failures)\n", failures ? "FAILED" : "all checks passed", failures);
    return failures != 0;
}
