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
 *   errors within 2*n*eps.
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

int main(void) {
    ft_set_num_threads(1);
    for (int i = 0; i < 3; i++)
        for (int kind = 0; kind < 3; kind++) roundtrip(kind, 64 * (1 << i) + 1);
    for (int kind = 0; kind < 3; kind++) columns(kind, 100);
    printf("%s (%d failures)\n", failures ? "FAILED" : "all checks passed", failures);
    return failures != 0;
}
