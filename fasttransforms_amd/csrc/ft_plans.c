/*
 * ft_plans.c -- rotation plans: FP64 tables built on the host, uploaded to
 * HBM once (FP64, and an FP32 copy rounded from FP64 on first FP32 use).
 *
 * The tables are the reference's closed forms (src/rotations.c:113-190) with
 * the same operation order, so they are bit-identical: integer products for
 * the sphere, left-to-right double sums for the triangle and the disk, then
 * one correctly rounded division and one sqrt per entry.  Compiled with
 * -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

static long idx_tri(int n, int l, int m) { return l + (long)m * (2 * n + 1 - m) / 2; }
static long idx_disk(int n, int l, int m) { return l + (long)m * n - (long)(m / 2) * ((m + 1) / 2); }

static int g_num_threads = 1;

/* fasttransforms.h:295.  Recorded only: no host threads compute here. */
void ft_set_num_threads(const int n) { g_num_threads = n; }

ft_rotation_plan *ftamd_plan_alloc(int kind, int family, int n, double alpha, double beta, double gamma) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_rotation_plan *RP = calloc(1, sizeof(ft_rotation_plan));
    if (!RP) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    RP->n = n;
    RP->kind = kind;
    RP->family = family;
    RP->alpha = alpha;
    RP->beta = beta;
    RP->gamma = gamma;
    RP->nsc = kind == FTAMD_DISK ? (long)n * n : (long)n * (n + 1) / 2;
    RP->s = malloc((size_t)RP->nsc * sizeof(double));
    RP->c = malloc((size_t)RP->nsc * sizeof(double));
    RP->device = -1;
    static unsigned long long next_serial = 0;
    RP->serial = __atomic_add_fetch(&next_serial, 1, __ATOMIC_RELAXED);
    pthread_mutex_init(&RP->lock, NULL);
    pthread_mutex_init(&RP->exec_lock, NULL);
    if (!RP->s || !RP->c) {
        free(RP->s);
        free(RP->c);
        free(RP);
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    return RP;
}

/* rotations.c:116-133 */
ft_rotation_plan *ft_plan_rotsphere(const int n) {
    ft_rotation_plan *RP = ftamd_plan_alloc(FTAMD_SPH, FT_GPU_SPH, n, 0.0, 0.0, 0.0);
    if (!RP) return NULL;
    for (int m = 0; m < n; m++)
        for (int l = 0; l < n - m; l++) {
            double nums = (double)((l + 1) * (l + 2));
            double numc = (double)((2 * m + 2) * (2 * l + 2 * m + 5));
            double den = (double)((l + 2 * m + 3) * (l + 2 * m + 4));
            RP->s[idx_tri(n, l, m)] = sqrt(nums / den);
            RP->c[idx_tri(n, l, m)] = sqrt(numc / den);
        }
    return RP;
}

/* rotations.c:143-160 */
ft_rotation_plan *ft_plan_rottriangle(const int n, const double alpha, const double beta, const double gamma) {
    ft_rotation_plan *RP = ftamd_plan_alloc(FTAMD_TRI, FT_GPU_TRI, n, alpha, beta, gamma);
    if (!RP) return NULL;
    for (int m = 0; m < n; m++)
        for (int l = 0; l < n - m; l++) {
            double nums = (double)(l + 1) * ((l + alpha) + 1);
            double numc = ((((2 * m) + beta) + gamma) + 2) * (((((2 * l + 2 * m) + alpha) + beta) + gamma) + 4);
            double den = ((((l + 2 * m) + beta) + gamma) + 3) * (((((l + 2 * m) + alpha) + beta) + gamma) + 3);
            RP->s[idx_tri(n, l, m)] = sqrt(nums / den);
            RP->c[idx_tri(n, l, m)] = sqrt(numc / den);
        }
    return RP;
}

/* rotations.c:173-190 */
ft_rotation_plan *ft_plan_rotdisk(const int n, const double alpha, const double beta) {
    ft_rotation_plan *RP = ftamd_plan_alloc(FTAMD_DISK, FT_GPU_DISK, n, alpha, beta, 0.0);
    if (!RP) return NULL;
    for (int m = 0; m < 2 * n - 1; m++)
        for (int l = 0; l < n - (m + 1) / 2; l++) {
            double nums = (double)(l + 1) * ((l + beta) + 1);
            double numc = ((m + alpha) + 1) * ((((2 * l + m) + alpha) + beta) + 3);
            double den = (((l + m) + alpha) + 2) * ((((l + m) + alpha) + beta) + 2);
            RP->s[idx_disk(n, l, m)] = -sqrt(nums / den);
            RP->c[idx_disk(n, l, m)] = sqrt(numc / den);
        }
    return RP;
}

/* rotations.c:443-475: s1, c1 (n x n, s1(l, m) at l + m*n) the sphere's
 * closed form, s2, c2 (|s| x n x n, s2(l, j, m) at l + j*n + m*|s|*n) for
 * l < n-m-j; every term an exact integer, as the reference computes them.
 * The sphere plan of size n beside it holds the same s1, c1 values in the
 * sphere's packed layout (the second stage runs on it). */
ft_spin_rotation_plan *ft_plan_rotspinsphere(const int n, const int s) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    const int as = s < 0 ? -s : s;
    ft_spin_rotation_plan *P = calloc(1, sizeof(ft_spin_rotation_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    const size_t nn = (size_t)n * n, n2 = (size_t)as * n * n;
    P->s1 = calloc(nn, sizeof(double));
    P->c1 = calloc(nn, sizeof(double));
    P->s2 = calloc(n2 > 0 ? n2 : 1, sizeof(double));
    P->c2 = calloc(n2 > 0 ? n2 : 1, sizeof(double));
    P->n = n;
    P->s = s;
    P->device = -1;
    pthread_mutex_init(&P->lock, NULL);
    P->sph = ft_plan_rotsphere(n);
    if (!P->s1 || !P->c1 || !P->s2 || !P->c2 || !P->sph) {
        ft_destroy_spin_rotation_plan(P);
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    for (int m = 0; m < n; m++)
        for (int l = 0; l < n - m; l++) {
            const double nums = (double)((long long)(l + 1) * (l + 2));
            const double numc = (double)((long long)(2 * m + 2) * (2 * l + 2 * m + 5));
            const double den = (double)((long long)(l + 2 * m + 3) * (l + 2 * m + 4));
            P->s1[l + (size_t)m * n] = sqrt(nums / den);
            P->c1[l + (size_t)m * n] = sqrt(numc / den);
        }
    for (int m = 0; m < n; m++)
        for (int j = 0; j < as; j++)
            for (int l = 0; l < n - m - j; l++) {
                const double nums = (double)((long long)(l + 1) * (l + m + 1));
                const double numc = (double)((long long)(2 * j + m + 1) * (2 * l + 2 * j + 2 * m + 3));
                const double den = (double)((long long)(l + 2 * j + m + 2) * (l + 2 * j + 2 * m + 2));
                const size_t i = l + (size_t)j * n + (size_t)m * as * n;
                P->s2[i] = sqrt(nums / den);
                P->c2[i] = sqrt(numc / den);
            }
    return P;
}

/* rotations.c:431-437 -- also frees the device copies and the sphere plan. */
void ft_destroy_spin_rotation_plan(ft_spin_rotation_plan *P) {
    if (!P) return;
    if (P->d_tab2) {
        ftamd_hip_sync(NULL);
        ftamd_hip_free(P->d_tab2);
    }
    ft_destroy_rotation_plan(P->sph);
    free(P->s1);
    free(P->c1);
    free(P->s2);
    free(P->c2);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* Uploads the (s2, c2) pairs on first use (caller holds P->lock). */
int ftamd_spin_upload(ft_spin_rotation_plan *P) {
    const int as = P->s < 0 ? -P->s : P->s;
    if (P->d_tab2 || as == 0) return FTAMD_OK;
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) return rc;
    if (P->device >= 0 && P->device != dev) return FTAMD_EDEVICE;
    const size_t count = (size_t)as * P->n * P->n;
    double *host = malloc(2 * count * sizeof(double));
    if (!host) return FTAMD_ENOMEM;
    for (size_t i = 0; i < count; i++) {
        host[2 * i] = P->s2[i];
        host[2 * i + 1] = P->c2[i];
    }
    void *d = NULL;
    rc = ftamd_hip_malloc(&d, 2 * count * sizeof(double));
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, host, 2 * count * sizeof(double), NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    free(host);
    if (rc != FTAMD_OK) {
        ftamd_hip_free(d);
        return rc;
    }
    P->d_tab2 = d;
    P->device = dev;
    return FTAMD_OK;
}

/* rotations.c:270-287.  The sphere's table layout and kernels; each term is
 * summed left to right in double as the reference writes it, the numerator
 * of s an exact integer product. */
ft_rotation_plan *ft_plan_rotrectdisk(const int n, const double beta) {
    ft_rotation_plan *RP = ftamd_plan_alloc(FTAMD_SPH, FT_GPU_RECTDISK, n, 0.0, beta, 0.0);
    if (!RP) return NULL;
    for (int m = 0; m < n; m++)
        for (int l = 0; l < n - m; l++) {
            double nums = (double)((l + 1) * (l + 2));
            double numc = (((2 * m) + 2 * beta) + 3) * ((((2 * l + 2 * m) + 2 * beta)) + 6);
            double den = (((l + 2 * m) + 2 * beta) + 4) * (((l + 2 * m) + 2 * beta) + 5);
            RP->s[idx_tri(n, l, m)] = sqrt(nums / den);
            RP->c[idx_tri(n, l, m)] = sqrt(numc / den);
        }
    return RP;
}

/* ------------------------------------------------------------ annulus -- */

/* The normalized Jacobi operator x P^{(a,b)} on N coefficients as the
 * reference's banded constructor evaluates it (banded_source.c:1797-1818,
 * norm 1): diagonal d[j], superdiagonal up[j] = (j, j+1), subdiagonal
 * lo[j] = (j+1, j).  The two off-diagonals are equal in exact arithmetic but
 * evaluated by different expressions; the QR steps read the first, the
 * Cholesky factor both. */
static void jacobi_operator(int N, double a, double b, double *d, double *up, double *lo) {
    for (int j = 0; j < N; j++) {
        double v;
        if (j == 1)
            v = 2 * sqrt(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)));
        else
            v = 2 * sqrt((j * (j + a) * (j + b) * (j + a + b)) /
                         ((2 * j + a + b - 1) * (2 * j + a + b) * (2 * j + a + b) * (2 * j + a + b + 1)));
        if (j > 0) up[j - 1] = v;
        if (j == 0) {
            d[0] = (b - a) / (a + b + 2);
            if (N > 1) lo[0] = 2 * sqrt(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)));
        } else {
            d[j] = ((b - a) * (a + b)) / ((2 * j + a + b) * (2 * j + a + b + 2));
            if (j + 1 < N)
                lo[j] = 2 * sqrt(((j + 1) * (j + a + 1) * (j + b + 1) * (j + a + b + 1)) /
                                 ((2 * j + a + b + 1) * (2 * j + a + b + 2) * (2 * j + a + b + 2) * (2 * j + a + b + 3)));
        }
    }
}

/* Givens rotation zeroing y against x (banded_source.c:1415-1425). */
static void givens(double x, double y, double *c, double *s, double *r) {
    *r = hypot(x, y);
    if (*r <= 0x1p-970) { /* floatmin / eps */
        *c = 1;
        *s = 0;
    } else {
        *c = x / *r;
        *s = y / *r;
    }
}

/* QR factorization by Givens rotations of the symmetric tridiagonal (a, b)
 * of size N >= 2 (banded_source.c:1429-1456): the N-1 rotations (c, s) and
 * the diagonal and first superdiagonal of R. */
static void tridiag_qr(int N, const double *a, const double *b, double *c, double *s, double *rd, double *ru) {
    double r, aip1 = a[0], bip1r = b[0];
    for (int i = 0; i < N - 2; i++) {
        givens(aip1, -b[i], c + i, s + i, &r);
        rd[i] = r;
        ru[i] = c[i] * bip1r - s[i] * a[i + 1];
        aip1 = c[i] * a[i + 1] + s[i] * bip1r;
        bip1r = c[i] * b[i + 1];
    }
    givens(aip1, -b[N - 2], c + N - 2, s + N - 2, &r);
    rd[N - 2] = r;
    ru[N - 2] = c[N - 2] * bip1r - s[N - 2] * a[N - 1];
    rd[N - 1] = c[N - 2] * a[N - 1] + s[N - 2] * bip1r;
}

/* The leading N-1 x N-1 block of R X R^{-1} for the symmetric tridiagonal X
 * = (a, b) of size N and the upper triangular R (diagonal rd, superdiagonal
 * ru), in place: a gets N-1 entries, b N-2 (banded_source.c:1609-1617,
 * 1641-1649). */
static void jacobi_similarity(int N, double *a, double *b, const double *rd, const double *ru) {
    double bprev = 0;
    for (int i = 0; i < N - 1; i++) {
        const double bq = i < N - 2 ? rd[i + 1] / rd[i] * b[i] : 0;
        if (i == 0) a[0] = a[0] + ru[0] / rd[0] * b[0];
        else a[i] = (rd[i] * a[i] + ru[i] * b[i] - ru[i - 1] * bprev) / rd[i];
        if (i < N - 2) b[i] = bq;
        bprev = bq;
    }
}

/* rotations.c:194-260 (gamma = 0): the rotations of the annulus with inner
 * radius rho, in the disk's layout and kernels.  With t = 1 + 2 rho^2 / (1 -
 * rho^2) and J the Jacobi operator of P^{(beta, alpha)} on n+1 coefficients,
 * the even orders m = 0, 2, ... take the Givens rotations of the QR
 * factorization of J + tI, then of R (J + tI) R^{-1} (one size smaller) and
 * so on; the odd orders start from U (J + tI) U^{-1} with U the Cholesky
 * factor of J + tI (banded_source.c:484-499).  Host FP64 in the reference's
 * operation order, with no contraction. */
ft_rotation_plan *ft_plan_rotannulus(const int n, const double alpha, const double beta, const double gamma,
                                     const double rho) {
    if (n < 1 || gamma != 0.0 || !(rho >= 0.0 && rho < 1.0)) {
        ftamd_set_error(FT_GPU_EINVAL); /* the reference implements gamma = 0 only (rotations.c:252-254) */
        return NULL;
    }
    ft_rotation_plan *RP = ftamd_plan_alloc(FTAMD_DISK, FTAMD_FAMILY_ANNULUS, n, alpha, beta, 0.0);
    if (!RP) return NULL;
    RP->rho = rho;
    const int N = n + 1;
    double *w = malloc(sizeof(double) * (size_t)(9 * N));
    if (!w) {
        ft_destroy_rotation_plan(RP);
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    double *d = w, *up = w + N, *lo = w + 2 * N, *a = w + 3 * N, *b = w + 4 * N, *c = w + 5 * N, *s = w + 6 * N,
           *rd = w + 7 * N, *ru = w + 8 * N;
    const double t = 1.0 + 2.0 * rho * rho / (1.0 - rho * rho);
    jacobi_operator(N, beta, alpha, d, up, lo);
    for (int par = 0; par < 2; par++) {
        int size = N;
        for (int i = 0; i < N; i++) a[i] = d[i] + t;
        for (int i = 0; i < N - 1; i++) b[i] = up[i];
        if (par == 1) {
            /* U: Cholesky factor of J + tI, rows j: U(j,j) = sqrt(p_j),
             * U(j,j+1) = up[j] / sqrt(p_j), p_{j+1} = (t + d[j+1]) - lo[j]/p_j * up[j] */
            double p = t + d[0];
            for (int j = 0; j < N; j++) {
                const double mult = j + 1 < N ? lo[j] / p : 0;
                rd[j] = sqrt(p);
                if (j + 1 < N) {
                    const double pn = (t + d[j + 1]) - mult * up[j];
                    ru[j] = up[j] / rd[j];
                    p = pn;
                }
            }
            jacobi_similarity(size, a, b, rd, ru);
            size--;
        }
        for (int m = par; m < 2 * n - 1; m += 2) {
            tridiag_qr(size, a, b, c, s, rd, ru);
            for (int l = 0; l < n - (m + 1) / 2; l++) {
                RP->s[idx_disk(n, l, m)] = s[l];
                RP->c[idx_disk(n, l, m)] = c[l];
            }
            jacobi_similarity(size, a, b, rd, ru);
            size--;
        }
    }
    free(w);
    return RP;
}

/* rotations.c:6-10, plus the device state. */
void ft_destroy_rotation_plan(ft_rotation_plan *RP) {
    if (!RP) return;
    while (RP->subs) {
        ft_rotation_plan *next = RP->subs->sub_next;
        ft_destroy_rotation_plan(RP->subs);
        RP->subs = next;
    }
    /* a cached batch schedule that uses this plan's tables goes with it */
    ftamd_batch_forget_plan(RP);
    if (RP->device >= 0) ftamd_hip_sync(NULL);
    ft_schedule *sc = RP->scheds;
    while (sc) {
        ft_schedule *next = sc->next;
        ftamd_schedule_release(sc);
        sc = next;
    }
    for (int d = 0; d < 2; d++)
        for (int p = 0; p < 2; p++) ftamd_sys_tables_free(RP->sys[d][p]);
    ftamd_hip_free(RP->d_tab[0]);
    ftamd_hip_free(RP->d_tab[1]);
    ftamd_hip_free(RP->d_stage);
    pthread_mutex_destroy(&RP->lock);
    pthread_mutex_destroy(&RP->exec_lock);
    free(RP->s);
    free(RP->c);
    free(RP);
}

/*
 * Upload the interleaved table {s0, c0, s1, c1, ...} with FTAMD_TABLE_PAD zero
 * entries on each side.  FP32 entries are (float) of the FP64 ones.  The
 * sphere's kernels run even and odd rows as separate chains, so each of its
 * columns is stored with the even rows l = 0, 2, ... first and the odd rows
 * after them (both then contiguous).  Caller holds RP->lock.
 */
int ftamd_plan_upload(ft_rotation_plan *RP, int precision) {
    if (RP->d_tab[precision]) return FTAMD_OK;
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) return rc;
    if (RP->device >= 0 && RP->device != dev) return FTAMD_EDEVICE;
    size_t esz = precision == 0 ? sizeof(double) : sizeof(float);
    size_t count = 2 * ((size_t)RP->nsc + 2 * FTAMD_TABLE_PAD);
    void *host = calloc(count, esz);
    if (!host) return FTAMD_ENOMEM;
    size_t base = 2 * (size_t)FTAMD_TABLE_PAD;
    int n = RP->n;
    int ncol = RP->kind == FTAMD_DISK ? 2 * n - 1 : n;
    for (int m = 0; m < ncol; m++) {
        long start = RP->kind == FTAMD_DISK ? idx_disk(n, 0, m) : idx_tri(n, 0, m);
        long rows = RP->kind == FTAMD_DISK ? n - (m + 1) / 2 : n - m;
        long nev = (rows + 1) / 2;
        for (long l = 0; l < rows; l++) {
            long src = start + l, dst = src;
            if (RP->kind == FTAMD_SPH) dst = start + (l % 2 == 0 ? l / 2 : nev + l / 2);
            if (precision == 0) {
                ((double *)host)[base + 2 * dst] = RP->s[src];
                ((double *)host)[base + 2 * dst + 1] = RP->c[src];
            } else {
                ((float *)host)[base + 2 * dst] = (float)RP->s[src];
                ((float *)host)[base + 2 * dst + 1] = (float)RP->c[src];
            }
        }
    }
    void *d = NULL;
    rc = ftamd_hip_malloc(&d, count * esz);
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, host, count * esz, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    free(host);
    if (rc != FTAMD_OK) {
        ftamd_hip_free(d);
        return rc;
    }
    RP->d_tab[precision] = d;
    RP->device = dev;
    return FTAMD_OK;
}

/* Device address of table entry 0 (past the leading padding). */
const void *ftamd_plan_table(const ft_rotation_plan *RP, int precision) {
    size_t esz = precision == 0 ? sizeof(double) : sizeof(float);
    return (const char *)RP->d_tab[precision] + 2 * (size_t)FTAMD_TABLE_PAD * esz;
}
