/*
 * ft_harmonic.c -- the harmonic transforms (drivers.c:793-1021): the rotation
 * stage on the GPU followed (or preceded) by the upper-triangular products
 * with the 1D connection matrices, which run as matrix-core launches
 * (ft_trmm.hip) instead of BLAS calls.
 *
 * sph2fourier / disk2cxf: TRANS 'N' is hi2lo then the products with P on the
 * column classes j mod 4 (the orders of even parity sit in j = 0, 3 and take
 * P[0], the odd ones in j = 1, 2 and take P[1]; the vector-field layout,
 * columns 2m+1 and 2m+2, swaps them, drivers.c:867-899); 'T' is the
 * transposed products then lo2hi; the inverses (fourier2sph, cxf2disk) use
 * Pinv with lo2hi after ('N') or hi2lo before ('T').  tri2cheb: hi2lo, P[0]
 * from the left on every column, P[1]^T from the right on every row, the 2D
 * Chebyshev normalization; cheb2tri runs the inverses in reverse order
 * (drivers.c:921-957).  ann2cxf: the disk's pattern with the annulus's
 * rotations, a bidiagonal solve (or product) on the columns j = 1, 2 mod 4
 * and one matrix P[0] on every column (drivers.c:1023-1121).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

/* A modified plan of the annulus (drivers.c:1027-1046): the identity (b 0)
 * or the upper bidiagonal R (b 1), R[0 .. n-1] its diagonal and R[n ..
 * 2n-2] its superdiagonal.  nu, nv as the reference's members. */
struct ft_modified_plan_s {
    int n, nu, nv, b;
    double *R;
};

struct ftamd_harmonic_dev {
    pthread_mutex_t lock; /* held for a whole call: guards the buffers below */
    int device;           /* HIP device of the uploads, -1 before */
    double *dP[3], *dPinv[3];
    double *dR; /* MP[1]'s bidiagonal factor (the annulus's plans) */
    void *dY;             /* product output, n x M */
    size_t y_bytes;
    void *dA;             /* staging of host matrices */
    size_t a_bytes;
};

void ft_destroy_harmonic_plan(ft_harmonic_plan *P) {
    if (!P) return;
    if (P->dev) {
        struct ftamd_harmonic_dev *d = P->dev;
        if (d->device >= 0) ftamd_hip_sync(NULL);
        for (int k = 0; k < 3; k++) {
            ftamd_hip_free(d->dP[k]);
            ftamd_hip_free(d->dPinv[k]);
        }
        ftamd_hip_free(d->dR);
        ftamd_hip_free(d->dY);
        ftamd_hip_free(d->dA);
        pthread_mutex_destroy(&d->lock);
        free(d);
    }
    for (int i = 0; P->RP && i < P->NRP; i++) ft_destroy_rotation_plan(P->RP[i]);
    free(P->RP);
    for (int i = 0; i < P->NP; i++) {
        if (P->P) free(P->P[i]);
        if (P->Pinv) free(P->Pinv[i]);
    }
    free(P->P);
    free(P->Pinv);
    for (int i = 0; P->MP && i < P->NMP; i++) {
        if (P->MP[i]) free(P->MP[i]->R);
        free(P->MP[i]);
    }
    free(P->MP);
    free(P);
}

/* A plan with one rotation plan and np matrix pairs, members allocated
 * (NULL if out of memory). */
static ft_harmonic_plan *plan_alloc(int np) {
    ft_harmonic_plan *P = calloc(1, sizeof(ft_harmonic_plan));
    if (!P) return NULL;
    P->dev = calloc(1, sizeof(struct ftamd_harmonic_dev));
    if (P->dev) {
        pthread_mutex_init(&P->dev->lock, NULL);
        P->dev->device = -1;
    }
    P->RP = calloc(1, sizeof(ft_rotation_plan *));
    P->P = calloc((size_t)np, sizeof(double *));
    P->Pinv = calloc((size_t)np, sizeof(double *));
    if (!P->dev || !P->RP || !P->P || !P->Pinv) {
        ft_destroy_harmonic_plan(P);
        return NULL;
    }
    P->NRP = 1;
    P->NP = np;
    return P;
}

/* The plan if every member was built, else NULL (destroyed; error set). */
static ft_harmonic_plan *plan_done(ft_harmonic_plan *P) {
    int ok = P && P->RP[0];
    for (int k = 0; ok && k < P->NP; k++) ok = P->P[k] && P->Pinv[k];
    if (ok) return P;
    ft_destroy_harmonic_plan(P);
    ftamd_set_error(FT_GPU_ENOMEM);
    return NULL;
}

/* drivers.c:816-831 */
ft_harmonic_plan *ft_plan_sph2fourier(const int n) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = plan_alloc(2);
    if (!P) return plan_done(NULL);
    P->RP[0] = ft_plan_rotsphere(n);
    P->P[0] = ftamd_plan_leg2cheb(1, 0, n);
    P->P[1] = ftamd_plan_ultra2ultra(1, 0, n, 1.5, 1.0);
    P->Pinv[0] = ftamd_plan_cheb2leg(0, 1, n);
    P->Pinv[1] = ftamd_plan_ultra2ultra(0, 1, n, 1.0, 1.5);
    return plan_done(P);
}

/* drivers.c:901-919 */
ft_harmonic_plan *ft_plan_tri2cheb(const int n, const double alpha, const double beta, const double gamma) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = plan_alloc(2);
    if (!P) return plan_done(NULL);
    P->RP[0] = ft_plan_rottriangle(n, alpha, beta, gamma);
    P->P[0] = ftamd_plan_jac2jac(1, 1, n, beta + gamma + 1.0, alpha, -0.5, -0.5);
    P->P[1] = ftamd_plan_jac2jac(1, 1, n, gamma, beta, -0.5, -0.5);
    P->Pinv[0] = ftamd_plan_jac2jac(1, 1, n, -0.5, -0.5, beta + gamma + 1.0, alpha);
    P->Pinv[1] = ftamd_plan_jac2jac(1, 1, n, -0.5, -0.5, gamma, beta);
    P->alpha = alpha;
    P->beta = beta;
    P->gamma = gamma;
    return plan_done(P);
}

/* drivers.c:1123-1141 */
ft_harmonic_plan *ft_plan_rectdisk2cheb(const int n, const double beta) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = plan_alloc(3);
    if (!P) return plan_done(NULL);
    P->RP[0] = ft_plan_rotrectdisk(n, beta);
    P->P[0] = ftamd_plan_ultra2cheb(1, 0, n, beta + 1.0);
    P->P[1] = ftamd_plan_ultra2ultra(1, 0, n, beta + 2.0, 1.0);
    P->P[2] = ftamd_plan_ultra2cheb(1, 0, n, beta + 0.5);
    P->Pinv[0] = ftamd_plan_cheb2ultra(0, 1, n, beta + 1.0);
    P->Pinv[1] = ftamd_plan_ultra2ultra(0, 1, n, 1.0, beta + 2.0);
    P->Pinv[2] = ftamd_plan_cheb2ultra(0, 1, n, beta + 0.5);
    P->beta = beta;
    return plan_done(P);
}

/* drivers.c:959-985: P[0] rows of the Chebyshev side and P[1] scaled by
 * 2^((alpha+beta+2)/2) (times sqrt(2/pi) for P[1]), the inverses by the
 * reciprocals, on and above the diagonal. */
ft_harmonic_plan *ft_plan_disk2cxf(const int n, const double alpha, const double beta) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = plan_alloc(2);
    if (!P) return plan_done(NULL);
    P->RP[0] = ft_plan_rotdisk(n, alpha, beta);
    P->P[0] = ftamd_plan_jac2cheb(1, 0, n, beta, alpha);
    P->P[1] = ftamd_plan_jac2jac(1, 1, n, beta, alpha + 1.0, -0.5, 0.5);
    P->Pinv[0] = ftamd_plan_cheb2jac(0, 1, n, beta, alpha);
    P->Pinv[1] = ftamd_plan_jac2jac(1, 1, n, -0.5, 0.5, beta, alpha + 1.0);
    if (P->P[0] && P->P[1] && P->Pinv[0] && P->Pinv[1]) {
        const double cst0 = pow(2.0, 0.5 * (alpha + beta + 2.0));
        const double cst1 = cst0 * 0.797884560802865355; /* sqrt(2/pi), ftinternal.h:58 */
        const double cst0inv = pow(2.0, -0.5 * (alpha + beta + 2.0));
        const double cst1inv = cst0inv * 1.253314137315500251; /* sqrt(pi/2), ftinternal.h:57 */
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++) {
                P->P[0][i + (size_t)j * n] *= cst0;
                P->P[1][i + (size_t)j * n] *= cst1;
                P->Pinv[0][i + (size_t)j * n] *= cst0inv;
                P->Pinv[1][i + (size_t)j * n] *= cst1inv;
            }
    }
    P->alpha = alpha;
    P->beta = beta;
    return plan_done(P);
}

/* drivers.c:1023-1071 with banded_source.c:484-499: MP[1] holds the upper
 * Cholesky factor of t I + J, J the normalized Jacobi(beta, alpha)
 * multiplication operator (banded_source.c:1797-1820), in the reference's
 * order: the subdiagonal entry divided by the pivot before its square root,
 * the next pivot less that quotient times the superdiagonal entry, the
 * superdiagonal entry divided by the square root. */
static ft_modified_plan *annulus_factor(int n, double alpha, double beta, double rho) {
    ft_modified_plan *F = calloc(1, sizeof(ft_modified_plan));
    double *R = calloc((size_t)(2 * n), sizeof(double)); /* the last entry is padding */
    if (!F || !R) {
        free(F);
        free(R);
        return NULL;
    }
    F->n = n;
    F->nu = 1;
    F->b = 1;
    F->R = R;
    const double a = beta, b = alpha; /* the operator's (alpha, beta) are the plan's (beta, alpha) */
    const double t = 1.0 + 2.0 * rho * rho / (1.0 - rho * rho);
    double *dg = R, *up = R + n;
    double p = t + (b - a) / (a + b + 2); /* pivot of row 0 */
    for (int j = 0; j < n; j++) {
        if (j + 1 < n) {
            /* entries (j+1, j) and (j, j+1) of J; equal formulas but for j = 0 */
            const double lo = j == 0 ? 2 * sqrt(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)))
                                     : 2 * sqrt(((j + 1) * (j + a + 1) * (j + b + 1) * (j + a + b + 1)) /
                                                ((2 * j + a + b + 1) * (2 * j + a + b + 2) * (2 * j + a + b + 2) *
                                                 (2 * j + a + b + 3)));
            const int k = j + 1;
            const double upj = k == 1 ? 2 * sqrt(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)))
                                      : 2 * sqrt((k * (k + a) * (k + b) * (k + a + b)) /
                                                 ((2 * k + a + b - 1) * (2 * k + a + b) * (2 * k + a + b) *
                                                  (2 * k + a + b + 1)));
            const double diag = ((b - a) * (a + b)) / ((2 * k + a + b) * (2 * k + a + b + 2));
            const double mult = lo / p;
            dg[j] = sqrt(p);
            up[j] = upj / dg[j];
            p = (t + diag) - mult * upj;
        } else {
            dg[j] = sqrt(p);
        }
    }
    return F;
}

/* drivers.c:1023-1071: gamma = 0 only. */
ft_harmonic_plan *ft_plan_ann2cxf(const int n, const double alpha, const double beta, const double gamma,
                                  const double rho) {
    if (n < 1 || gamma != 0.0 || !(rho >= 0.0 && rho < 1.0)) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = plan_alloc(1);
    if (!P) return plan_done(NULL);
    P->RP[0] = ft_plan_rotannulus(n, alpha, beta, gamma, rho);
    P->P[0] = ftamd_plan_jac2cheb(1, 0, n, beta, alpha);
    P->Pinv[0] = ftamd_plan_cheb2jac(0, 1, n, beta, alpha);
    P->MP = calloc(2, sizeof(ft_modified_plan *));
    if (P->MP) {
        P->NMP = 2;
        P->MP[0] = calloc(1, sizeof(ft_modified_plan));
        if (P->MP[0]) P->MP[0]->n = n; /* the identity: R NULL, b 0 */
        P->MP[1] = annulus_factor(n, alpha, beta, rho);
    }
    if (!P->MP || !P->MP[0] || !P->MP[1]) {
        ft_destroy_harmonic_plan(P);
        return plan_done(NULL);
    }
    if (P->P[0] && P->Pinv[0]) {
        const double cst = sqrt(2.0) * pow(2.0 / (1.0 - rho * rho), 0.5 * (alpha + beta + gamma + 1.0));
        const double cstinv = 1.0 / cst;
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++) {
                P->P[0][i + (size_t)j * n] *= cst;
                P->Pinv[0][i + (size_t)j * n] *= cstinv;
            }
    }
    P->alpha = alpha;
    P->beta = beta;
    P->gamma = gamma;
    P->rho = rho;
    return plan_done(P);
}

/* Device copies of the connection matrices (caller holds the lock). */
static int upload(ft_harmonic_plan *P) {
    struct ftamd_harmonic_dev *d = P->dev;
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) return rc;
    if (d->device >= 0) return d->device == dev ? FTAMD_OK : FTAMD_EINVAL;
    const int n = P->RP[0]->n;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    for (int k = 0; rc == FTAMD_OK && k < P->NP; k++) {
        rc = ftamd_hip_malloc((void **)&d->dP[k], bytes);
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d->dP[k], P->P[k], bytes, NULL);
        if (rc == FTAMD_OK) rc = ftamd_hip_malloc((void **)&d->dPinv[k], bytes);
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d->dPinv[k], P->Pinv[k], bytes, NULL);
    }
    if (rc == FTAMD_OK && P->NMP == 2) {
        const size_t rbytes = sizeof(double) * (size_t)(2 * n);
        rc = ftamd_hip_malloc((void **)&d->dR, rbytes);
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d->dR, P->MP[1]->R, rbytes, NULL);
    }
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    if (rc == FTAMD_OK) d->device = dev;
    return rc;
}

static int reserve(void **buf, size_t *have, size_t bytes) {
    if (*have >= bytes) return FTAMD_OK;
    ftamd_hip_device_sync(); /* an earlier launch, on any stream, may still use the old buffer */
    ftamd_hip_free(*buf);
    *buf = NULL;
    *have = 0;
    int rc = ftamd_hip_malloc(buf, bytes);
    if (rc == FTAMD_OK) *have = bytes;
    return rc;
}

/* The rotation executor kind of each transform. */
static int rotation_kind(int transform) {
    switch (transform) {
    case FT_GPU_SPH2FOURIER:
    case FT_GPU_FOURIER2SPH: return FT_GPU_SPH;
    case FT_GPU_SPHV2FOURIER:
    case FT_GPU_FOURIER2SPHV: return FT_GPU_SPHV;
    case FT_GPU_DISK2CXF:
    case FT_GPU_CXF2DISK:
    case FT_GPU_ANN2CXF:
    case FT_GPU_CXF2ANN: return FT_GPU_DISK;
    case FT_GPU_TRI2CHEB:
    case FT_GPU_CHEB2TRI: return FT_GPU_TRI;
    case FT_GPU_RECTDISK2CHEB:
    case FT_GPU_CHEB2RECTDISK: return FT_GPU_RECTDISK;
    }
    return -1;
}

static int valid(int transform, char TRANS, const ft_harmonic_plan *P, int N, int M) {
    const int rk = rotation_kind(transform);
    if (!P || !P->dev || !P->RP || !P->RP[0] || rk < 0 || (TRANS != 'N' && TRANS != 'T')) return 0;
    /* the plan's rotation tables must be the transform's geometry */
    const int geo = rk == FT_GPU_TRI ? FTAMD_TRI : rk == FT_GPU_DISK ? FTAMD_DISK : FTAMD_SPH;
    const int ann = transform == FT_GPU_ANN2CXF || transform == FT_GPU_CXF2ANN;
    const int np = rk == FT_GPU_RECTDISK ? 3 : ann ? 1 : 2; /* matrix pairs the transform uses */
    /* the annulus's transforms need its plan (rotations and factor), the disk's a disk plan */
    if (ann != (P->RP[0]->family == FTAMD_FAMILY_ANNULUS) || (ann && (P->NMP != 2 || !P->MP))) return 0;
    return P->RP[0]->kind == geo && P->NP >= np && N == P->RP[0]->n && M >= 1 && ftamd_valid_M(rk, N, M);
}

/* One product into the work buffer and back into dA. */
static int product(struct ftamd_harmonic_dev *d, FtamdTrmmLaunch *t, void *dA, size_t bytes, void *stream) {
    t->X = dA;
    t->Y = d->dY;
    int rc = ftamd_hip_trmm_launch(t, stream);
    if (rc == FTAMD_OK) rc = ftamd_hip_d2d(dA, d->dY, bytes, stream);
    return rc;
}

/* The products of the column-class transforms (sph, sphv, disk). */
static int class_products(struct ftamd_harmonic_dev *d, double *const *H, int vec, int trans, int n, int M, void *dA,
                          size_t bytes, void *stream) {
    FtamdTrmmLaunch t;
    memset(&t, 0, sizeof(t));
    for (int c = 0; c < 4; c++) {
        const int even = (c == 0 || c == 3); /* columns of even-parity orders (sph and disk layouts) */
        t.U[c] = H[even != vec ? 0 : 1];
        const int count = M > c ? (M - c + 3) / 4 : 0;
        t.ctiles[c] = (count + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS;
    }
    t.n = n;
    t.M = M;
    t.cstride = 4;
    t.ncls = 4;
    t.trans = trans;
    return product(d, &t, dA, bytes, stream);
}

/* One triangle product on every column (left) or every row (right). */
static int tri_product(struct ftamd_harmonic_dev *d, const double *U, int right, int trans, int n, int M, void *dA,
                       size_t bytes, void *stream) {
    FtamdTrmmLaunch t;
    memset(&t, 0, sizeof(t));
    t.U[0] = U;
    t.ctiles[0] = (M + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS;
    t.n = n;
    t.M = M;
    t.cstride = 1;
    t.ncls = 1;
    t.trans = trans;
    t.right = right;
    return product(d, &t, dA, bytes, stream);
}

/* The triangle's transforms (drivers.c:921-957): tri2cheb 'N' is hi2lo, P[0]
 * from the left unless (alpha, beta + gamma) = (-1/2, -3/2), P[1]^T from the
 * right unless beta = gamma = -1/2, then the Chebyshev normalization;
 * cheb2tri inverts it. */
static int run_tri(ft_harmonic_plan *P, int inverse, char TRANS, void *dA, int M, size_t bytes, void *stream) {
    struct ftamd_harmonic_dev *d = P->dev;
    const int n = P->RP[0]->n;
    const int use0 = (P->beta + P->gamma != -1.5) || (P->alpha != -0.5);
    const int use1 = (P->gamma != -0.5) || (P->beta != -0.5);
    const double s1 = inverse ? 1.41421356237309504880 : 0.70710678118654752440; /* M_SQRT2, M_SQRT1_2 */
    const double s2 = inverse ? 1.57079632679489661923 : 0.63661977236758134308; /* M_PI_2, M_2_PI */
    double *const *H = inverse ? d->dPinv : d->dP;
    int rc = FTAMD_OK;
    if (!inverse && TRANS == 'N') {
        rc = ft_gpu_execute(P->RP[0], FT_GPU_TRI, 0, FT_GPU_FP64, dA, M, stream);
        if (rc == FTAMD_OK && use0) rc = tri_product(d, H[0], 0, 0, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK && use1) rc = tri_product(d, H[1], 1, 1, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ftamd_hip_cheb_scale(dA, n, M, s1, s2, stream);
    } else if (!inverse) {
        rc = ftamd_hip_cheb_scale(dA, n, M, s1, s2, stream);
        if (rc == FTAMD_OK && use1) rc = tri_product(d, H[1], 1, 0, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK && use0) rc = tri_product(d, H[0], 0, 1, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ft_gpu_execute(P->RP[0], FT_GPU_TRI, 1, FT_GPU_FP64, dA, M, stream);
    } else if (TRANS == 'N') {
        rc = ftamd_hip_cheb_scale(dA, n, M, s1, s2, stream);
        if (rc == FTAMD_OK && use1) rc = tri_product(d, H[1], 1, 1, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK && use0) rc = tri_product(d, H[0], 0, 0, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ft_gpu_execute(P->RP[0], FT_GPU_TRI, 1, FT_GPU_FP64, dA, M, stream);
    } else {
        rc = ft_gpu_execute(P->RP[0], FT_GPU_TRI, 0, FT_GPU_FP64, dA, M, stream);
        if (rc == FTAMD_OK && use0) rc = tri_product(d, H[0], 0, 1, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK && use1) rc = tri_product(d, H[1], 1, 0, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ftamd_hip_cheb_scale(dA, n, M, s1, s2, stream);
    }
    return rc;
}

/* The rectangle-disk's transforms (drivers.c:1143-1169): rectdisk2cheb 'N'
 * is hi2lo, P[0] from the left on the even columns and P[1] on the odd ones
 * (column classes j mod 2), then P[2]^T from the right on every row;
 * cheb2rectdisk inverts it with Pinv in reverse order. */
static int run_rectdisk(ft_harmonic_plan *P, int inverse, char TRANS, void *dA, int M, size_t bytes, void *stream) {
    struct ftamd_harmonic_dev *d = P->dev;
    const int n = P->RP[0]->n;
    double *const *H = inverse ? d->dPinv : d->dP;
    FtamdTrmmLaunch t;
    memset(&t, 0, sizeof(t));
    t.n = n;
    t.M = M;
    t.cstride = 2;
    t.ncls = 2;
    for (int c = 0; c < 2; c++) {
        t.U[c] = H[c];
        const int count = M > c ? (M - c + 1) / 2 : 0;
        t.ctiles[c] = (count + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS;
    }
    const int fwd = !inverse ? TRANS == 'N' : TRANS == 'T'; /* rotations before the products */
    int rc = FTAMD_OK;
    if (fwd && !inverse) {
        rc = ft_gpu_execute(P->RP[0], FT_GPU_RECTDISK, 0, FT_GPU_FP64, dA, M, stream);
        t.trans = 0;
        if (rc == FTAMD_OK) rc = product(d, &t, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = tri_product(d, H[2], 1, 1, n, M, dA, bytes, stream);
    } else if (!inverse) {
        rc = tri_product(d, H[2], 1, 0, n, M, dA, bytes, stream);
        t.trans = 1;
        if (rc == FTAMD_OK) rc = product(d, &t, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ft_gpu_execute(P->RP[0], FT_GPU_RECTDISK, 1, FT_GPU_FP64, dA, M, stream);
    } else if (TRANS == 'N') {
        rc = tri_product(d, H[2], 1, 1, n, M, dA, bytes, stream);
        t.trans = 0;
        if (rc == FTAMD_OK) rc = product(d, &t, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ft_gpu_execute(P->RP[0], FT_GPU_RECTDISK, 1, FT_GPU_FP64, dA, M, stream);
    } else {
        rc = ft_gpu_execute(P->RP[0], FT_GPU_RECTDISK, 0, FT_GPU_FP64, dA, M, stream);
        t.trans = 1;
        if (rc == FTAMD_OK) rc = product(d, &t, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = tri_product(d, H[2], 1, 0, n, M, dA, bytes, stream);
    }
    return rc;
}

/* The annulus's transforms (drivers.c:1073-1121): ann2cxf 'N' is hi2lo,
 * R^{-1} on the columns j = 1, 2 mod 4 (the identity MP[0] on j = 0, 3 is
 * skipped), P[0] on every column; 'T' is P[0]^T, R^{-T}, lo2hi.  cxf2ann 'N'
 * is Pinv[0], R, lo2hi; 'T' is hi2lo, R^T, Pinv[0]^T. */
static int run_annulus(ft_harmonic_plan *P, int inverse, char TRANS, void *dA, int M, size_t bytes, void *stream) {
    struct ftamd_harmonic_dev *d = P->dev;
    const int n = P->RP[0]->n, trans = TRANS == 'T';
    const double *H = inverse ? d->dPinv[0] : d->dP[0];
    const int rot_first = inverse ? trans : !trans;
    int rc = FTAMD_OK;
    if (rot_first) {
        rc = ft_gpu_execute(P->RP[0], FT_GPU_DISK, 0, FT_GPU_FP64, dA, M, stream);
        if (rc == FTAMD_OK) rc = ftamd_hip_bidiag(dA, n, M, d->dR, !inverse, trans, stream);
        if (rc == FTAMD_OK) rc = tri_product(d, H, 0, trans, n, M, dA, bytes, stream);
    } else {
        rc = tri_product(d, H, 0, trans, n, M, dA, bytes, stream);
        if (rc == FTAMD_OK) rc = ftamd_hip_bidiag(dA, n, M, d->dR, !inverse, trans, stream);
        if (rc == FTAMD_OK) rc = ft_gpu_execute(P->RP[0], FT_GPU_DISK, 1, FT_GPU_FP64, dA, M, stream);
    }
    return rc;
}

/* The transform on device matrix dA, asynchronously on `stream` (caller
 * holds the lock and has uploaded the matrices). */
static int run(ft_harmonic_plan *P, int transform, char TRANS, void *dA, int M, void *stream) {
    struct ftamd_harmonic_dev *d = P->dev;
    const int n = P->RP[0]->n;
    const int rk = rotation_kind(transform), inverse = transform & 1;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)M;
    int rc = reserve(&d->dY, &d->y_bytes, bytes);
    if (rc != FTAMD_OK) return rc;
    if (rk == FT_GPU_TRI) return run_tri(P, inverse, TRANS, dA, M, bytes, stream);
    if (rk == FT_GPU_RECTDISK) return run_rectdisk(P, inverse, TRANS, dA, M, bytes, stream);
    if (transform == FT_GPU_ANN2CXF || transform == FT_GPU_CXF2ANN)
        return run_annulus(P, inverse, TRANS, dA, M, bytes, stream);
    /* drivers.c:833-899, 987-1021: rotations first for the forward 'N' and the inverse 'T' */
    const int rot_first = inverse ? TRANS == 'T' : TRANS == 'N';
    const int rot_dir = rot_first ? 0 : 1;
    if (rot_first) rc = ft_gpu_execute(P->RP[0], rk, rot_dir, FT_GPU_FP64, dA, M, stream);
    if (rc == FTAMD_OK)
        rc = class_products(d, inverse ? d->dPinv : d->dP, rk == FT_GPU_SPHV, TRANS == 'T', n, M, dA, bytes, stream);
    if (rc == FTAMD_OK && !rot_first) rc = ft_gpu_execute(P->RP[0], rk, rot_dir, FT_GPU_FP64, dA, M, stream);
    return rc;
}

int ft_gpu_execute_harmonic(int transform, char TRANS, const ft_harmonic_plan *P_, void *dA, int N, int M,
                            void *stream) {
    ft_harmonic_plan *P = (ft_harmonic_plan *)P_;
    if (!valid(transform, TRANS, P, N, M) || !dA) return FT_GPU_EINVAL;
    pthread_mutex_lock(&P->dev->lock);
    int rc = upload(P);
    if (rc == FTAMD_OK) rc = run(P, transform, TRANS, dA, M, stream);
    pthread_mutex_unlock(&P->dev->lock);
    return rc;
}

/* The reference entry points: synchronous, host or device A. */
static void harmonic(int transform, char TRANS, const ft_harmonic_plan *P_, double *A, int N, int M) {
    ft_harmonic_plan *P = (ft_harmonic_plan *)P_;
    if (!valid(transform, TRANS, P, N, M) || !A) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    const size_t bytes = sizeof(double) * (size_t)N * (size_t)M;
    struct ftamd_harmonic_dev *d = P->dev;
    pthread_mutex_lock(&d->lock);
    rc = upload(P);
    void *dA = A;
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) {
        rc = reserve(&d->dA, &d->a_bytes, bytes);
        dA = d->dA;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(dA, A, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = run(P, transform, TRANS, dA, M, NULL);
    /* waits and reads the rotation launch's error word */
    if (rc == FTAMD_OK) rc = ft_gpu_synchronize(NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, dA, bytes, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&d->lock);
    ftamd_set_error(rc);
}

/* drivers.c:833-848 */
void ft_execute_sph2fourier(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_SPH2FOURIER, TRANS, P, A, N, M);
}
/* drivers.c:850-865 */
void ft_execute_fourier2sph(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_FOURIER2SPH, TRANS, P, A, N, M);
}
/* drivers.c:867-882 */
void ft_execute_sphv2fourier(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_SPHV2FOURIER, TRANS, P, A, N, M);
}
/* drivers.c:884-899 */
void ft_execute_fourier2sphv(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_FOURIER2SPHV, TRANS, P, A, N, M);
}
/* drivers.c:921-938 */
void ft_execute_tri2cheb(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_TRI2CHEB, TRANS, P, A, N, M);
}
/* drivers.c:940-957 */
void ft_execute_cheb2tri(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_CHEB2TRI, TRANS, P, A, N, M);
}
/* drivers.c:1143-1156 */
void ft_execute_rectdisk2cheb(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_RECTDISK2CHEB, TRANS, P, A, N, M);
}
/* drivers.c:1158-1171 */
void ft_execute_cheb2rectdisk(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_CHEB2RECTDISK, TRANS, P, A, N, M);
}
/* drivers.c:987-1003 */
void ft_execute_disk2cxf(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_DISK2CXF, TRANS, P, A, N, M);
}
/* drivers.c:1005-1021 */
void ft_execute_cxf2disk(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_CXF2DISK, TRANS, P, A, N, M);
}
/* drivers.c:1073-1096 */
void ft_execute_ann2cxf(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_ANN2CXF, TRANS, P, A, N, M);
}
/* drivers.c:1098-1121 */
void ft_execute_cxf2ann(const char TRANS, const ft_harmonic_plan *P, double *A, const int N, const int M) {
    harmonic(FT_GPU_CXF2ANN, TRANS, P, A, N, M);
}
