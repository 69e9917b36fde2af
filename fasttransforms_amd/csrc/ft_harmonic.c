/*
 * ft_harmonic.c -- the spherical-harmonic <-> bivariate Fourier transforms
 * (drivers.c:793-899): the rotation stage on the GPU followed (or preceded)
 * by the strided upper-triangular products with the 1D connection matrices,
 * which run as one matrix-core launch (ft_trmm.hip) instead of four BLAS
 * calls.
 *
 * TRANS 'N' of sph2fourier is hi2lo then the products with P; 'T' is the
 * transposed products then lo2hi; fourier2sph uses Pinv with lo2hi after
 * ('N') or hi2lo before ('T').  Column classes j mod 4: the sphere's orders
 * of even parity sit in j = 0, 3 and take P[0], the odd ones in j = 1, 2 and
 * take P[1]; the vector-field layout (columns 2m+1, 2m+2) swaps them
 * (drivers.c:867-899).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

struct ftamd_harmonic_dev {
    pthread_mutex_t lock; /* held for a whole call: guards the buffers below */
    int device;           /* HIP device of the uploads, -1 before */
    double *dP[2], *dPinv[2];
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
        for (int k = 0; k < 2; k++) {
            ftamd_hip_free(d->dP[k]);
            ftamd_hip_free(d->dPinv[k]);
        }
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
    free(P);
}

/* drivers.c:816-831 */
ft_harmonic_plan *ft_plan_sph2fourier(const int n) {
    if (n < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    ft_harmonic_plan *P = calloc(1, sizeof(ft_harmonic_plan));
    if (!P) goto oom;
    P->dev = calloc(1, sizeof(struct ftamd_harmonic_dev));
    if (P->dev) {
        pthread_mutex_init(&P->dev->lock, NULL);
        P->dev->device = -1;
    }
    P->RP = calloc(1, sizeof(ft_rotation_plan *));
    P->P = calloc(2, sizeof(double *));
    P->Pinv = calloc(2, sizeof(double *));
    if (!P->dev || !P->RP || !P->P || !P->Pinv) goto oom;
    P->NRP = 1;
    P->NP = 2;
    P->RP[0] = ft_plan_rotsphere(n);
    P->P[0] = ftamd_plan_leg2cheb(1, 0, n);
    P->P[1] = ftamd_plan_ultra2ultra(1, 0, n, 1.5, 1.0);
    P->Pinv[0] = ftamd_plan_cheb2leg(0, 1, n);
    P->Pinv[1] = ftamd_plan_ultra2ultra(0, 1, n, 1.0, 1.5);
    if (!P->RP[0] || !P->P[0] || !P->P[1] || !P->Pinv[0] || !P->Pinv[1]) goto oom;
    return P;
oom:
    ft_destroy_harmonic_plan(P);
    ftamd_set_error(FT_GPU_ENOMEM);
    return NULL;
}

/* Device copies of the connection matrices (caller holds the lock). */
static int upload(ft_harmonic_plan *P) {
    struct ftamd_harmonic_dev *d = P->dev;
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) return rc;
    if (d->device >= 0) return d->device == dev ? FTAMD_OK : FTAMD_EINVAL;
    const int n = P->RP[0]->n;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    for (int k = 0; rc == FTAMD_OK && k < 2; k++) {
        rc = ftamd_hip_malloc((void **)&d->dP[k], bytes);
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d->dP[k], P->P[k], bytes, NULL);
        if (rc == FTAMD_OK) rc = ftamd_hip_malloc((void **)&d->dPinv[k], bytes);
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d->dPinv[k], P->Pinv[k], bytes, NULL);
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

static int valid(int transform, char TRANS, const ft_harmonic_plan *P, int N, int M) {
    if (!P || !P->dev || !P->RP || !P->RP[0] || transform < 0 || transform > 3 || (TRANS != 'N' && TRANS != 'T'))
        return 0;
    const int vec = transform >= FT_GPU_SPHV2FOURIER;
    return N == P->RP[0]->n && M >= 1 && ftamd_valid_M(vec ? FT_GPU_SPHV : FT_GPU_SPH, N, M);
}

/* The transform on device matrix dA, asynchronously on `stream` (caller
 * holds the lock and has uploaded the matrices). */
static int run(ft_harmonic_plan *P, int transform, char TRANS, void *dA, int M, void *stream) {
    struct ftamd_harmonic_dev *d = P->dev;
    const int n = P->RP[0]->n;
    const int vec = transform >= FT_GPU_SPHV2FOURIER, inverse = transform & 1;
    const int rot_kind = vec ? FT_GPU_SPHV : FT_GPU_SPH;
    /* drivers.c:833-899: rotations first for sph2fourier 'N' and fourier2sph 'T' */
    const int rot_first = inverse ? TRANS == 'T' : TRANS == 'N';
    const int rot_dir = rot_first ? 0 : 1;
    double *const *H = inverse ? d->dPinv : d->dP;
    FtamdTrmmLaunch t;
    memset(&t, 0, sizeof(t));
    for (int c = 0; c < 4; c++) {
        const int even = (c == 0 || c == 3); /* columns of even-parity orders (sph layout) */
        t.U[c] = H[even != vec ? 0 : 1];
        const int count = M > c ? (M - c + 3) / 4 : 0;
        t.ctiles[c] = (count + FTAMD_TRMM_COLS - 1) / FTAMD_TRMM_COLS;
    }
    t.X = dA;
    t.n = n;
    t.M = M;
    t.cstride = 4;
    t.ncls = 4;
    t.trans = TRANS == 'T';
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)M;
    int rc = reserve(&d->dY, &d->y_bytes, bytes);
    t.Y = d->dY;
    if (rc == FTAMD_OK && rot_first) rc = ft_gpu_execute(P->RP[0], rot_kind, rot_dir, FT_GPU_FP64, dA, M, stream);
    if (rc == FTAMD_OK) rc = ftamd_hip_trmm_launch(&t, stream);
    if (rc == FTAMD_OK) rc = ftamd_hip_d2d(dA, d->dY, bytes, stream);
    if (rc == FTAMD_OK && !rot_first) rc = ft_gpu_execute(P->RP[0], rot_kind, rot_dir, FT_GPU_FP64, dA, M, stream);
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
