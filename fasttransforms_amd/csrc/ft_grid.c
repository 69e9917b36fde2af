/*
 * ft_grid.c -- the sphere's synthesis and analysis plans and executors
 * (fftw.c:62-270, fasttransforms.h:466-493) on the GPU.
 *
 * A plan fixes the theta transforms of its column classes (synthesis: DCT-III
 * on the cosine class, DST-III on the sine class; analysis: DCT-II, DST-II;
 * the vector-field plans swap the classes) and the direction of the phi DFT
 * (synthesis c2r, analysis r2c).  An executor fixes the order of the steps,
 * the scalings and which rows it doubles or halves (fftw.c:126-270):
 *
 *   synthesis 'N' and analysis 'T' take a synthesis plan: fix-ups, theta,
 *   scalings, phi;  synthesis 'T' and analysis 'N' take an analysis plan:
 *   phi, scalings, theta, fix-ups (the reference's tests pair them so,
 *   test_fftw.c:52-64).  A plan of the other kind sets FT_GPU_EINVAL.
 *
 * X is the reference's N x M column-major array (M odd: fftw.c:40-60 reads
 * column 2*(M/2) for every M), host or device, overwritten in place.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct ft_sphere_fftw_plan_s {
    int N, M;
    int synthesis; /* 1: DCT-III / DST-III and c2r; 0: DCT-II / DST-II and r2c */
    int sphv;      /* 0 the sphere's, 1 the vector field's column classes, 2 the disk's
                      (DCT-IV, REDFT11, on the sine class in both plan kinds), 3 the
                      annulus's (one class, radial weights) */
    double rho;    /* the annulus's inner radius */
    int device;
    FtamdGrid *g;
    void *d_stage; /* staging buffer of host-pointer calls */
    size_t stage_bytes;
    pthread_mutex_t lock; /* held for the whole of a call */
};

/* the reference's constants (ftinternal.h:48-56) */
#define FT_SQRT2 1.4142135623730950488
#define FT_SQRT1_2 0.70710678118654752440
#define FT_1_4_SQRT_PI 0.141047395886939072
#define FT_1_2_SQRT_PI 0.282094791773878143
#define FT_4_SQRT_PI 7.089815403622064109
#define FT_2_SQRT_PI 3.544907701811032055

static ft_sphere_fftw_plan *plan_grid(int N, int M, int synthesis, int sphv, double rho) {
    /* the disk's and the annulus's plans are the same object */
    if (N < 1 || M < 1 || M % 2 == 0 || (sphv == 3 && !(rho >= 0.0 && rho < 1.0))) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return NULL;
    }
    ft_sphere_fftw_plan *P = calloc(1, sizeof(ft_sphere_fftw_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    P->N = N;
    P->M = M;
    P->synthesis = synthesis;
    P->sphv = sphv;
    P->rho = rho;
    P->device = dev;
    pthread_mutex_init(&P->lock, NULL);
    double *w = NULL;
    if (sphv == 3) {
        /* fftw.c:540-548: the radial weights at theta = (2n+1) pi / 2N */
        w = malloc(sizeof(double) * (size_t)N);
        if (!w) {
            pthread_mutex_destroy(&P->lock);
            free(P);
            ftamd_set_error(FT_GPU_ENOMEM);
            return NULL;
        }
        for (int n = 0; n < N; n++) {
            const double t = (N - n - 0.5) * M_PI / (2.0 * N);
            const double ct = sin(t), st = cos(t);
            w[n] = sqrt(2.0 * (ct * ct + rho * rho * st * st) / (1.0 - rho * rho));
        }
    }
    rc = ftamd_grid_create(N, M, synthesis, sphv, w, &P->g);
    free(w);
    if (rc != FTAMD_OK) {
        pthread_mutex_destroy(&P->lock);
        free(P);
        ftamd_set_error(rc);
        return NULL;
    }
    return P;
}

/* fasttransforms.h:481-485 (fftw.c:108-126).  flags (FFTW's planner flags)
 * are accepted and ignored. */
ft_sphere_fftw_plan *ft_plan_sph_synthesis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 1, 0, 0.0);
}
ft_sphere_fftw_plan *ft_plan_sph_analysis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 0, 0, 0.0);
}
ft_sphere_fftw_plan *ft_plan_sphv_synthesis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 1, 1, 0.0);
}
ft_sphere_fftw_plan *ft_plan_sphv_analysis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 0, 1, 0.0);
}

/* fasttransforms.h:477 (fftw.c:66-74) */
void ft_destroy_sphere_fftw_plan(ft_sphere_fftw_plan *P) {
    if (!P) return;
    ftamd_grid_destroy(P->g);
    ftamd_hip_free(P->d_stage);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* The step order, fix-ups and scalings of executor `which` (0 sph synthesis,
 * 1 sph analysis, 2 sphv synthesis, 3 sphv analysis, 4 disk synthesis, 5
 * disk analysis) and TRANS; *need_syn: the plan kind it runs on.  The disk's
 * executors (fftw.c:460-520) are the sphere's without the last-row fix-ups
 * of the sine class. */
typedef struct {
    int need_syn, exec_sphv, wmode;
    double f0, f1, a, b;
} grid_step;

static int grid_steps(int which, char TRANS, int N, int M, grid_step *s) {
    const int analysis = which & 1, sphv = (which >> 1) == 1, disk = (which >> 1) == 2, ann = (which >> 1) == 3;
    if (TRANS != 'N' && TRANS != 'T') return 0;
    const int trans = TRANS == 'T';
    s->exec_sphv = ann ? 3 : sphv;
    /* the annulus's weights (fftw.c:583-655): multiplied by synthesis, divided by analysis */
    s->wmode = ann ? (analysis ? 2 : 1) : 0;
    if (!analysis && !trans) { /* fftw.c:128-146: x2 row 0 (cosine class), III, x 1/(4 sqrt(pi)), col 0 x sqrt 2, c2r */
        s->need_syn = 1, s->f0 = 2.0, s->f1 = 1.0, s->a = FT_1_4_SQRT_PI, s->b = FT_SQRT2;
    } else if (!analysis) { /* fftw.c:148-163: r2c, x 1/(2 sqrt(pi)), col 0 x 1/sqrt 2, II, x1/2 last row (sine class) */
        s->need_syn = 0, s->f0 = 1.0, s->f1 = 0.5, s->a = FT_1_2_SQRT_PI, s->b = FT_SQRT1_2;
    } else if (!trans) { /* fftw.c:167-183: r2c, x 4 sqrt(pi)/(2NM), col 0 x 1/sqrt 2, II, x1/2 row 0 (cosine class) */
        s->need_syn = 0, s->f0 = 0.5, s->f1 = 1.0, s->a = FT_4_SQRT_PI / (2.0 * N * M), s->b = FT_SQRT1_2;
    } else { /* fftw.c:185-200: x2 last row (sine class), III, x 2 sqrt(pi)/(2NM), col 0 x sqrt 2, c2r */
        s->need_syn = 1, s->f0 = 1.0, s->f1 = 2.0, s->a = FT_2_SQRT_PI / (2.0 * N * M), s->b = FT_SQRT2;
    }
    if (disk || ann) s->f1 = 1.0;
    return 1;
}

/* The executor's geometry must be the plan's (a disk plan's sine class is
 * another transform) and its kind the one the executor takes. */
static int valid_grid(const ft_sphere_fftw_plan *P, int which, int N, int M, const grid_step *s) {
    const int geo = which >> 1;
    return P && P->N == N && P->M == M && P->synthesis == s->need_syn && (geo == 2) == (P->sphv == 2) &&
           (geo == 3) == (P->sphv == 3);
}

int ft_gpu_execute_sph_grid(int which, char TRANS, const ft_sphere_fftw_plan *P_, double *dX, int N, int M,
                            void *stream) {
    ft_sphere_fftw_plan *P = (ft_sphere_fftw_plan *)P_;
    grid_step s;
    if (which < 0 || which > 7 || !grid_steps(which, TRANS, N, M, &s) || !valid_grid(P, which, N, M, &s) || !dX)
        return FT_GPU_EINVAL;
    pthread_mutex_lock(&P->lock);
    int rc = ftamd_grid_run(P->g, dX, s.exec_sphv, s.f0, s.f1, s.a, s.b, s.wmode, stream);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

static void execute_grid(int which, char TRANS, const ft_sphere_fftw_plan *P_, double *X, int N, int M) {
    ft_sphere_fftw_plan *P = (ft_sphere_fftw_plan *)P_;
    grid_step s;
    if (!grid_steps(which, TRANS, N, M, &s) || !valid_grid(P, which, N, M, &s) || !X) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(X, &pk);
    const size_t bytes = (size_t)N * (size_t)M * sizeof(double);
    pthread_mutex_lock(&P->lock);
    void *d = X;
    if (pk != FTAMD_PTR_DEVICE) {
        if (P->stage_bytes < bytes) {
            ftamd_hip_free(P->d_stage);
            P->d_stage = NULL;
            P->stage_bytes = 0;
            rc = ftamd_hip_malloc(&P->d_stage, bytes);
            if (rc == FTAMD_OK) P->stage_bytes = bytes;
        }
        d = P->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, X, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = ftamd_grid_run(P->g, d, s.exec_sphv, s.f0, s.f1, s.a, s.b, s.wmode, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(X, d, bytes, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&P->lock);
    ftamd_set_error(rc);
}

/* fasttransforms.h:488-493 (fftw.c:126-270) */
void ft_execute_sph_synthesis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(0, TRANS, P, X, N, M);
}
void ft_execute_sph_analysis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(1, TRANS, P, X, N, M);
}
void ft_execute_sphv_synthesis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(2, TRANS, P, X, N, M);
}
void ft_execute_sphv_analysis(const char TRANS, const ft_sphere_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(3, TRANS, P, X, N, M);
}

/* fasttransforms.h:542-556 (fftw.c:400-520): the disk's plans, REDFT01 or
 * REDFT10 on columns 0 and 3 mod 4, REDFT11 on columns 1 and 2 mod 4, the
 * phi DFT as the sphere's. */
ft_disk_fftw_plan *ft_plan_disk_synthesis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 1, 2, 0.0);
}
ft_disk_fftw_plan *ft_plan_disk_analysis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 0, 2, 0.0);
}
void ft_destroy_disk_fftw_plan(ft_disk_fftw_plan *P) { ft_destroy_sphere_fftw_plan(P); }
void ft_execute_disk_synthesis(const char TRANS, const ft_disk_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(4, TRANS, P, X, N, M);
}
void ft_execute_disk_analysis(const char TRANS, const ft_disk_fftw_plan *P, double *X, const int N, const int M) {
    execute_grid(5, TRANS, P, X, N, M);
}

/* fasttransforms.h:558-575 (fftw.c:530-655): the annulus's plans, REDFT01 or
 * REDFT10 on every column, the radial weights on columns 1 and 2 mod 4, the
 * phi DFT as the sphere's; rho in [0, 1). */
ft_annulus_fftw_plan *ft_plan_annulus_synthesis(const int N, const int M, const double rho, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 1, 3, rho);
}
ft_annulus_fftw_plan *ft_plan_annulus_analysis(const int N, const int M, const double rho, const unsigned flags) {
    (void)flags;
    return plan_grid(N, M, 0, 3, rho);
}
double ft_get_rho_annulus_fftw_plan(ft_annulus_fftw_plan *P) { return P ? P->rho : 0.0; }
void ft_destroy_annulus_fftw_plan(ft_annulus_fftw_plan *P) { ft_destroy_sphere_fftw_plan(P); }
void ft_execute_annulus_synthesis(const char TRANS, const ft_annulus_fftw_plan *P, double *X, const int N,
                                  const int M) {
    execute_grid(6, TRANS, P, X, N, M);
}
void ft_execute_annulus_analysis(const char TRANS, const ft_annulus_fftw_plan *P, double *X, const int N,
                                 const int M) {
    execute_grid(7, TRANS, P, X, N, M);
}

/* ------------------------------------------------------------ triangle -- */

/* fasttransforms.h:497-512 (fftw.c:275-330): the plan fixes the kinds of
 * FFTW's r2r_2d over X as an N x M row-major buffer (synthesis REDFT01 on
 * both axes, analysis REDFT10); the executors scale before or after it and
 * do nothing unless N > 1 and M > 1, as the reference's do. */
struct ft_triangle_fftw_plan_s {
    int N, M, synthesis;
    FtamdTriGrid *g;
    void *d_stage;
    size_t stage_bytes;
    pthread_mutex_t lock;
};

static ft_triangle_fftw_plan *plan_tri_grid(int N, int M, int synthesis) {
    if (N < 1 || M < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return NULL;
    }
    ft_triangle_fftw_plan *P = calloc(1, sizeof(ft_triangle_fftw_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    P->N = N;
    P->M = M;
    P->synthesis = synthesis;
    pthread_mutex_init(&P->lock, NULL);
    rc = ftamd_trigrid_create(N, M, synthesis, &P->g);
    if (rc != FTAMD_OK) {
        pthread_mutex_destroy(&P->lock);
        free(P);
        ftamd_set_error(rc);
        return NULL;
    }
    return P;
}

ft_triangle_fftw_plan *ft_plan_tri_synthesis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_tri_grid(N, M, 1);
}
ft_triangle_fftw_plan *ft_plan_tri_analysis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_tri_grid(N, M, 0);
}

void ft_destroy_triangle_fftw_plan(ft_triangle_fftw_plan *P) {
    if (!P) return;
    ftamd_trigrid_destroy(P->g);
    ftamd_hip_free(P->d_stage);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* The scalings of executor `analysis` and TRANS (fftw.c:290-330). */
static int tri_steps(int analysis, char TRANS, int N, int M, int *pre_fix, double *pre_mul, double *pre_div,
                     int *post_fix, double *post_div) {
    if (TRANS != 'N' && TRANS != 'T') return 0;
    *pre_fix = 0, *pre_mul = 1.0, *pre_div = 0.0, *post_fix = 0, *post_div = 0.0;
    const double nm = (double)N * (double)M;
    if (!analysis && TRANS == 'N') *pre_fix = 1, *pre_mul = 0.25;
    else if (!analysis) *pre_mul = 0.25;
    else if (TRANS == 'N') *post_fix = 1, *post_div = nm;
    else *pre_div = nm;
    return 1;
}

int ft_gpu_execute_tri_grid(int analysis, char TRANS, const ft_triangle_fftw_plan *P_, double *dX, int N, int M,
                            void *stream) {
    ft_triangle_fftw_plan *P = (ft_triangle_fftw_plan *)P_;
    int pf, qf;
    double pm, pd, qd;
    if ((analysis != 0 && analysis != 1) || !tri_steps(analysis, TRANS, N, M, &pf, &pm, &pd, &qf, &qd) || !P ||
        P->N != N || P->M != M || !dX)
        return FT_GPU_EINVAL;
    if (N < 2 || M < 2) return FT_GPU_OK;
    pthread_mutex_lock(&P->lock);
    int rc = ftamd_trigrid_run(P->g, dX, pf, pm, pd, qf, qd, stream);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

static void execute_tri_grid(int analysis, char TRANS, const ft_triangle_fftw_plan *P_, double *X, int N, int M) {
    ft_triangle_fftw_plan *P = (ft_triangle_fftw_plan *)P_;
    int pf, qf;
    double pm, pd, qd;
    if (!tri_steps(analysis, TRANS, N, M, &pf, &pm, &pd, &qf, &qd) || !P || P->N != N || P->M != M || !X) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    if (N < 2 || M < 2) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(X, &pk);
    const size_t bytes = (size_t)N * (size_t)M * sizeof(double);
    pthread_mutex_lock(&P->lock);
    void *d = X;
    if (pk != FTAMD_PTR_DEVICE) {
        if (P->stage_bytes < bytes) {
            ftamd_hip_free(P->d_stage);
            P->d_stage = NULL;
            P->stage_bytes = 0;
            rc = ftamd_hip_malloc(&P->d_stage, bytes);
            if (rc == FTAMD_OK) P->stage_bytes = bytes;
        }
        d = P->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, X, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = ftamd_trigrid_run(P->g, d, pf, pm, pd, qf, qd, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(X, d, bytes, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&P->lock);
    ftamd_set_error(rc);
}

/* fasttransforms.h:509-512 (fftw.c:290-330) */
void ft_execute_tri_synthesis(const char TRANS, const ft_triangle_fftw_plan *P, double *X, const int N, const int M) {
    execute_tri_grid(0, TRANS, P, X, N, M);
}
void ft_execute_tri_analysis(const char TRANS, const ft_triangle_fftw_plan *P, double *X, const int N, const int M) {
    execute_tri_grid(1, TRANS, P, X, N, M);
}

/* --------------------------------------------------------- tetrahedron -- */

/* fasttransforms.h:514-524 (fftw.c:333-395): FFTW's r2r_3d(N, L, M) over X as
 * a row-major N x L x M buffer (synthesis REDFT01 on every axis, analysis
 * REDFT10) with the executors' face scalings on the column-major view
 * (x2 / x0.5 where i, j or k is 0); nothing unless N, L and M exceed 1. */
struct ft_tetrahedron_fftw_plan_s {
    int N, L, M, synthesis;
    FtamdTriGrid *g;
    void *d_stage;
    size_t stage_bytes;
    pthread_mutex_t lock;
};

static ft_tetrahedron_fftw_plan *plan_tet_grid(int N, int L, int M, int synthesis) {
    if (N < 1 || L < 1 || M < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return NULL;
    }
    ft_tetrahedron_fftw_plan *P = calloc(1, sizeof(ft_tetrahedron_fftw_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    P->N = N;
    P->L = L;
    P->M = M;
    P->synthesis = synthesis;
    pthread_mutex_init(&P->lock, NULL);
    const int dims[3] = {N, L, M};
    rc = ftamd_boxgrid_create(3, dims, synthesis, &P->g);
    if (rc != FTAMD_OK) {
        pthread_mutex_destroy(&P->lock);
        free(P);
        ftamd_set_error(rc);
        return NULL;
    }
    return P;
}

ft_tetrahedron_fftw_plan *ft_plan_tet_synthesis(const int N, const int L, const int M, const unsigned flags) {
    (void)flags;
    return plan_tet_grid(N, L, M, 1);
}
ft_tetrahedron_fftw_plan *ft_plan_tet_analysis(const int N, const int L, const int M, const unsigned flags) {
    (void)flags;
    return plan_tet_grid(N, L, M, 0);
}

void ft_destroy_tetrahedron_fftw_plan(ft_tetrahedron_fftw_plan *P) {
    if (!P) return;
    ftamd_trigrid_destroy(P->g);
    ftamd_hip_free(P->d_stage);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* The scalings of the tetrahedron's executors (fftw.c:349-395). */
static int tet_steps(int analysis, char TRANS, int N, int L, int M, int *pf, double *pm, double *pd, int *qf, double *qm,
                     double *qd) {
    if (TRANS != 'N' && TRANS != 'T') return 0;
    *pf = 0, *pm = 1.0, *pd = 0.0, *qf = 0, *qm = 1.0, *qd = 0.0;
    const double nlm = (double)N * (double)L * (double)M;
    if (!analysis && TRANS == 'N') *pf = 1, *pm = 0.125;
    else if (!analysis) *qm = 0.125;
    else if (TRANS == 'N') *qf = 1, *qd = nlm;
    else *pd = nlm;
    return 1;
}

static int run_tet_grid(ft_tetrahedron_fftw_plan *P, int analysis, char TRANS, double *d, void *stream) {
    int pf, qf;
    double pm, pd, qm, qd;
    tet_steps(analysis, TRANS, P->N, P->L, P->M, &pf, &pm, &pd, &qf, &qm, &qd);
    return ftamd_boxgrid_run(P->g, d, pf, pm, pd, qf, qm, qd, stream);
}

int ft_gpu_execute_tet_grid(int analysis, char TRANS, const ft_tetrahedron_fftw_plan *P_, double *dX, int N, int L,
                            int M, void *stream) {
    ft_tetrahedron_fftw_plan *P = (ft_tetrahedron_fftw_plan *)P_;
    if ((analysis != 0 && analysis != 1) || (TRANS != 'N' && TRANS != 'T') || !P || P->N != N || P->L != L ||
        P->M != M || !dX)
        return FT_GPU_EINVAL;
    if (N < 2 || L < 2 || M < 2) return FT_GPU_OK;
    pthread_mutex_lock(&P->lock);
    int rc = run_tet_grid(P, analysis, TRANS, dX, stream);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

static void execute_tet_grid(int analysis, char TRANS, const ft_tetrahedron_fftw_plan *P_, double *X, int N, int L,
                             int M) {
    ft_tetrahedron_fftw_plan *P = (ft_tetrahedron_fftw_plan *)P_;
    if ((TRANS != 'N' && TRANS != 'T') || !P || P->N != N || P->L != L || P->M != M || !X) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    if (N < 2 || L < 2 || M < 2) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(X, &pk);
    const size_t bytes = (size_t)N * L * M * sizeof(double);
    pthread_mutex_lock(&P->lock);
    void *d = X;
    if (pk != FTAMD_PTR_DEVICE) {
        if (P->stage_bytes < bytes) {
            ftamd_hip_free(P->d_stage);
            P->d_stage = NULL;
            P->stage_bytes = 0;
            rc = ftamd_hip_malloc(&P->d_stage, bytes);
            if (rc == FTAMD_OK) P->stage_bytes = bytes;
        }
        d = P->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, X, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = run_tet_grid(P, analysis, TRANS, d, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(X, d, bytes, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&P->lock);
    ftamd_set_error(rc);
}

/* fasttransforms.h:526-527 (fftw.c:349-395) */
void ft_execute_tet_synthesis(const char TRANS, const ft_tetrahedron_fftw_plan *P, double *X, const int N, const int L,
                              const int M) {
    execute_tet_grid(0, TRANS, P, X, N, L, M);
}
void ft_execute_tet_analysis(const char TRANS, const ft_tetrahedron_fftw_plan *P, double *X, const int N, const int L,
                             const int M) {
    execute_tet_grid(1, TRANS, P, X, N, L, M);
}

/* ------------------------------------------------------ rectangle-disk -- */

/* fasttransforms.h:577-590 (fftw.c:657-740): the plan fixes the kinds (x:
 * REDFT01 on the even columns, RODFT01 on the odd ones, y: REDFT01 along the
 * rows for synthesis plans; REDFT10 / RODFT10 / REDFT10 for analysis plans);
 * the executors fix the order of the two passes and the scalings. */
struct ft_rectdisk_fftw_plan_s {
    int N, M, synthesis;
    FtamdRectGrid *g;
    void *d_stage;
    size_t stage_bytes;
    pthread_mutex_t lock;
};

static ft_rectdisk_fftw_plan *plan_rect_grid(int N, int M, int synthesis) {
    if (N < 1 || M < 1) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return NULL;
    }
    ft_rectdisk_fftw_plan *P = calloc(1, sizeof(ft_rectdisk_fftw_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    P->N = N;
    P->M = M;
    P->synthesis = synthesis;
    pthread_mutex_init(&P->lock, NULL);
    rc = ftamd_rectgrid_create(N, M, synthesis, &P->g);
    if (rc != FTAMD_OK) {
        pthread_mutex_destroy(&P->lock);
        free(P);
        ftamd_set_error(rc);
        return NULL;
    }
    return P;
}

ft_rectdisk_fftw_plan *ft_plan_rectdisk_synthesis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_rect_grid(N, M, 1);
}
ft_rectdisk_fftw_plan *ft_plan_rectdisk_analysis(const int N, const int M, const unsigned flags) {
    (void)flags;
    return plan_rect_grid(N, M, 0);
}

void ft_destroy_rectdisk_fftw_plan(ft_rectdisk_fftw_plan *P) {
    if (!P) return;
    ftamd_rectgrid_destroy(P->g);
    ftamd_hip_free(P->d_stage);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* The two passes of the executors (fftw.c:690-740).  Masks: 1 row 0 of the
 * even columns, 2 the last row of the odd columns, 3 column 0. */
static int run_rect_grid(ft_rectdisk_fftw_plan *P, int analysis, char TRANS, double *d, void *stream) {
    FtamdRectGrid *g = P->g;
    const double nm = (double)P->N * (double)P->M;
    int rc;
    if (!analysis && TRANS == 'N') {
        /* x2 row 0 of the even columns, x, x2 column 0, x0.25, y */
        rc = ftamd_rectgrid_pass(g, d, 0, 1, 2.0, 3, 2.0, 0.25, 0.0, stream);
        if (rc == FTAMD_OK) rc = ftamd_rectgrid_pass(g, d, 1, 0, 1.0, 0, 1.0, 1.0, 0.0, stream);
    } else if (!analysis) {
        /* y, x0.25, x, x0.5 the last row of the odd columns */
        rc = ftamd_rectgrid_pass(g, d, 1, 0, 1.0, 0, 1.0, 0.25, 0.0, stream);
        if (rc == FTAMD_OK) rc = ftamd_rectgrid_pass(g, d, 0, 0, 1.0, 2, 0.5, 1.0, 0.0, stream);
    } else if (TRANS == 'N') {
        /* y, / NM, x0.5 column 0, x, x0.5 row 0 of the even columns */
        rc = ftamd_rectgrid_pass(g, d, 1, 0, 1.0, 3, 0.5, 1.0, nm, stream);
        if (rc == FTAMD_OK) rc = ftamd_rectgrid_pass(g, d, 0, 0, 1.0, 1, 0.5, 1.0, 0.0, stream);
    } else {
        /* x2 the last row of the odd columns, x, y, / NM */
        rc = ftamd_rectgrid_pass(g, d, 0, 2, 2.0, 0, 1.0, 1.0, 0.0, stream);
        if (rc == FTAMD_OK) rc = ftamd_rectgrid_pass(g, d, 1, 0, 1.0, 0, 1.0, 1.0, nm, stream);
    }
    return rc;
}

int ft_gpu_execute_rectdisk_grid(int analysis, char TRANS, const ft_rectdisk_fftw_plan *P_, double *dX, int N, int M,
                                 void *stream) {
    ft_rectdisk_fftw_plan *P = (ft_rectdisk_fftw_plan *)P_;
    if ((analysis != 0 && analysis != 1) || (TRANS != 'N' && TRANS != 'T') || !P || P->N != N || P->M != M || !dX)
        return FT_GPU_EINVAL;
    pthread_mutex_lock(&P->lock);
    int rc = run_rect_grid(P, analysis, TRANS, dX, stream);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

static void execute_rect_grid(int analysis, char TRANS, const ft_rectdisk_fftw_plan *P_, double *X, int N, int M) {
    ft_rectdisk_fftw_plan *P = (ft_rectdisk_fftw_plan *)P_;
    if ((TRANS != 'N' && TRANS != 'T') || !P || P->N != N || P->M != M || !X) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(X, &pk);
    const size_t bytes = (size_t)N * M * sizeof(double);
    pthread_mutex_lock(&P->lock);
    void *d = X;
    if (pk != FTAMD_PTR_DEVICE) {
        if (P->stage_bytes < bytes) {
            ftamd_hip_free(P->d_stage);
            P->d_stage = NULL;
            P->stage_bytes = 0;
            rc = ftamd_hip_malloc(&P->d_stage, bytes);
            if (rc == FTAMD_OK) P->stage_bytes = bytes;
        }
        d = P->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, X, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = run_rect_grid(P, analysis, TRANS, d, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(X, d, bytes, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&P->lock);
    ftamd_set_error(rc);
}

/* fasttransforms.h:589-590 (fftw.c:690-740) */
void ft_execute_rectdisk_synthesis(const char TRANS, const ft_rectdisk_fftw_plan *P, double *X, const int N,
                                   const int M) {
    execute_rect_grid(0, TRANS, P, X, N, M);
}
void ft_execute_rectdisk_analysis(const char TRANS, const ft_rectdisk_fftw_plan *P, double *X, const int N,
                                  const int M) {
    execute_rect_grid(1, TRANS, P, X, N, M);
}

/* -------------------------------------------------- spin-weighted sphere -- */

/* fasttransforms.h:592-617 (fftw.c:742-937): complex N x M arrays; the plan
 * fixes the theta kinds (the sphere's column classes for even spin S, the
 * vector field's for odd) and the row DFT's sign (synthesis backward,
 * analysis forward); synthesis 'N' and analysis 'C' take a synthesis plan,
 * synthesis 'C' and analysis 'N' an analysis plan (test_fftw.c pairs them
 * so); another pairing sets FT_GPU_EINVAL. */
struct ft_spinsphere_fftw_plan_s {
    int N, M, S, synthesis;
    FtamdSpinGrid *g;
    void *d_stage;
    size_t stage_bytes;
    pthread_mutex_t lock;
};

#define FT_1_2_SQRT_2PI 0.199471140200716338
#define FT_2_SQRT_2PI 5.013256549262001005

static ft_spinsphere_fftw_plan *plan_spin_grid(int N, int M, int S, int synthesis) {
    if (N < 1 || M < 1 || M % 2 == 0) {
        ftamd_set_error(FT_GPU_EINVAL);
        return NULL;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return NULL;
    }
    ft_spinsphere_fftw_plan *P = calloc(1, sizeof(ft_spinsphere_fftw_plan));
    if (!P) {
        ftamd_set_error(FT_GPU_ENOMEM);
        return NULL;
    }
    P->N = N;
    P->M = M;
    P->S = S;
    P->synthesis = synthesis;
    pthread_mutex_init(&P->lock, NULL);
    rc = ftamd_spingrid_create(N, M, synthesis, S % 2 != 0, &P->g);
    if (rc != FTAMD_OK) {
        pthread_mutex_destroy(&P->lock);
        free(P);
        ftamd_set_error(rc);
        return NULL;
    }
    return P;
}

ft_spinsphere_fftw_plan *ft_plan_spinsph_synthesis(const int N, const int M, const int S, const unsigned flags) {
    (void)flags;
    return plan_spin_grid(N, M, S, 1);
}
ft_spinsphere_fftw_plan *ft_plan_spinsph_analysis(const int N, const int M, const int S, const unsigned flags) {
    (void)flags;
    return plan_spin_grid(N, M, S, 0);
}
int ft_get_spin_spinsphere_fftw_plan(const ft_spinsphere_fftw_plan *P) { return P ? P->S : 0; }

void ft_destroy_spinsphere_fftw_plan(ft_spinsphere_fftw_plan *P) {
    if (!P) return;
    ftamd_spingrid_destroy(P->g);
    ftamd_hip_free(P->d_stage);
    pthread_mutex_destroy(&P->lock);
    free(P);
}

/* The fix-ups and scaling of the executors (fftw.c:800-937). */
static int spin_steps(const ft_spinsphere_fftw_plan *P, int analysis, char TRANS, int N, int M, double *f0, double *f1,
                      double *a) {
    if (!P || P->N != N || P->M != M || (TRANS != 'N' && TRANS != 'C')) return 0;
    const int need_syn = analysis == (TRANS == 'C');
    if (P->synthesis != need_syn) return 0;
    if (!analysis && TRANS == 'N') *f0 = 2.0, *f1 = 1.0, *a = FT_1_2_SQRT_2PI;
    else if (!analysis) *f0 = 1.0, *f1 = 0.5, *a = FT_1_2_SQRT_2PI;
    else if (TRANS == 'N') *f0 = 0.5, *f1 = 1.0, *a = FT_2_SQRT_2PI / (2.0 * N * M);
    else *f0 = 1.0, *f1 = 2.0, *a = FT_2_SQRT_2PI / (2.0 * N * M);
    return 1;
}

int ft_gpu_execute_spinsph_grid(int analysis, char TRANS, const ft_spinsphere_fftw_plan *P_, void *dX, int N, int M,
                                void *stream) {
    ft_spinsphere_fftw_plan *P = (ft_spinsphere_fftw_plan *)P_;
    double f0, f1, a;
    if ((analysis != 0 && analysis != 1) || !spin_steps(P, analysis, TRANS, N, M, &f0, &f1, &a) || !dX)
        return FT_GPU_EINVAL;
    pthread_mutex_lock(&P->lock);
    int rc = ftamd_spingrid_run(P->g, dX, f0, f1, a, stream);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

static void execute_spin_grid(int analysis, char TRANS, const ft_spinsphere_fftw_plan *P_, ft_complex *X, int N,
                              int M) {
    ft_spinsphere_fftw_plan *P = (ft_spinsphere_fftw_plan *)P_;
    double f0, f1, a;
    if (!spin_steps(P, analysis, TRANS, N, M, &f0, &f1, &a) || !X) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(X, &pk);
    const size_t bytes = 2 * (size_t)N * M * sizeof(double);
    pthread_mutex_lock(&P->lock);
    void *d = X;
    if (pk != FTAMD_PTR_DEVICE) {
        if (P->stage_bytes < bytes) {
            ftamd_hip_free(P->d_stage);
            P->d_stage = NULL;
            P->stage_bytes = 0;
            rc = ftamd_hip_malloc(&P->d_stage, bytes);
            if (rc == FTAMD_OK) P->stage_bytes = bytes;
        }
        d = P->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, X, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = ftamd_spingrid_run(P->g, d, f0, f1, a, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(X, d, bytes, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&P->lock);
    ftamd_set_error(rc);
}

/* fasttransforms.h:614-617 (fftw.c:800-937) */
void ft_execute_spinsph_synthesis(const char TRANS, const ft_spinsphere_fftw_plan *P, ft_complex *X, const int N,
                                  const int M) {
    execute_spin_grid(0, TRANS, P, X, N, M);
}
void ft_execute_spinsph_analysis(const char TRANS, const ft_spinsphere_fftw_plan *P, ft_complex *X, const int N,
                                 const int M) {
    execute_spin_grid(1, TRANS, P, X, N, M);
}
