/*
 * ft_tuning.c -- the autotuning table: for a kind, precision, direction,
 * single transform or batch, and degree band, the engine and its shape that
 * measured fastest on the device (the table fasttransforms_amd/tune.py
 * writes after a sweep on the GPU, compiled in from ft_tuned_gfx950.h), or a
 * table the caller installs at run time (ft_gpu_set_tuning_table).  No
 * reference counterpart: the reference picks its SIMD kernel by CPU feature
 * alone (src/drivers.c:58-110).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

typedef struct {
    int kind, precision, dir, batch, nlo, nhi, engine, maxpart, L, W, wg_per_cu;
} ft_tune_row;

#include "ft_tuned_gfx950.h"

static pthread_mutex_t g_tune_lock = PTHREAD_MUTEX_INITIALIZER;
static ft_tune_row *g_user = NULL;  /* installed table, or NULL */
static int g_user_n = 0;
static int g_mode = 0;              /* 0 compiled table, 1 installed table, -1 none */

static const ft_tune_row *compiled_rows(int *n) {
    int k = 0;
    while (k < (int)(sizeof(ft_tuned_gfx950) / sizeof(ft_tuned_gfx950[0])) && ft_tuned_gfx950[k].kind >= 0) k++;
    *n = k;
    return ft_tuned_gfx950;
}

static int row_valid(const ft_tune_row *r) {
    return r->kind >= 0 && r->kind <= 2 && (r->precision == 0 || r->precision == 1) && (r->dir == 0 || r->dir == 1) &&
           (r->batch == 0 || r->batch == 1) && r->nlo >= 1 && r->nhi >= r->nlo &&
           (r->engine == FT_GPU_ENGINE_PIPELINE || r->engine == FT_GPU_ENGINE_SYSTOLIC) && r->maxpart >= 0 &&
           r->maxpart <= FTAMD_SYS_WAVES && (r->L == 0 || r->L == 2 || r->L == 4 || r->L == 8) &&
           (r->W == 0 || r->W == 4 || r->W == 8 || r->W == 16) && r->wg_per_cu >= 0 && r->wg_per_cu <= 8;
}

int ft_gpu_set_tuning_table(const int *rows, int count) {
    if (count < -1 || (count > 0 && !rows)) return FT_GPU_EINVAL;
    ft_tune_row *copy = NULL;
    if (count > 0) {
        copy = malloc(sizeof(ft_tune_row) * (size_t)count);
        if (!copy) return FT_GPU_ENOMEM;
        for (int i = 0; i < count; i++) {
            memcpy(&copy[i], rows + (size_t)i * FT_GPU_TUNE_COLS, sizeof(ft_tune_row));
            if (!row_valid(&copy[i])) {
                free(copy);
                return FT_GPU_EINVAL;
            }
        }
    }
    pthread_mutex_lock(&g_tune_lock);
    free(g_user);
    g_user = copy;
    g_user_n = count > 0 ? count : 0;
    g_mode = count > 0 ? 1 : count;
    pthread_mutex_unlock(&g_tune_lock);
    return FT_GPU_OK;
}

int ft_gpu_tuning_table(int *rows, int cap) {
    pthread_mutex_lock(&g_tune_lock);
    int n = 0;
    const ft_tune_row *t = NULL;
    if (g_mode == 0) t = compiled_rows(&n);
    else if (g_mode == 1) t = g_user, n = g_user_n;
    for (int i = 0; rows && i < n && i < cap; i++) memcpy(rows + (size_t)i * FT_GPU_TUNE_COLS, &t[i], sizeof(ft_tune_row));
    pthread_mutex_unlock(&g_tune_lock);
    return n;
}

/* The active table's row for (kind, precision, dir, batch, n): 1 and *out,
 * or 0.  The compiled table applies only to the architecture it was measured
 * on (or where the device cannot be queried: table inspection without a GPU). */
int ftamd_tuned_lookup(int kind, int precision, int dir, int batch, int n, int *engine, int *maxpart, int *L, int *W,
                       int *wg_per_cu) {
    int found = 0;
    pthread_mutex_lock(&g_tune_lock);
    int cnt = 0;
    const ft_tune_row *t = NULL;
    if (g_mode == 0) {
        char arch[64];
        if (ftamd_hip_arch(arch, sizeof(arch)) != FTAMD_OK || strncmp(arch, FT_TUNED_ARCH, strlen(FT_TUNED_ARCH)) == 0)
            t = compiled_rows(&cnt);
    } else if (g_mode == 1) {
        t = g_user;
        cnt = g_user_n;
    }
    for (int i = 0; i < cnt; i++) {
        const ft_tune_row *r = &t[i];
        if (r->kind == kind && r->precision == precision && r->dir == dir && r->batch == batch && n >= r->nlo &&
            n <= r->nhi) {
            *engine = r->engine;
            *maxpart = r->maxpart;
            *L = r->L;
            *W = r->W;
            *wg_per_cu = r->wg_per_cu;
            found = 1;
            break;
        }
    }
    pthread_mutex_unlock(&g_tune_lock);
    return found;
}
