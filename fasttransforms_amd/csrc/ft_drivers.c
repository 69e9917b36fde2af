/*
 * ft_drivers.c -- the reference's executors and single-vector kernels on the
 * GPU, plus the schedules the pipelined kernel runs.
 *
 * src/drivers.c:58-92, 278-312, 378-412 define ft_execute_{sph,tri,disk}_
 * {hi2lo,lo2hi}: every column j >= 1 of order m (m = (j+1)/2 for sph/disk,
 * m = j for tri) is taken from order m down to m%2 (sph/disk) or 0 (tri), or
 * back up.  Here those columns are grouped 64 at a time (one wavefront, one
 * column per lane) by what sequences they need, and each group's chain of
 * sequences is cut into pipeline stages and workgroup-sized tasks.  The
 * reference's warp/permute passes (drivers.c:114-115, 125-126) have no
 * counterpart: the grouping is only a lane-to-column map.
 */
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

/* ------------------------------------------------------------ errors -- */

static __thread int t_last_error = 0;

void ftamd_set_error(int code) {
    if (code != FTAMD_OK && t_last_error == FTAMD_OK) t_last_error = code;
}

int ft_gpu_last_error(void) {
    int e = t_last_error;
    t_last_error = FTAMD_OK;
    return e;
}

const char *ft_gpu_error_string(int code) {
    switch (code) {
        case FT_GPU_OK: return "no error";
        case FT_GPU_EINVAL: return "invalid argument";
        case FT_GPU_ENOMEM: return "device allocation failed";
        case FT_GPU_EDEVICE: return "HIP device or runtime error";
        case FT_GPU_ETIMEOUT: return "kernel wait expired; results are invalid";
        case FT_GPU_EPTR: return "pointer is not host memory or device memory of the current device";
        case FT_GPU_EINCOMPLETE: return "a launch ended with work not done; results are invalid";
        default: return "unknown error";
    }
}

/* ------------------------------------------------------------ tuning -- */

/* Process-wide knobs, read and written atomically (a call reads each once). */
static int g_L = 0, g_W = 0, g_wg_per_cu = 0;
#define KNOB_GET(v) __atomic_load_n(&(v), __ATOMIC_RELAXED)
#define KNOB_SET(v, x) __atomic_store_n(&(v), (x), __ATOMIC_RELAXED)

int ft_gpu_set_tuning(int L, int W, int wg_per_cu) {
    if (L != 0 && L != 2 && L != 4 && L != 8) return FT_GPU_EINVAL;
    if (W != 0 && W != 4 && W != 8 && W != 16) return FT_GPU_EINVAL;
    if (wg_per_cu < 0 || wg_per_cu > 8) return FT_GPU_EINVAL;
    KNOB_SET(g_L, L);
    KNOB_SET(g_W, W);
    KNOB_SET(g_wg_per_cu, wg_per_cu);
    return FT_GPU_OK;
}

static int g_engine = FT_GPU_ENGINE_AUTO;
static int g_spin_limit = 0;

int ft_gpu_set_spin_limit(int limit) {
    if (limit < 0) return FT_GPU_EINVAL;
    KNOB_SET(g_spin_limit, limit);
    return FT_GPU_OK;
}

int ft_gpu_set_engine(int engine) {
    if (engine != FT_GPU_ENGINE_AUTO && engine != FT_GPU_ENGINE_PIPELINE && engine != FT_GPU_ENGINE_SYSTOLIC)
        return FT_GPU_EINVAL;
    KNOB_SET(g_engine, engine);
    return FT_GPU_OK;
}

int ftamd_engine(void) { return KNOB_GET(g_engine); }

/* Bands per chain part of the band-systolic engine (longer chains run in
 * parts on several workgroups of one XCD). */
static int g_sys_maxpart = 0;

int ft_gpu_set_systolic(int maxpart) {
    if (maxpart < 0 || maxpart > FTAMD_SYS_WAVES) return FT_GPU_EINVAL;
    KNOB_SET(g_sys_maxpart, maxpart);
    return FT_GPU_OK;
}

static __thread int t_last_engine = 0;
int ft_gpu_last_engine(void) { return t_last_engine; }

void ftamd_choose(int kind, int precision, int dir, int batch, int n, ftamd_choice *c) {
    int te = 0, tm = 0, tL = 0, tW = 0, tg = 0;
    const int found = ftamd_tuned_lookup(kind, precision, dir, batch, n, &te, &tm, &tL, &tW, &tg);
    const int ke = KNOB_GET(g_engine), km = KNOB_GET(g_sys_maxpart), kL = KNOB_GET(g_L), kW = KNOB_GET(g_W),
              kg = KNOB_GET(g_wg_per_cu);
    /* built-in: the band-systolic engine for single transforms, the pipelined
     * one for batches (many short chains leave the systolic one's CUs idle),
     * whole chains of up to 16 bands, L 4, W 16, batches at 2 workgroups per
     * CU (512 sph transforms of degree 32..1018: 9.9 ms capped, 10.4 ms not) */
    c->forced = ke != FT_GPU_ENGINE_AUTO;
    c->engine = c->forced ? ke : found ? te : (batch ? FT_GPU_ENGINE_PIPELINE : FT_GPU_ENGINE_SYSTOLIC);
    c->maxpart = km ? km : (found && tm) ? tm : FTAMD_SYS_WAVES;
    c->L = kL ? kL : (found && tL) ? tL : 4;
    c->W = kW ? kW : (found && tW) ? tW : 16;
    c->wg_per_cu = kg ? kg : (found && tg) ? tg : (batch ? 2 : 0);
}

static int kind_geometry(int kind, int *layout);

/* Choices are made per table geometry: sphv and rectdisk launches use the
 * sphere's. */
int ft_gpu_tuned_choice(int kind, int precision, int direction, int batch, int n, int *choice) {
    int layout, geo = kind_geometry(kind, &layout);
    if (geo < 0 || (precision != 0 && precision != 1) || (direction != 0 && direction != 1) ||
        (batch != 0 && batch != 1) || n < 1 || !choice)
        return FT_GPU_EINVAL;
    ftamd_choice c;
    ftamd_choose(geo, precision, direction, batch, n, &c);
    choice[0] = c.engine;
    choice[1] = c.maxpart;
    choice[2] = c.L;
    choice[3] = c.W;
    choice[4] = c.wg_per_cu;
    return FT_GPU_OK;
}

/* Workgroups are dealt to the XCDs in turn and each takes tasks only from its
 * own XCD's list, so the grid holds as many workgroups per XCD as the longest
 * list has tasks: every task gets a workgroup as soon as the hardware has room
 * (workgroups that find their list empty exit at once).  A schedule's
 * wg_per_cu (the knob at launch time if set) caps the grid at that many
 * workgroups per CU. */
int ftamd_tuning_grid(const ft_schedule *S) {
    int grid = S->launch.nxcd * S->xcd_tasks;
    int knob = KNOB_GET(g_wg_per_cu);
    int per_cu = knob ? knob : S->wg_per_cu;
    if (per_cu) {
        int cu = 0;
        if (ftamd_hip_device(NULL, &cu) != FTAMD_OK || cu <= 0) cu = 256;
        if (grid > cu * per_cu) grid = cu * per_cu;
    }
    return grid;
}

/* ---------------------------------------------------------- geometry -- */

static int geo_step(int kind) { return kind == FTAMD_TRI ? 1 : 2; }
static int geo_top(int kind, int n, int j) {
    if (kind == FTAMD_SPH) return n - 3 - j;
    if (kind == FTAMD_TRI) return n - 2 - j;
    return n - 2 - (j + 1) / 2;
}

/* --------------------------------------------------------- schedules -- */

/* One column to be transformed. */
typedef struct {
    void *A;       /* device matrix */
    int n, rs;
    long long off; /* element offset of the column */
    int jmax;      /* largest sequence j it needs */
    int jmin;      /* smallest (hi2lo end / lo2hi start) */
} ft_col;

typedef struct {
    FtamdGroup g;
    int off[64];
    int jmax[64];
    int crit;      /* ordering key: longer chains first */
    double work;   /* rotations of all its columns */
    int xcd;       /* XCD whose workgroups run every task of the group */
} ft_hgroup;

static double column_rotations(int kind, int n, int jmin, int jmax);

/* Group up to 64 consecutive columns of `cols` (already sorted so that
 * columns sharing sequences are adjacent) with equal jmin and A. */
static int make_groups(const ft_col *cols, int ncols, int kind, int dir, int L, int W, ft_hgroup **out,
                       int *ngroups, int *nprog) {
    /* The sphere couples rows k and k+2, so its even and odd rows are two
     * independent chains: every sphere group is run as two groups (rpar). */
    const int split = kind == FTAMD_SPH;
    ft_hgroup *gs = calloc((size_t)(2 * ncols + 2), sizeof(ft_hgroup));
    if (!gs) return FTAMD_ENOMEM;
    int step = geo_step(kind), G = 0, prog = 0;
    int i = 0;
    while (i < ncols) {
        int cnt = 0;
        ft_hgroup *h = &gs[G];
        int jmaxmax = -1, jmaxmin = 1 << 30;
        h->work = 0;
        while (i < ncols && cnt < 64 && (cnt == 0 || (cols[i].jmin == cols[i - 1].jmin && cols[i].A == cols[i - 1].A))) {
            h->work += column_rotations(kind, cols[i].n, cols[i].jmin, cols[i].jmax);
            h->off[cnt] = (int)cols[i].off;
            h->jmax[cnt] = cols[i].jmax;
            if (cols[i].jmax > jmaxmax) jmaxmax = cols[i].jmax;
            if (cols[i].jmax < jmaxmin) jmaxmin = cols[i].jmax;
            cnt++;
            i++;
        }
        const ft_col *c0 = &cols[i - cnt];
        for (int l = cnt; l < 64; l++) {
            h->off[l] = -1;
            h->jmax[l] = -1;
        }
        int jmin = c0->jmin;
        if (jmaxmax < jmin) continue; /* nothing to do for these columns */
        int Q = (jmaxmax - jmin) / step + 1;
        h->g.A = c0->A;
        h->g.n = c0->n;
        h->g.rs = c0->rs;
        h->g.j_first = dir == 0 ? jmin + (Q - 1) * step : jmin;
        h->g.Q = Q;
        h->g.nstages = (Q + L - 1) / L;
        h->g.npass = (h->g.nstages + (W - 2) - 1) / (W - 2);  /* W-2 stage waves per task */
        h->g.jfull = jmaxmin;
        h->g.prog_base = prog;
        h->crit = (split ? c0->n / 2 : c0->n) + Q;
        h->g.rpar = split ? 0 : -1;
        prog += h->g.npass;
        G++;
        if (split) {
            gs[G] = *h;
            gs[G].g.rpar = 1;
            gs[G].g.prog_base = prog;
            h->work *= 0.5;
            gs[G].work = h->work;
            prog += gs[G].g.npass;
            G++;
        }
    }
    *out = gs;
    *ngroups = G;
    *nprog = prog;
    return FTAMD_OK;
}

/* Rotations one column performs (for the FLOP count). */
static double column_rotations(int kind, int n, int jmin, int jmax) {
    double r = 0;
    for (int j = jmin; j <= jmax; j += geo_step(kind)) {
        int t = geo_top(kind, n, j);
        if (t >= 0) r += t + 1;
    }
    return r;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/*
 * Upload groups, lane maps, the task order and a zeroed workspace.
 *
 * Every group is handed to one XCD (largest work first, onto the least loaded
 * XCD), because its tasks pass rows to each other through A and only the
 * workgroups of one XCD share an L2.  Each XCD's list runs pass by pass
 * (every group's pass p before any group's pass p+1), longest chains first
 * within a pass, so a task's producer is always handed out before it.
 */
static int finish_schedule(ft_schedule *S, ft_hgroup *gs, int G, int nprog, const void *table, int ntab) {
    int nxcd = 1;
    if (ftamd_hip_xcd_count(&nxcd) != FTAMD_OK || nxcd < 1) nxcd = 1;
    int maxpass = 0;
    for (int g = 0; g < G; g++)
        if (gs[g].g.npass > maxpass) maxpass = gs[g].g.npass;
    int *order = malloc(sizeof(int) * (size_t)(G ? G : 1));
    if (!order) return FTAMD_ENOMEM;
    /* LPT assignment of groups to XCDs */
    for (int g = 0; g < G; g++) order[g] = g;
    for (int a = 1; a < G; a++) {
        int v = order[a], b = a - 1;
        while (b >= 0 && gs[order[b]].work < gs[v].work) {
            order[b + 1] = order[b];
            b--;
        }
        order[b + 1] = v;
    }
    double load[FTAMD_MAX_XCD] = {0};
    for (int a = 0; a < G; a++) {
        int best = 0;
        for (int x = 1; x < nxcd; x++)
            if (load[x] < load[best]) best = x;
        gs[order[a]].xcd = best;
        load[best] += gs[order[a]].work;
    }
    /* within an XCD: pass-major, longest chains first */
    for (int g = 0; g < G; g++) order[g] = g;
    for (int a = 1; a < G; a++) {
        int v = order[a], b = a - 1;
        while (b >= 0 && gs[order[b]].crit < gs[v].crit) {
            order[b + 1] = order[b];
            b--;
        }
        order[b + 1] = v;
    }
    /* issue priority by chain length, and the boundary buffers: npass-1 of
     * [rows of the sub-problem][64 lanes] per group */
    int maxcrit = 1;
    for (int g = 0; g < G; g++)
        if (gs[g].crit > maxcrit) maxcrit = gs[g].crit;
    long long bnd_elems = 0;
    for (int g = 0; g < G; g++) {
        long long c4 = 4LL * gs[g].crit;
        gs[g].g.prio = c4 > 3LL * maxcrit ? 3 : c4 > 2LL * maxcrit ? 2 : c4 > maxcrit ? 1 : 0;
        int ne = gs[g].g.rpar >= 0 ? (gs[g].g.n - gs[g].g.rpar + 1) / 2 : gs[g].g.n;
        gs[g].g.bnd = bnd_elems;
        bnd_elems += (long long)(gs[g].g.npass - 1) * (ne + 2 * FTAMD_BND_PAD) * 64;
    }
    int T = 0;
    for (int g = 0; g < G; g++) T += gs[g].g.npass;
    int *tasks = malloc(sizeof(int) * 2 * (size_t)(T ? T : 1));
    int xoff[FTAMD_MAX_XCD + 1];
    if (!tasks) {
        free(order);
        return FTAMD_ENOMEM;
    }
    int t = 0;
    for (int x = 0; x < nxcd; x++) {
        xoff[x] = t;
        for (int p = 0; p < maxpass; p++)
            for (int a = 0; a < G; a++) {
                int g = order[a];
                if (gs[g].xcd == x && p < gs[g].g.npass) {
                    tasks[2 * t] = g;
                    tasks[2 * t + 1] = p;
                    t++;
                }
            }
    }
    xoff[nxcd] = t;
    free(order);
    int xcd_tasks = 0;
    for (int x = 0; x < nxcd; x++)
        if (xoff[x + 1] - xoff[x] > xcd_tasks) xcd_tasks = xoff[x + 1] - xoff[x];
    S->xcd_tasks = xcd_tasks;

    size_t o_groups = 0;
    size_t o_off = align256(o_groups + sizeof(FtamdGroup) * (size_t)G);
    size_t o_jmax = align256(o_off + sizeof(int) * 64 * (size_t)G);
    size_t o_tasks = align256(o_jmax + sizeof(int) * 64 * (size_t)G);
    size_t o_xoff = align256(o_tasks + sizeof(int) * 2 * (size_t)T);
    size_t o_prog = align256(o_xoff + sizeof(int) * (FTAMD_MAX_XCD + 1));
    size_t o_ctl = align256(o_prog + sizeof(int) * (size_t)nprog);
    size_t total = align256(o_ctl + (FTAMD_MAX_XCD + 4) * sizeof(int));
    char *host = calloc(1, total);
    if (!host) {
        free(tasks);
        return FTAMD_ENOMEM;
    }
    for (int g = 0; g < G; g++) {
        memcpy(host + o_groups + sizeof(FtamdGroup) * g, &gs[g].g, sizeof(FtamdGroup));
        memcpy(host + o_off + sizeof(int) * 64 * g, gs[g].off, sizeof(int) * 64);
        memcpy(host + o_jmax + sizeof(int) * 64 * g, gs[g].jmax, sizeof(int) * 64);
    }
    memcpy(host + o_tasks, tasks, sizeof(int) * 2 * (size_t)T);
    memcpy(host + o_xoff, xoff, sizeof(int) * (size_t)(nxcd + 1));
    for (int i = 0; i < nprog; i++) ((int *)(host + o_prog))[i] = FTAMD_PROG_NONE;
    free(tasks);
    void *d = NULL, *dbnd = NULL;
    int rc = ftamd_hip_malloc(&d, total);
    if (rc == FTAMD_OK && bnd_elems > 0)
        rc = ftamd_hip_malloc(&dbnd, (size_t)bnd_elems * (S->precision == 0 ? sizeof(double) : sizeof(float)));
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, host, total, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    free(host);
    if (rc != FTAMD_OK) {
        ftamd_hip_free(d);
        ftamd_hip_free(dbnd);
        return rc;
    }
    char *db = d;
    S->d_mem = d;
    S->d_bnd = dbnd;
    S->launch.bnd = dbnd;
    S->ngroups = G;
    S->ntasks = T;
    FtamdLaunch *a = &S->launch;
    a->groups = (const FtamdGroup *)(db + o_groups);
    a->lane_off = (const int *)(db + o_off);
    a->lane_jmax = (const int *)(db + o_jmax);
    a->tasks = (const int *)(db + o_tasks);
    a->xcd_off = (const int *)(db + o_xoff);
    a->ntasks = T;
    a->nxcd = nxcd;
    a->sc = table;
    a->ntab = ntab;
    a->prog = (int *)(db + o_prog);
    a->nprog = nprog;
    a->ticket = (unsigned *)(db + o_ctl);
    a->done = (unsigned *)(db + o_ctl + FTAMD_MAX_XCD * sizeof(int));
    a->completed = (unsigned *)(db + o_ctl + (FTAMD_MAX_XCD + 1) * sizeof(int));
    a->err = (int *)(db + o_ctl + (FTAMD_MAX_XCD + 2) * sizeof(int));
    return FTAMD_OK;
}

void ftamd_schedule_free_device(ft_schedule *S) {
    if (!S) return;
    ftamd_hip_free(S->d_mem);
    ftamd_hip_free(S->d_bnd);
    ftamd_hip_free(S->d_sys);
    S->d_mem = S->d_bnd = S->d_sys = NULL;
}

/* ------------------------------------------------- launch error words --
 * Every schedule has a device error word its kernels set (FTAMD_ERR_*).  The
 * synchronous entry points clear it before and read it after their launch.
 * Asynchronous launches (ft_gpu_execute, batches) register the schedule here
 * with its stream; ft_gpu_synchronize(stream) waits for the stream and reads
 * and clears the words of the schedules launched on it. */
static pthread_mutex_t g_pend_lock = PTHREAD_MUTEX_INITIALIZER;
static ft_schedule **g_pend = NULL;
static int g_npend = 0, g_cap_pend = 0;

static void pending_add(ft_schedule *S, void *stream) {
    pthread_mutex_lock(&g_pend_lock);
    if (!S->pending) {
        if (g_npend == g_cap_pend) {
            int cap = g_cap_pend ? 2 * g_cap_pend : 16;
            ft_schedule **grown = realloc(g_pend, sizeof(*grown) * (size_t)cap);
            if (!grown) {
                pthread_mutex_unlock(&g_pend_lock);
                return; /* the error word is still read by a later synchronous call */
            }
            g_pend = grown;
            g_cap_pend = cap;
        }
        g_pend[g_npend++] = S;
        S->pending = 1;
    }
    S->pending_stream = stream;
    pthread_mutex_unlock(&g_pend_lock);
}

static void pending_remove_locked(ft_schedule *S) {
    for (int i = 0; i < g_npend; i++)
        if (g_pend[i] == S) {
            g_pend[i] = g_pend[--g_npend];
            break;
        }
    S->pending = 0;
}

/* Frees a schedule (its device memory and the struct). */
void ftamd_schedule_release(ft_schedule *S) {
    if (!S) return;
    pthread_mutex_lock(&g_pend_lock);
    if (S->pending) pending_remove_locked(S);
    pthread_mutex_unlock(&g_pend_lock);
    ftamd_schedule_free_device(S);
    free(S);
}

/* Reads and clears a device error word (stream-ordered). */
static int read_error_at(int *err_dev, void *stream) {
    int err = 0;
    int rc = ftamd_hip_d2h(&err, err_dev, sizeof(int), stream);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(stream);
    if (rc != FTAMD_OK) return rc;
    if (err) {
        static const int zero = 0;
        rc = ftamd_hip_h2d(err_dev, &zero, sizeof(int), stream);
        if (rc == FTAMD_OK) rc = ftamd_hip_sync(stream);
        if (rc != FTAMD_OK) return rc;
        return (err & FTAMD_ERR_TIMEOUT) ? FTAMD_ETIMEOUT : FTAMD_EINCOMPLETE;
    }
    return FTAMD_OK;
}

static int read_error_word(ft_schedule *S, void *stream) { return read_error_at(S->launch.err, stream); }

/* Launches of the allocation-free entry point not yet checked, per device
 * (their error word is the device scratch's first word). */
static struct {
    int *err;
    int pending;
    void *stream;
} g_na_pend[FTAMD_MAX_DEVICES];

int ft_gpu_synchronize(void *stream) {
    int rc = ftamd_hip_sync(stream);
    if (rc != FTAMD_OK) return rc;
    pthread_mutex_lock(&g_pend_lock);
    for (int i = 0; i < g_npend;) {
        ft_schedule *S = g_pend[i];
        if (S->pending_stream != stream) {
            i++;
            continue;
        }
        int r = read_error_word(S, stream);
        if (rc == FTAMD_OK) rc = r;
        pending_remove_locked(S); /* swaps the last entry into slot i */
    }
    int dev = -1;
    if (ftamd_hip_device(&dev, NULL) == FTAMD_OK && dev >= 0 && dev < FTAMD_MAX_DEVICES && g_na_pend[dev].pending &&
        g_na_pend[dev].stream == stream) {
        int r = read_error_at(g_na_pend[dev].err, stream);
        if (rc == FTAMD_OK) rc = r;
        g_na_pend[dev].pending = 0;
    }
    pthread_mutex_unlock(&g_pend_lock);
    return rc;
}

/* Columns of one n x M matrix, ordered so that equal-need columns are adjacent:
 * sph, sphv, rectdisk and disk by order parity then increasing order, tri by
 * increasing order.  Column j of order m: sph and disk j = 2m-1, 2m
 * (drivers.c:94-110, 414-430), sphv j = 2m+1, 2m+2 (drivers.c:204-220),
 * rectdisk and tri j = m (drivers.c:496-508, 314-326). */
static int matrix_columns(int kind, int layout, int n, int M, void *dA, ft_col **out, int *ncols) {
    /* the kernels index a matrix with 32-bit element offsets */
    if ((long long)n * M > 0x7fffffffLL) return FTAMD_EINVAL;
    ft_col *cols = malloc(sizeof(ft_col) * (size_t)(M > 0 ? M : 1));
    if (!cols) return FTAMD_ENOMEM;
    int k = 0;
    if (kind == FTAMD_TRI) {
        for (int m = 1; m < M; m++) {
            ft_col c = {dA, n, 1, (long long)n * m, m - 1, 0};
            cols[k++] = c;
        }
    } else {
        for (int par = 0; par < 2; par++)
            for (int j = 2; j < M; j++) {
                int m = layout == FTAMD_LAYOUT_RECTDISK ? j : layout == FTAMD_LAYOUT_SPHV ? (j - 1) / 2 : (j + 1) / 2;
                if (m < 2 || m % 2 != par) continue;
                /* sphv: orders m <= M/2-1 only (the last column of an even M is not one of theirs) */
                if (layout == FTAMD_LAYOUT_SPHV && m > M / 2 - 1) continue;
                ft_col c = {dA, n, 1, (long long)n * j, m - 2, par};
                cols[k++] = c;
            }
    }
    *out = cols;
    *ncols = k;
    return FTAMD_OK;
}

static double cols_flops(int kind, const ft_col *cols, int ncols) {
    double r = 0;
    for (int i = 0; i < ncols; i++) r += column_rotations(kind, cols[i].n, cols[i].jmin, cols[i].jmax);
    return 6.0 * r;
}

static int build_from_columns(ft_schedule *S, int kind, int dir, int L, int W, const ft_col *cols, int ncols,
                              const void *table, int ntab) {
    ft_hgroup *gs = NULL;
    int G = 0, nprog = 0;
    int rc = make_groups(cols, ncols, kind, dir, L, W, &gs, &G, &nprog);
    if (rc != FTAMD_OK) return rc;
    rc = finish_schedule(S, gs, G, nprog, table, ntab);
    free(gs);
    S->flops = cols_flops(kind, cols, ncols);
    return rc;
}

/* Schedule for the whole matrix (cached in the plan; caller holds RP->lock). */
static ft_schedule *matrix_schedule(ft_rotation_plan *RP, int layout, int dir, int precision, int M, void *dA,
                                    int *rc) {
    ftamd_choice ch;
    ftamd_choose(RP->kind, precision, dir, 0, RP->n, &ch);
    const int L = ch.L, W = ch.W;
    for (ft_schedule *s = RP->scheds; s; s = s->next)
        if (!s->single && !s->sys && s->layout == layout && s->M == M && s->dir == dir && s->precision == precision &&
            s->L == L && s->W == W && s->wg_per_cu == ch.wg_per_cu) {
            /* Matrix pointer may change between calls: groups hold it. */
            return s;
        }
    ft_schedule *S = calloc(1, sizeof(ft_schedule));
    if (!S) {
        *rc = FTAMD_ENOMEM;
        return NULL;
    }
    S->M = M;
    S->layout = layout;
    S->dir = dir;
    S->precision = precision;
    S->L = L;
    S->W = W;
    S->wg_per_cu = ch.wg_per_cu;
    ft_col *cols = NULL;
    int ncols = 0;
    *rc = matrix_columns(RP->kind, layout, RP->n, M, dA, &cols, &ncols);
    if (*rc == FTAMD_OK)
        *rc = build_from_columns(S, RP->kind, dir, L, W, cols, ncols, ftamd_plan_table(RP, precision), RP->n);
    free(cols);
    if (*rc != FTAMD_OK) {
        free(S);
        return NULL;
    }
    S->next = RP->scheds;
    RP->scheds = S;
    return S;
}

/* Slots of a band-systolic launch of nch chain parts: nxcd XCDs of *nsx
 * slots each, about one slot per CU (the kernel's LDS allows one workgroup
 * per CU), fewer if there are fewer parts, and at least as many per XCD as
 * the chain with the most parts has.  Without split chains the slots form
 * one list (nxcd 1) that workgroups on every XCD take in order. */
static void sys_slots(const FtamdSysChain *ch, int nch, int *nxcd, int *nsx) {
    int x = 1, cu = 0;
    if (ftamd_hip_xcd_count(&x) != FTAMD_OK || x < 1) x = 1;
    if (x > FTAMD_MAX_XCD) x = FTAMD_MAX_XCD;
    if (ftamd_hip_device(NULL, &cu) != FTAMD_OK || cu <= 0) cu = 256;
    int per = cu / x > 0 ? cu / x : 1;
    int longest = 1, run = 1;
    for (int c = 1; c < nch; c++) {
        run = (ch[c].split_in >= 0 && ch[c].split_in == ch[c - 1].split_out) ? run + 1 : 1;
        if (run > longest) longest = run;
    }
    if (longest == 1) {
        /* no part waits on another: one list of slots, whatever XCD runs them */
        *nxcd = 1;
        *nsx = nch < cu ? (nch > 0 ? nch : 1) : cu;
        return;
    }
    int need = (nch + x - 1) / x;
    int s = need < per ? need : per;
    if (s < longest) s = longest;
    *nxcd = x;
    *nsx = s > 0 ? s : 1;
}

/* Fills the launch's work-list fields for the arrays at device address d
 * laid out as sys_layout describes. */
typedef struct {
    size_t prog, chains, extra, items, slot_items, xcd_slot, total;
} sys_layout;

/* [256 B control words][progress counters][chains][extra][items][slot_items]
 * [xcd_slot], each part 256-byte aligned. */
static sys_layout sys_layout_of(int nch, int nprog, size_t nextra, int nslots, int nxcd) {
    sys_layout L;
    L.prog = FTAMD_SYS_CTL_BYTES;
    L.chains = L.prog + align256(sizeof(int) * (size_t)(nprog > 0 ? nprog : 1));
    L.extra = L.chains + align256(sizeof(FtamdSysChain) * (size_t)(nch > 0 ? nch : 1));
    L.items = L.extra + align256(nextra);
    L.slot_items = L.items + align256(sizeof(FtamdSysItem) * (size_t)(nch > 0 ? nch : 1));
    L.xcd_slot = L.slot_items + align256(sizeof(int) * (size_t)(nslots + 1));
    L.total = L.xcd_slot + sizeof(int) * (size_t)(nxcd + 1);
    return L;
}

static void sys_fill_launch(FtamdSysLaunch *a, char *d, const sys_layout *L, int nch, int nprog, int nslots,
                            int nxcd, int nsx) {
    a->ctl = (int *)d;
    a->prog = (int *)(d + L->prog);
    a->nprog = nprog;
    a->chains = (const FtamdSysChain *)(d + L->chains);
    a->nchains = nch;
    a->items = (const FtamdSysItem *)(d + L->items);
    a->slot_items = (const int *)(d + L->slot_items);
    a->xcd_slot = (const int *)(d + L->xcd_slot);
    a->nslots = nslots;
    a->nxcd = nxcd;
    a->grid = nxcd * nsx;
}

/* Deals the chain parts to slots and uploads the work list (sys_layout;
 * extra: nextra bytes the caller places there, for example a batch's merged
 * bands) into S->d_sys.  Fills the launch's work-list pointers; *extra_dev
 * gets the extra bytes' address. */
static int sys_upload(ft_schedule *S, const FtamdSysChain *ch, int nch, int nprog, const FtamdSysBand *hbands,
                      const void *extra, size_t nextra, void **extra_dev) {
    int nxcd = 1, nsx = 1;
    sys_slots(ch, nch, &nxcd, &nsx);
    const int nslots = nxcd * nsx;
    const sys_layout L = sys_layout_of(nch, nprog, nextra, nslots, nxcd);
    FtamdSysItem *items = malloc(sizeof(FtamdSysItem) * (size_t)(nch > 0 ? nch : 1));
    int *slot_items = malloc(sizeof(int) * (size_t)(nslots + 1));
    int xcd_slot[FTAMD_MAX_XCD + 1];
    long *ws_load = malloc(sizeof(long) * ((size_t)nslots * (1 + FTAMD_SYS_WAVES) + (size_t)nxcd));
    int *ws_chain = malloc(sizeof(int) * 2 * (size_t)(nch > 0 ? nch : 1));
    int rc = (items && slot_items && ws_load && ws_chain) ? FTAMD_OK : FTAMD_ENOMEM;
    if (rc == FTAMD_OK)
        rc = ftamd_sys_assign(ch, nch, hbands, nxcd, nsx, items, slot_items, xcd_slot, ws_load, ws_chain);
    char *d = NULL;
    if (rc == FTAMD_OK) rc = ftamd_hip_malloc((void **)&d, L.total);
    if (rc == FTAMD_OK) rc = ftamd_hip_memset(d, 0, L.chains); /* control words and progress counters */
    if (rc == FTAMD_OK && nch > 0) rc = ftamd_hip_h2d(d + L.chains, ch, sizeof(FtamdSysChain) * (size_t)nch, NULL);
    if (rc == FTAMD_OK && nextra > 0) rc = ftamd_hip_h2d(d + L.extra, extra, nextra, NULL);
    if (rc == FTAMD_OK && nch > 0) rc = ftamd_hip_h2d(d + L.items, items, sizeof(FtamdSysItem) * (size_t)nch, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d + L.slot_items, slot_items, sizeof(int) * (size_t)(nslots + 1), NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d + L.xcd_slot, xcd_slot, sizeof(int) * (size_t)(nxcd + 1), NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    free(items);
    free(slot_items);
    free(ws_load);
    free(ws_chain);
    if (rc != FTAMD_OK) {
        ftamd_hip_free(d);
        return rc;
    }
    S->d_sys = d;
    sys_fill_launch(&S->sys_launch, d, &L, nch, nprog, nslots, nxcd, nsx);
    S->launch.err = S->sys_launch.ctl + FTAMD_SYS_CTL_ERR;
    if (extra_dev) *extra_dev = d + L.extra;
    return FTAMD_OK;
}

/* Whether both columns of every chain row share one order (sph, sphv and
 * disk pair an order's two columns; tri and rectdisk have one per order). */
static int sys_rowmask(int kind, int layout) { return kind != FTAMD_TRI && layout != FTAMD_LAYOUT_RECTDISK; }

/* Band-systolic schedule of the whole matrix (cached in the plan; caller
 * holds RP->lock).  Returns NULL with *rc == FTAMD_OK where the engine does
 * not apply. */
static ft_schedule *sys_schedule(ft_rotation_plan *RP, int layout, int dir, int precision, int M, int *rc) {
    *rc = FTAMD_OK;
    int Sq = ftamd_sys_choose_S(RP->kind, RP->n, precision);
    if (Sq == 0 || (long long)RP->n * M > 0x7fffffffLL) return NULL;
    ftamd_choice choice;
    ftamd_choose(RP->kind, precision, dir, 0, RP->n, &choice);
    const int maxpart = choice.maxpart;
    for (ft_schedule *s = RP->scheds; s; s = s->next)
        if (s->sys && s->layout == layout && s->M == M && s->dir == dir && s->precision == precision &&
            s->maxpart == maxpart)
            return s;
    if (!RP->sys[dir][precision]) {
        *rc = ftamd_sys_tables_build(RP, dir, precision, Sq, &RP->sys[dir][precision]);
        if (*rc != FTAMD_OK) return NULL;
    }
    const ft_sys_tables *T = RP->sys[dir][precision];
    FtamdSysChain *ch = NULL;
    int nch = 0, nprog = 0;
    *rc = ftamd_sys_chains(RP, T, M, layout, maxpart, &ch, &nch, &nprog);
    if (*rc != FTAMD_OK) return NULL;
    ft_schedule *S = calloc(1, sizeof(ft_schedule));
    if (!S) {
        free(ch);
        *rc = FTAMD_ENOMEM;
        return NULL;
    }
    S->sys = 1;
    S->M = M;
    S->layout = layout;
    S->dir = dir;
    S->precision = precision;
    S->maxpart = maxpart;
    *rc = sys_upload(S, ch, nch, nprog, ftamd_sys_tables_host_bands(T), NULL, 0, NULL);
    free(ch);
    if (*rc != FTAMD_OK) {
        free(S);
        return NULL;
    }
    FtamdSysLaunch *a = &S->sys_launch;
    a->bands = ftamd_sys_tables_dev_bands(T);
    a->rowmask = sys_rowmask(RP->kind, layout);
    a->tab = ftamd_sys_tables_dev_tab(T);
    S->L = ftamd_sys_tables_S(T);
    S->next = RP->scheds;
    RP->scheds = S;
    return S;
}



/* dA: the matrix of every group (single-matrix schedules), or NULL to use the
 * matrices recorded in the groups (batches). */
static int launch(ft_schedule *S, int precision, int kind, int dir, void *dA, void *stream) {
    if (S->sys) {
        FtamdSysLaunch b = S->sys_launch;
        if (dA) b.A = dA;  /* batches keep their own base */
        int spin = KNOB_GET(g_spin_limit);
        b.spin_limit = spin ? spin : 1 << 22;
        b.trace = NULL;
        return ftamd_hip_sys_launch(&b, precision, S->L, stream);
    }
    FtamdLaunch a = S->launch;
    a.trace = NULL;
    a.A = dA;
    return ftamd_hip_launch(&a, precision, kind, dir, S->L, S->W, ftamd_tuning_grid(S), stream);
}

/* Synchronous launches: clear the schedule's error word first (a word an
 * unchecked asynchronous launch left set is not this call's), launch, wait,
 * read it. */
static int launch_checked(ft_schedule *S, int precision, int kind, int dir, void *dA) {
    static const int zero = 0;
    pthread_mutex_lock(&g_pend_lock);
    if (S->pending) pending_remove_locked(S);
    pthread_mutex_unlock(&g_pend_lock);
    int rc = ftamd_hip_h2d(S->launch.err, &zero, sizeof(int), NULL);
    if (rc == FTAMD_OK) rc = launch(S, precision, kind, dir, dA, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    if (rc == FTAMD_OK) rc = read_error_word(S, NULL);
    return rc;
}

/* --------------------------------------------------------- executors -- */

/* Table geometry (FTAMD_SPH, _TRI, _DISK) of a public kind (FT_GPU_SPH ..
 * FT_GPU_RECTDISK) and its column layout; -1 for an unknown kind.  sphv and
 * rectdisk run the sphere's kernels (drivers.c:204-220, rotations_default.c:73-85). */
static int kind_geometry(int kind, int *layout) {
    *layout = FTAMD_LAYOUT_STD;
    switch (kind) {
    case FT_GPU_SPH: return FTAMD_SPH;
    case FT_GPU_TRI: return FTAMD_TRI;
    case FT_GPU_DISK: return FTAMD_DISK;
    case FT_GPU_SPHV: *layout = FTAMD_LAYOUT_SPHV; return FTAMD_SPH;
    case FT_GPU_RECTDISK: *layout = FTAMD_LAYOUT_RECTDISK; return FTAMD_SPH;
    }
    return -1;
}

/* Matrix widths each executor accepts: every column it touches exists and
 * every sequence it applies is in the plan (n x n tables; disk n x 2n-1). */
static int valid_M(int kind, int n, int M) { return ftamd_valid_M(kind, n, M); }
int ftamd_valid_M(int kind, int n, int M) {
    if (M < 0) return 0;
    switch (kind) {
    case FT_GPU_TRI:
    case FT_GPU_RECTDISK: return M <= n;
    case FT_GPU_SPH: return M == 0 || (M % 2 == 1 && M <= 2 * n - 1);
    case FT_GPU_SPHV: return M == 0 || (M % 2 == 1 && M <= 2 * n + 1);
    case FT_GPU_DISK: return M == 0 || (M % 2 == 1 && M <= 4 * n - 3);
    }
    return 0;
}

/* A plan of the kind's table geometry and a valid width (the reference does
 * not check; a plan of another geometry would be read out of bounds).
 * Sets *geo and *layout. */
static int valid_call(const ft_rotation_plan *RP, int kind, int M, int *geo, int *layout) {
    *geo = kind_geometry(kind, layout);
    return RP && *geo >= 0 && RP->kind == *geo && valid_M(kind, RP->n, M);
}

/* Schedule for one launch on RP (caller holds no lock). */
static int pick_schedule(ft_rotation_plan *RP, int layout, int dir, int precision, void *dA, int M,
                         ft_schedule **out) {
    int rc = FTAMD_OK;
    ftamd_choice ch;
    ftamd_choose(RP->kind, precision, dir, 0, RP->n, &ch);
    pthread_mutex_lock(&RP->lock);
    rc = ftamd_plan_upload(RP, precision);
    ft_schedule *S = NULL;
    if (rc == FTAMD_OK && ch.engine == FT_GPU_ENGINE_SYSTOLIC) {
        S = sys_schedule(RP, layout, dir, precision, M, &rc);
        if (rc == FTAMD_OK && !S && ch.forced) rc = FTAMD_EINVAL;
    }
    if (rc == FTAMD_OK && !S) S = matrix_schedule(RP, layout, dir, precision, M, dA, &rc);
    pthread_mutex_unlock(&RP->lock);
    *out = S;
    if (rc == FTAMD_OK) t_last_engine = S->sys ? FT_GPU_ENGINE_SYSTOLIC : FT_GPU_ENGINE_PIPELINE;
    return rc;
}

int ft_gpu_execute(const ft_rotation_plan *RP_, int kind, int direction, int precision, void *dA, int M,
                   void *stream) {
    ft_rotation_plan *RP = (ft_rotation_plan *)RP_;
    int geo, layout;
    if (!valid_call(RP, kind, M, &geo, &layout) || !dA || (direction != 0 && direction != 1) ||
        (precision != 0 && precision != 1))
        return FT_GPU_EINVAL;
    if (M == 0) return FT_GPU_OK;
    ft_schedule *S = NULL;
    int rc = pick_schedule(RP, layout, direction, precision, dA, M, &S);
    if (rc != FTAMD_OK) return rc;
    rc = launch(S, precision, RP->kind, direction, dA, stream);
    if (rc == FTAMD_OK) pending_add(S, stream);
    return rc;
}

/*
 * The allocation-free variant of ft_gpu_execute (no reference counterpart).
 * It runs the band-systolic engine with its work list and error word in the
 * library's fixed device scratch array (FTAMD_NOALLOC_SCRATCH_BYTES per
 * device, ft_noalloc.hip) instead of a schedule the library allocates: the
 * chain records are built on the host into a fixed host buffer and copied
 * into the scratch on `stream`, then the kernel runs.  It calls no hipMalloc
 * or hipFree.  The plan's band tables for this direction and precision must
 * already be resident (any earlier executor call on the plan with the same
 * direction and precision leaves them so); shapes the engine does not run, or
 * whose chain records exceed the scratch, are rejected with FT_GPU_EINVAL.
 * Every call on one device shares its scratch: calls that overlap on a device
 * (other streams, other threads) race.
 */
static pthread_mutex_t g_na_lock = PTHREAD_MUTEX_INITIALIZER;
#define NA_MAX_SLOTS 1024
#define NA_MAX_CHAINS                                                                                            \
    ((FTAMD_NOALLOC_SCRATCH_BYTES - FTAMD_SYS_CTL_BYTES - 6 * 256 - sizeof(int) * (NA_MAX_SLOTS + FTAMD_MAX_XCD + 2)) / \
     (sizeof(FtamdSysChain) + sizeof(FtamdSysItem) + sizeof(int)))
static FtamdSysChain g_na_chains[NA_MAX_CHAINS];
static FtamdSysItem g_na_items[NA_MAX_CHAINS];
static int g_na_slot_items[NA_MAX_SLOTS + 1];
static int g_na_xcd_slot[FTAMD_MAX_XCD + 1];
static long g_na_ws_load[NA_MAX_SLOTS * (1 + FTAMD_SYS_WAVES) + FTAMD_MAX_XCD];
static int g_na_ws_chain[2 * NA_MAX_CHAINS];
static const int g_na_zeros[NA_MAX_CHAINS];

int ft_gpu_execute_noalloc(const ft_rotation_plan *RP_, int kind, int direction, int precision, void *dA, int M,
                           void *stream) {
    ft_rotation_plan *RP = (ft_rotation_plan *)RP_;
    int geo, layout;
    if (!valid_call(RP, kind, M, &geo, &layout) || !dA || (direction != 0 && direction != 1) ||
        (precision != 0 && precision != 1))
        return FT_GPU_EINVAL;
    if (M == 0) return FT_GPU_OK;
    if ((long long)RP->n * M > 0x7fffffffLL) return FT_GPU_EINVAL;
    int dev = -1, rc = ftamd_hip_device(&dev, NULL);
    if (rc != FTAMD_OK) return rc;
    if (dev < 0 || dev >= FTAMD_MAX_DEVICES) return FT_GPU_EDEVICE;
    pthread_mutex_lock(&RP->lock);
    const ft_sys_tables *T = RP->device == dev ? RP->sys[direction][precision] : NULL;
    pthread_mutex_unlock(&RP->lock);
    if (!T) return FT_GPU_EINVAL; /* tables not resident on this device */
    void *scr = NULL;
    size_t scr_bytes = 0;
    rc = ftamd_hip_scratch(&scr, &scr_bytes);
    if (rc != FTAMD_OK) return rc;
    pthread_mutex_lock(&g_na_lock);
    /* the scratch holds the work list as a schedule does (sys_layout) */
    int nfull = 0, nch = 0, nprog = 0, nxcd = 1, nsx = 1;
    rc = ftamd_sys_chains_fill(RP, T, M, layout, g_na_chains, (int)NA_MAX_CHAINS, &nfull);
    if (rc == FTAMD_OK) {
        ftamd_choice ch;
        ftamd_choose(geo, precision, direction, 0, RP->n, &ch);
        rc = ftamd_sys_split(g_na_chains, nfull, (int)NA_MAX_CHAINS, ch.maxpart, &nch, &nprog);
    }
    if (rc == FTAMD_OK) {
        sys_slots(g_na_chains, nch, &nxcd, &nsx);
        if (nxcd * nsx > NA_MAX_SLOTS) rc = FTAMD_EINVAL;
    }
    if (rc == FTAMD_OK)
        rc = ftamd_sys_assign(g_na_chains, nch, ftamd_sys_tables_host_bands(T), nxcd, nsx, g_na_items,
                              g_na_slot_items, g_na_xcd_slot, g_na_ws_load, g_na_ws_chain);
    if (rc != FTAMD_OK) rc = FTAMD_EINVAL; /* the work list does not fit */
    const int nslots = nxcd * nsx;
    const sys_layout L = sys_layout_of(nch, nprog, 0, nslots, nxcd);
    char *d = (char *)scr;
    if (rc == FTAMD_OK && L.total > scr_bytes) rc = FTAMD_EINVAL;
    /* progress counters start at zero (earlier calls may have left other
     * records there); the control words are zero after every launch */
    if (rc == FTAMD_OK && nprog > 0) rc = ftamd_hip_h2d(d + L.prog, g_na_zeros, sizeof(int) * (size_t)nprog, stream);
    if (rc == FTAMD_OK && nch > 0) rc = ftamd_hip_h2d(d + L.chains, g_na_chains, sizeof(FtamdSysChain) * (size_t)nch, stream);
    if (rc == FTAMD_OK && nch > 0) rc = ftamd_hip_h2d(d + L.items, g_na_items, sizeof(FtamdSysItem) * (size_t)nch, stream);
    if (rc == FTAMD_OK)
        rc = ftamd_hip_h2d(d + L.slot_items, g_na_slot_items, sizeof(int) * (size_t)(nslots + 1), stream);
    if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d + L.xcd_slot, g_na_xcd_slot, sizeof(int) * (size_t)(nxcd + 1), stream);
    if (rc == FTAMD_OK && nch > 0) {
        FtamdSysLaunch b;
        memset(&b, 0, sizeof(b));
        sys_fill_launch(&b, d, &L, nch, nprog, nslots, nxcd, nsx);
        b.bands = ftamd_sys_tables_dev_bands(T);
        b.rowmask = sys_rowmask(geo, layout);
        b.tab = ftamd_sys_tables_dev_tab(T);
        b.A = dA;
        int spin = KNOB_GET(g_spin_limit);
        b.spin_limit = spin ? spin : 1 << 22;
        rc = ftamd_hip_sys_launch(&b, precision, ftamd_sys_tables_S(T), stream);
    }
    pthread_mutex_unlock(&g_na_lock);
    if (rc == FTAMD_OK) {
        t_last_engine = FT_GPU_ENGINE_SYSTOLIC;
        pthread_mutex_lock(&g_pend_lock);
        g_na_pend[dev].err = (int *)scr + FTAMD_SYS_CTL_ERR;
        g_na_pend[dev].pending = 1;
        g_na_pend[dev].stream = stream;
        pthread_mutex_unlock(&g_pend_lock);
    }
    return rc;
}

/* Grows the plan's staging buffer to `bytes` (caller holds RP->exec_lock). */
static int stage_reserve(ft_rotation_plan *RP, size_t bytes) {
    if (RP->stage_bytes >= bytes) return FTAMD_OK;
    ftamd_hip_free(RP->d_stage);
    RP->d_stage = NULL;
    RP->stage_bytes = 0;
    int rc = ftamd_hip_malloc(&RP->d_stage, bytes);
    if (rc == FTAMD_OK) RP->stage_bytes = bytes;
    return rc;
}

/* The reference entry points: synchronous, host or device A.  The reference
 * only reads the plan, so concurrent calls on one plan with distinct A are
 * safe there (drivers.c:120-124); here every call holds the plan's exec_lock
 * from staging to the last copy back, so calls on one plan run one after
 * another and never share the staging buffer or a schedule's workspace. */
static void execute(const ft_rotation_plan *RP_, int kind, int dir, double *A, int M) {
    ft_rotation_plan *RP = (ft_rotation_plan *)RP_;
    int geo, layout;
    if (!valid_call(RP, kind, M, &geo, &layout) || !A) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    if (M == 0) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    size_t bytes = (size_t)RP->n * (size_t)M * sizeof(double);
    ft_schedule *S = NULL;
    pthread_mutex_lock(&RP->exec_lock);
    void *d = A;
    if (pk != FTAMD_PTR_DEVICE) {
        rc = stage_reserve(RP, bytes);
        d = RP->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, A, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = pick_schedule(RP, layout, dir, 0, d, M, &S);
    if (rc == FTAMD_OK) rc = launch_checked(S, 0, RP->kind, dir, d);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, d, bytes, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&RP->exec_lock);
    ftamd_set_error(rc);
}

/* drivers.c:58-92 */
void ft_execute_sph_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_SPH, 0, A, M);
}
void ft_execute_sph_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_SPH, 1, A, M);
}
/* drivers.c:168-202 */
void ft_execute_sphv_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_SPHV, 0, A, M);
}
void ft_execute_sphv_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_SPHV, 1, A, M);
}
/* drivers.c:278-312 */
void ft_execute_tri_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_TRI, 0, A, M);
}
void ft_execute_tri_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_TRI, 1, A, M);
}
/* drivers.c:378-412 */
void ft_execute_disk_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_DISK, 0, A, M);
}
void ft_execute_disk_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_DISK, 1, A, M);
}
/* drivers.c:488-494 */
void ft_execute_rectdisk_hi2lo(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_RECTDISK, 0, A, M);
}
void ft_execute_rectdisk_lo2hi(const ft_rotation_plan *RP, double *A, double *B, const int M) {
    (void)B;
    execute(RP, FT_GPU_RECTDISK, 1, A, M);
}

/* ------------------------------------------------------ tetrahedron -- */

/* The block rotations of slabs of L < n columns read the triangle's (s, c)
 * for l + j <= L - 2 only, which are those of the triangle plan of size L
 * with RP's parameters (the entries do not depend on the size,
 * rotations.c:143-160): that plan, kept in RP's list and built on first use
 * (caller holds RP->lock). */
static ft_rotation_plan *tet_block_plan(ft_rotation_plan *RP, int L, int *rc) {
    *rc = FTAMD_OK;
    if (L == RP->n) return RP;
    for (ft_rotation_plan *p = RP->subs; p; p = p->sub_next)
        if (p->n == L) return p;
    ft_rotation_plan *p = ft_plan_rottriangle(L, RP->alpha, RP->beta, RP->gamma);
    if (!p) {
        *rc = FTAMD_ENOMEM;
        return NULL;
    }
    p->sub_next = RP->subs;
    RP->subs = p;
    return p;
}

/* The band-systolic schedule of a tetrahedron stage (ftamd_sys_chains_tet_fill)
 * on `count` slabs of N rows and L columns from order m0, on P, the plan whose
 * tables it runs: cached in P if `cache`, else the caller's to release
 * (caller holds no lock). */
static int tet_schedule(ft_rotation_plan *P, int part, int N, int L, int m0, int count, int dir, int cache,
                        ft_schedule **out) {
    *out = NULL;
    const int layout = part == 0 ? FTAMD_LAYOUT_TET : FTAMD_LAYOUT_TETBLOCK;
    const int Sq = ftamd_sys_choose_S(FTAMD_TRI, P->n, 0);
    if (Sq == 0) return FTAMD_EINVAL; /* a build without the band-systolic engine */
    ftamd_choice choice;
    ftamd_choose(FTAMD_TRI, 0, dir, 0, P->n, &choice);
    const int maxpart = choice.maxpart;
    pthread_mutex_lock(&P->lock);
    int rc = ftamd_plan_upload(P, 0);
    ft_schedule *S = NULL;
    for (ft_schedule *s = rc == FTAMD_OK && cache ? P->scheds : NULL; s; s = s->next)
        if (s->sys && s->layout == layout && s->M == count && s->dir == dir && s->precision == 0 &&
            s->maxpart == maxpart && s->slab == L && s->order0 == m0 && s->rs == N) {
            S = s;
            break;
        }
    if (rc == FTAMD_OK && !S && !P->sys[dir][0]) rc = ftamd_sys_tables_build(P, dir, 0, Sq, &P->sys[dir][0]);
    if (rc == FTAMD_OK && !S) {
        const ft_sys_tables *T = P->sys[dir][0];
        FtamdSysChain *ch = NULL;
        int nch = 0, nprog = 0;
        rc = ftamd_sys_chains_tet(P, T, part, N, L, m0, count, maxpart, &ch, &nch, &nprog);
        if (rc == FTAMD_OK) {
            S = calloc(1, sizeof(ft_schedule));
            if (!S) rc = FTAMD_ENOMEM;
        }
        if (rc == FTAMD_OK) {
            S->sys = 1;
            S->M = count;
            S->layout = layout;
            S->dir = dir;
            S->precision = 0;
            S->maxpart = maxpart;
            S->slab = L;
            S->order0 = m0;
            S->rs = N;
            rc = sys_upload(S, ch, nch, nprog, ftamd_sys_tables_host_bands(T), NULL, 0, NULL);
            if (rc != FTAMD_OK) {
                free(S);
                S = NULL;
            }
        }
        free(ch);
        if (rc == FTAMD_OK) {
            FtamdSysLaunch *a = &S->sys_launch;
            a->bands = ftamd_sys_tables_dev_bands(T);
            a->rowmask = 0;
            a->tab = ftamd_sys_tables_dev_tab(T);
            S->L = ftamd_sys_tables_S(T);
            if (cache) {
                S->next = P->scheds;
                P->scheds = S;
            }
        }
    }
    pthread_mutex_unlock(&P->lock);
    *out = S;
    return rc;
}

/* A tetrahedron call: triangle plans of one size N (the reference reads RP1's
 * size as the slab's rows and RP2's as their stride, drivers.c:510-533,
 * rotations.c:297-321), 0 <= L <= N (a slab of more columns would read the
 * block rotations beyond RP2's table), M >= 0. */
static int tet_valid(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, int L, int M) {
    return RP1 && RP2 && RP1->kind == FTAMD_TRI && RP2->kind == FTAMD_TRI && RP1->family == FT_GPU_TRI &&
           RP2->family == FT_GPU_TRI && RP1->n == RP2->n && L >= 0 && L <= RP1->n && M >= 0 &&
           (long long)L * M <= 0x7fffffffLL;
}

/* The stages' schedules of a tetrahedron call (S[0] the slab columns on RP1,
 * S[1] the slab rows on RP2's block plan; part_mask selects them), m0 and
 * count slabs, cached or the caller's to release. */
static int tet_schedules(ft_rotation_plan *RP1, ft_rotation_plan *RP2, int part_mask, int L, int m0, int count,
                         int dir, int cache, ft_schedule **S) {
    const int N = RP2->n;
    int rc = FTAMD_OK;
    S[0] = S[1] = NULL;
    if (part_mask & 1) rc = tet_schedule(RP1, 0, N, L, m0, count, dir, cache, &S[0]);
    if (rc == FTAMD_OK && (part_mask & 2)) {
        pthread_mutex_lock(&RP2->lock);
        ft_rotation_plan *B = tet_block_plan(RP2, L, &rc);
        pthread_mutex_unlock(&RP2->lock);
        if (rc == FTAMD_OK) rc = tet_schedule(B, 1, N, L, m0, count, dir, cache, &S[1]);
    }
    return rc;
}

/* Launch order of a tetrahedron call: hi2lo runs the slab columns' rotations,
 * then the block rotations; lo2hi the reverse (drivers.c:510-533). */
static int tet_launch(ft_schedule *const *S, int dir, void *dA, void *stream, int checked) {
    int rc = FTAMD_OK;
    for (int i = 0; i < 2 && rc == FTAMD_OK; i++) {
        ft_schedule *x = S[dir == 0 ? i : 1 - i];
        if (!x) continue;
        if (checked) {
            rc = launch_checked(x, 0, FTAMD_TRI, dir, dA);
        } else {
            rc = launch(x, 0, FTAMD_TRI, dir, dA, stream);
            if (rc == FTAMD_OK) pending_add(x, stream);
        }
    }
    return rc;
}

int ft_gpu_execute_tet(const ft_rotation_plan *RP1_, const ft_rotation_plan *RP2_, int direction, double *dA, int L,
                       int M, void *stream) {
    ft_rotation_plan *RP1 = (ft_rotation_plan *)RP1_, *RP2 = (ft_rotation_plan *)RP2_;
    if (!tet_valid(RP1, RP2, L, M) || !dA || (direction != 0 && direction != 1)) return FT_GPU_EINVAL;
    if (L < 2 || M == 0) return FT_GPU_OK;
    ft_schedule *S[2];
    int rc = tet_schedules(RP1, RP2, 3, L, 0, M, direction, 1, S);
    if (rc == FTAMD_OK) rc = tet_launch(S, direction, dA, stream, 0);
    if (rc == FTAMD_OK) t_last_engine = FT_GPU_ENGINE_SYSTOLIC;
    return rc;
}

/* Both plans' exec locks, in a fixed order (by serial) so that two calls
 * with the plans swapped cannot deadlock. */
static void lock_pair(ft_rotation_plan *a, ft_rotation_plan *b, int lock) {
    ft_rotation_plan *first = a, *second = b;
    if (a == b) second = NULL;
    else if (b->serial < a->serial) first = b, second = a;
    if (lock) {
        pthread_mutex_lock(&first->exec_lock);
        if (second) pthread_mutex_lock(&second->exec_lock);
    } else {
        if (second) pthread_mutex_unlock(&second->exec_lock);
        pthread_mutex_unlock(&first->exec_lock);
    }
}

/* The synchronous tetrahedron calls, host or device A of `slabs` slabs of
 * N x L (staged through RP1's buffer): both stages (part_mask 3, schedules
 * cached) for the executors, the block rotations of one slab (part_mask 2,
 * a schedule per call: callers loop over the orders) for the kernels. */
static void execute_tet(const ft_rotation_plan *RP1_, const ft_rotation_plan *RP2_, int part_mask, int dir, double *A,
                        int L, int m0, int slabs) {
    ft_rotation_plan *RP1 = (ft_rotation_plan *)RP1_, *RP2 = (ft_rotation_plan *)RP2_;
    if (!tet_valid(RP1, RP2, L, m0 + slabs) || !A || m0 < 0) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    if (L < 2 || slabs == 0) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    size_t bytes = (size_t)RP1->n * (size_t)L * (size_t)slabs * sizeof(double);
    lock_pair(RP1, RP2, 1);
    void *d = A;
    if (pk != FTAMD_PTR_DEVICE) {
        rc = stage_reserve(RP1, bytes);
        d = RP1->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, A, bytes, NULL);
    }
    ft_schedule *S[2] = {NULL, NULL};
    const int cache = part_mask == 3;
    if (rc == FTAMD_OK) rc = tet_schedules(RP1, RP2, part_mask, L, m0, slabs, dir, cache, S);
    if (rc == FTAMD_OK) rc = tet_launch(S, dir, d, NULL, 1);
    if (!cache) {
        ftamd_schedule_release(S[0]);
        ftamd_schedule_release(S[1]);
    }
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, d, bytes, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    lock_pair(RP1, RP2, 0);
    ftamd_set_error(rc);
}

/* drivers.c:510-533 */
void ft_execute_tet_hi2lo(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, double *A, const int L,
                          const int M) {
    execute_tet(RP1, RP2, 3, 0, A, L, 0, M);
}
void ft_execute_tet_lo2hi(const ft_rotation_plan *RP1, const ft_rotation_plan *RP2, double *A, const int L,
                          const int M) {
    execute_tet(RP1, RP2, 3, 1, A, L, 0, M);
}

/* rotations.c:297-321: the block rotations of the N x L slab A (N = RP->n)
 * of order m. */
void ft_kernel_tet_hi2lo(const ft_rotation_plan *RP, const int L, const int m, double *A) {
    execute_tet(RP, RP, 2, 0, A, L, m, 1);
}
void ft_kernel_tet_lo2hi(const ft_rotation_plan *RP, const int L, const int m, double *A) {
    execute_tet(RP, RP, 2, 1, A, L, m, 1);
}

/* ------------------------------------------------- single vectors -- */

/* Single-vector schedules a plan keeps (the reference's callers loop over
 * orders with the same few (m1, m2, S)); the oldest beyond this are freed. */
#define FTAMD_SINGLE_CACHE 32

/* The cached single-vector schedule of (dir, jlo, jhi, S) on RP, built on
 * first use; the column is a launch argument.  Caller holds RP->exec_lock. */
static int single_schedule(ft_rotation_plan *RP, int kind, int dir, int jlo, int jhi, int S, ft_schedule **out) {
    ftamd_choice ch;
    ftamd_choose(kind, 0, dir, 0, RP->n, &ch);
    const int L = ch.L, W = ch.W;
    int rc = FTAMD_OK;
    pthread_mutex_lock(&RP->lock);
    ft_schedule *prev = NULL, *hit = NULL;
    int count = 0;
    for (ft_schedule *s = RP->scheds; s; prev = s, s = s->next) {
        if (!s->single) continue;
        if (!hit && s->dir == dir && s->m1 == jlo && s->m2 == jhi && s->rs == S && s->L == L && s->W == W) {
            hit = s;
            /* move to the front: the list's tail holds the least recently used */
            if (prev) {
                prev->next = s->next;
                s->next = RP->scheds;
                RP->scheds = s;
            }
            break;
        }
    }
    if (!hit) {
        rc = ftamd_plan_upload(RP, 0);
        ft_schedule *Sd = rc == FTAMD_OK ? calloc(1, sizeof(ft_schedule)) : NULL;
        if (rc == FTAMD_OK && !Sd) rc = FTAMD_ENOMEM;
        if (rc == FTAMD_OK) {
            Sd->single = 1;
            Sd->dir = dir;
            Sd->m1 = jlo;
            Sd->m2 = jhi;
            Sd->rs = S;
            Sd->L = L;
            Sd->W = W;
            ft_col c = {NULL, RP->n, S, 0, jhi, jlo};
            rc = build_from_columns(Sd, kind, dir, L, W, &c, 1, ftamd_plan_table(RP, 0), RP->n);
            if (rc != FTAMD_OK) {
                ftamd_schedule_free_device(Sd);
                free(Sd);
            } else {
                Sd->next = RP->scheds;
                RP->scheds = Sd;
                hit = Sd;
            }
        }
        /* evict the least recently used single-vector schedules beyond the cap */
        ft_schedule **link = &RP->scheds;
        count = 0;
        while (*link) {
            ft_schedule *s = *link;
            if (s->single && ++count > FTAMD_SINGLE_CACHE) {
                *link = s->next;
                ftamd_schedule_release(s);
                continue;
            }
            link = &s->next;
        }
    }
    pthread_mutex_unlock(&RP->lock);
    *out = hit;
    return rc;
}

/*
 * ft_kernel_<kind>_<dir>(RP, m1, m2, A, S) (rotations.c:135-141, 162-168,
 * 262-268; loops of rotations_default.c:31-71): hi2lo runs j = m2-step,
 * m2-2*step, ... >= m1; lo2hi runs j = m1, m1+step, ... < m2.  Sequences whose
 * row range is empty (j beyond the plan) do nothing, as in the reference.
 */
/* The sequences kernel_<kind>_<dir>(RP, m1, m2, ., S) applies: *jlo .. *jhi
 * in steps of the geometry's (empty if *jhi < *jlo), those with rows only. */
static void kernel_range(const ft_rotation_plan *RP, int kind, int dir, int m1, int m2, int *jlo_, int *jhi_) {
    int step = geo_step(kind), n = RP->n;
    int jlo, jhi;
    if (dir == 0) {
        /* j = m2-step, m2-2*step, ... >= m1: the parity of m2, not of m1 */
        jhi = m2 - step;
        jlo = jhi >= m1 ? jhi - ((jhi - m1) / step) * step : m1;
    } else {
        jlo = m1;
        jhi = m1 + ((m2 - 1 - m1) / step) * step;
        if (m2 <= m1) jhi = jlo - 1;
    }
    /* keep sequences with rows; j must also index an existing table column */
    int jcap = kind == FTAMD_DISK ? 2 * n - 2 : n - 1;
    while (jhi >= jlo && (geo_top(kind, n, jhi) < 0 || jhi > jcap)) jhi -= step;
    *jlo_ = jlo;
    *jhi_ = jhi;
}

/* One vector at device address d with stride S (caller holds RP->exec_lock). */
static int kernel_dev(ft_rotation_plan *RP, int kind, int dir, int m1, int m2, void *d, int S) {
    int jlo, jhi;
    kernel_range(RP, kind, dir, m1, m2, &jlo, &jhi);
    if (jhi < jlo) return FTAMD_OK;
    ft_schedule *Sd = NULL;
    int rc = single_schedule(RP, kind, dir, jlo, jhi, S, &Sd);
    if (rc == FTAMD_OK) rc = launch_checked(Sd, 0, kind, dir, d);
    return rc;
}

static void kernel_one(const ft_rotation_plan *RP_, int kind, int dir, int m1, int m2, double *A, int S) {
    ft_rotation_plan *RP = (ft_rotation_plan *)RP_;
    if (!RP || !A || RP->kind != kind || S < 1 || m1 < 0 || (long long)(RP->n - 1) * S > 0x7fffffffLL) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int n = RP->n, jlo, jhi;
    kernel_range(RP, kind, dir, m1, m2, &jlo, &jhi);
    if (jhi < jlo) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    size_t span = ((size_t)(n - 1) * (size_t)S + 1) * sizeof(double);
    pthread_mutex_lock(&RP->exec_lock);
    void *d = A;
    if (pk != FTAMD_PTR_DEVICE) {
        rc = stage_reserve(RP, span);
        d = RP->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, A, span, NULL);
    }
    if (rc == FTAMD_OK) rc = kernel_dev(RP, kind, dir, m1, m2, d, S);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, d, span, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&RP->exec_lock);
    ftamd_set_error(rc);
}

void ft_kernel_sph_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_SPH, 0, m1, m2, A, S);
}
void ft_kernel_sph_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_SPH, 1, m1, m2, A, S);
}
void ft_kernel_tri_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_TRI, 0, m1, m2, A, S);
}
void ft_kernel_tri_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_TRI, 1, m1, m2, A, S);
}
void ft_kernel_disk_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_DISK, 0, m1, m2, A, S);
}
void ft_kernel_disk_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_DISK, 1, m1, m2, A, S);
}
/* rotations.c:289-295: the sphere's kernels on a rectangle-disk plan
 * (rotations_default.c:73-85) */
void ft_kernel_rectdisk_hi2lo(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_SPH, 0, m1, m2, A, S);
}
void ft_kernel_rectdisk_lo2hi(const ft_rotation_plan *RP, const int m1, const int m2, double *A, const int S) {
    kernel_one(RP, FTAMD_SPH, 1, m1, m2, A, S);
}

/* ----------------------------------------------- spin-weighted sphere -- */

/* The second stage's band-systolic schedule on the spin plan's sphere plan
 * (cached there; caller holds no lock). */
static int spin_schedule(ft_spin_rotation_plan *P, int dir, int M, ft_schedule **out) {
    ft_rotation_plan *R = P->sph;
    *out = NULL;
    const int Sq = ftamd_sys_choose_S(FTAMD_SPH, R->n, 0);
    if (Sq == 0) return FTAMD_EINVAL; /* a build without the band-systolic engine */
    ftamd_choice choice;
    ftamd_choose(FTAMD_SPH, 0, dir, 0, R->n, &choice);
    const int maxpart = choice.maxpart;
    pthread_mutex_lock(&R->lock);
    int rc = ftamd_plan_upload(R, 0);
    ft_schedule *S = NULL;
    for (ft_schedule *x = rc == FTAMD_OK ? R->scheds : NULL; x; x = x->next)
        if (x->sys && x->layout == FTAMD_LAYOUT_SPIN && x->M == M && x->dir == dir && x->precision == 0 &&
            x->maxpart == maxpart && x->order0 == P->s) {
            S = x;
            break;
        }
    if (rc == FTAMD_OK && !S && !R->sys[dir][0]) rc = ftamd_sys_tables_build(R, dir, 0, Sq, &R->sys[dir][0]);
    if (rc == FTAMD_OK && !S) {
        const ft_sys_tables *T = R->sys[dir][0];
        FtamdSysChain *ch = NULL;
        int nch = 0, nprog = 0;
        rc = ftamd_sys_chains_spin(R, T, P->s, M, maxpart, &ch, &nch, &nprog);
        if (rc == FTAMD_OK) {
            S = calloc(1, sizeof(ft_schedule));
            if (!S) rc = FTAMD_ENOMEM;
        }
        if (rc == FTAMD_OK) {
            S->sys = 1;
            S->M = M;
            S->layout = FTAMD_LAYOUT_SPIN;
            S->dir = dir;
            S->precision = 0;
            S->maxpart = maxpart;
            S->order0 = P->s;
            rc = sys_upload(S, ch, nch, nprog, ftamd_sys_tables_host_bands(T), NULL, 0, NULL);
            if (rc != FTAMD_OK) {
                free(S);
                S = NULL;
            }
        }
        free(ch);
        if (rc == FTAMD_OK) {
            FtamdSysLaunch *a = &S->sys_launch;
            a->bands = ftamd_sys_tables_dev_bands(T);
            a->rowmask = 1; /* a lane row is one complex column: one order */
            a->tab = ftamd_sys_tables_dev_tab(T);
            S->L = ftamd_sys_tables_S(T);
            S->next = R->scheds;
            R->scheds = S;
        }
    }
    pthread_mutex_unlock(&R->lock);
    *out = S;
    return rc;
}

/* An n x M complex matrix of the executors' layout: M odd (drivers.c:687-705
 * touches column 2*(M/2), which is M for even M), M <= 2n-1 (orders up to
 * n-1), and 2nM doubles indexable in 32 bits by the chain records. */
static int spin_valid(const ft_spin_rotation_plan *P, int M) {
    return P && P->sph && M >= 0 && (M == 0 || (M % 2 == 1 && M <= 2 * P->n - 1)) &&
           2LL * P->n * M <= 0x7fffffffLL;
}

static int spin_upload(ft_spin_rotation_plan *P) {
    pthread_mutex_lock(&P->lock);
    int rc = ftamd_spin_upload(P);
    pthread_mutex_unlock(&P->lock);
    return rc;
}

/* Both stages on the device matrix d (drivers.c:687-705 with rotations_default.c:
 * 87-117): hi2lo the first stage on columns 1..M-1 (order 0 has none), then
 * the sphere's sequences; lo2hi the reverse.  checked: synchronous, with the
 * error word read. */
static int spin_run(ft_spin_rotation_plan *P, int dir, void *d, int M, void *stream, int checked) {
    ft_schedule *S = NULL;
    int rc = spin_upload(P);
    if (rc == FTAMD_OK) rc = spin_schedule(P, dir, M, &S);
    for (int stage = 0; stage < 2 && rc == FTAMD_OK; stage++) {
        if ((stage == 0) == (dir == 0)) {
            rc = ftamd_hip_spin_stage1(P->d_tab2, P->n, P->s, (double *)d + 2LL * P->n, 1, M - 1, P->n, 1, 0, 0,
                                       dir, checked ? NULL : stream);
        } else if (checked) {
            rc = launch_checked(S, 0, FTAMD_SPH, dir, d);
        } else {
            rc = launch(S, 0, FTAMD_SPH, dir, d, stream);
            if (rc == FTAMD_OK) pending_add(S, stream);
        }
    }
    if (rc == FTAMD_OK && checked) rc = ftamd_hip_sync(NULL);
    return rc;
}

int ft_gpu_execute_spinsph(const ft_spin_rotation_plan *P_, int direction, void *dA, int M, void *stream) {
    ft_spin_rotation_plan *P = (ft_spin_rotation_plan *)P_;
    if (!spin_valid(P, M) || !dA || (direction != 0 && direction != 1)) return FT_GPU_EINVAL;
    if (M == 0) return FT_GPU_OK;
    int rc = spin_run(P, direction, dA, M, stream, 0);
    if (rc == FTAMD_OK) t_last_engine = FT_GPU_ENGINE_SYSTOLIC;
    return rc;
}

/* The synchronous executors: host or device A (n x M ft_complex), staged
 * through the sphere plan's buffer under its execution lock. */
static void execute_spin(const ft_spin_rotation_plan *P_, int dir, ft_complex *A, int M) {
    ft_spin_rotation_plan *P = (ft_spin_rotation_plan *)P_;
    if (!spin_valid(P, M) || !A) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    if (M == 0) return;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    ft_rotation_plan *R = P->sph;
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    size_t bytes = 2 * (size_t)P->n * (size_t)M * sizeof(double);
    pthread_mutex_lock(&R->exec_lock);
    void *d = A;
    if (pk != FTAMD_PTR_DEVICE) {
        rc = stage_reserve(R, bytes);
        d = R->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, A, bytes, NULL);
    }
    if (rc == FTAMD_OK) rc = spin_run(P, dir, d, M, NULL, 1);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, d, bytes, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&R->exec_lock);
    ftamd_set_error(rc);
}

/* drivers.c:645-681 */
void ft_execute_spinsph_hi2lo(const ft_spin_rotation_plan *P, ft_complex *A, ft_complex *B, const int M) {
    (void)B;
    execute_spin(P, 0, A, M);
}
void ft_execute_spinsph_lo2hi(const ft_spin_rotation_plan *P, ft_complex *A, ft_complex *B, const int M) {
    (void)B;
    execute_spin(P, 1, A, M);
}

/* rotations.c:477-483 (rotations_default.c:87-117): one complex vector of
 * order m with stride S (complex entries).  The second stage is the sphere's
 * kernel of orders d .. d%2 (d = |am - as|) on its real and imaginary parts,
 * real vectors of stride 2S. */
static void kernel_spin(const ft_spin_rotation_plan *P_, int dir, int m, ft_complex *A, int S) {
    ft_spin_rotation_plan *P = (ft_spin_rotation_plan *)P_;
    if (!P || !P->sph || !A || S < 1 || 2LL * (P->n - 1) * S + 1 > 0x7fffffffLL || m <= -0x40000000 ||
        m >= 0x40000000) {
        ftamd_set_error(FT_GPU_EINVAL);
        return;
    }
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) {
        ftamd_set_error(rc);
        return;
    }
    ft_rotation_plan *R = P->sph;
    const int as = P->s < 0 ? -P->s : P->s, am = m < 0 ? -m : m;
    const int dd = am > as ? am - as : as - am;
    int pk = FTAMD_PTR_HOST;
    ftamd_hip_pointer_kind(A, &pk);
    size_t span = 2 * ((size_t)(P->n - 1) * (size_t)S + 1) * sizeof(double);
    pthread_mutex_lock(&R->exec_lock);
    void *d = A;
    if (pk != FTAMD_PTR_DEVICE) {
        rc = stage_reserve(R, span);
        d = R->d_stage;
        if (rc == FTAMD_OK) rc = ftamd_hip_h2d(d, A, span, NULL);
    }
    if (rc == FTAMD_OK) rc = spin_upload(P);
    for (int stage = 0; stage < 2 && rc == FTAMD_OK; stage++) {
        if ((stage == 0) == (dir == 0)) {
            rc = ftamd_hip_spin_stage1(P->d_tab2, P->n, P->s, d, 0, 1, 0, S, 1, m, dir, NULL);
            if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
        } else {
            for (int part = 0; part < 2 && rc == FTAMD_OK; part++)
                rc = kernel_dev(R, FTAMD_SPH, dir, dd % 2, dd, (double *)d + part, 2 * S);
        }
    }
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_d2h(A, d, span, NULL);
    if (rc == FTAMD_OK && pk != FTAMD_PTR_DEVICE) rc = ftamd_hip_sync(NULL);
    pthread_mutex_unlock(&R->exec_lock);
    ftamd_set_error(rc);
}

void ft_kernel_spinsph_hi2lo(const ft_spin_rotation_plan *P, const int m, ft_complex *A, const int S) {
    kernel_spin(P, 0, m, A, S);
}
void ft_kernel_spinsph_lo2hi(const ft_spin_rotation_plan *P, const int m, ft_complex *A, const int S) {
    kernel_spin(P, 1, m, A, S);
}

/* ------------------------------------------------------------ batch -- */

/*
 * Many transforms of one kind (and one set of plan parameters) but uneven
 * degree, in one launch.  Band-systolic: the chains of every transform (each
 * plan's own band tables, one merged band array, per-chain matrix offset and
 * column stride), longest first.  Pipelined: the sines and cosines s(l,m), c(l,m) are
 * closed forms in (l, m) only (rotations.c:122-126, 149-153, 179-183), so the
 * table of the largest plan serves every transform; only its packed index
 * uses that plan's n.  Each transform contributes its own groups; all tasks
 * go into the per-XCD, pass-ordered lists, so small transforms fill in around
 * the long chains of the large ones.
 *
 * The schedule of the last argument list is cached.  It is keyed on the whole
 * list, compared entry by entry: every plan pointer with the plan's serial
 * number (a plan created at a freed plan's address gets another serial) and
 * n, every matrix pointer and M, plus kind, direction, precision, engine and
 * tuning (L, W and the band-systolic part size).  Destroying a plan the cached schedule uses drops the schedule.
 */
typedef struct {
    pthread_mutex_t lock;
    ft_schedule *sched;
    int count, kind, dir, precision, L, W, engine, part;
    const ft_rotation_plan **plans;
    unsigned long long *serial;
    int *n;
    void **dA;
    int *M;
} ft_batch_cache;

static ft_batch_cache g_batch = {PTHREAD_MUTEX_INITIALIZER, NULL, 0, 0, 0, 0, 0, 0, 0, 0, NULL, NULL, NULL, NULL, NULL};

static void batch_cache_clear_locked(void) {
    if (g_batch.sched) {
        ftamd_hip_sync(NULL);
        ftamd_schedule_release(g_batch.sched);
    }
    g_batch.sched = NULL;
    free(g_batch.plans);
    free(g_batch.serial);
    free(g_batch.n);
    free(g_batch.dA);
    free(g_batch.M);
    g_batch.plans = NULL;
    g_batch.serial = NULL;
    g_batch.n = NULL;
    g_batch.dA = NULL;
    g_batch.M = NULL;
    g_batch.count = 0;
}

void ftamd_batch_forget_plan(const ft_rotation_plan *RP) {
    pthread_mutex_lock(&g_batch.lock);
    for (int i = 0; g_batch.sched && i < g_batch.count; i++)
        if (g_batch.plans[i] == RP && g_batch.serial[i] == RP->serial) {
            batch_cache_clear_locked();
            break;
        }
    pthread_mutex_unlock(&g_batch.lock);
}

static int batch_cache_hit(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                           void *const *dA, const int *M, int count, int L, int W, int engine, int part) {
    if (!g_batch.sched || g_batch.count != count || g_batch.kind != kind || g_batch.dir != direction ||
        g_batch.precision != precision || g_batch.L != L || g_batch.W != W || g_batch.engine != engine ||
        g_batch.part != part)
        return 0;
    for (int i = 0; i < count; i++)
        if (g_batch.plans[i] != plans[i] || g_batch.serial[i] != plans[i]->serial || g_batch.n[i] != plans[i]->n ||
            g_batch.dA[i] != dA[i] || g_batch.M[i] != M[i])
            return 0;
    return 1;
}

static int batch_cache_store(ft_schedule *S, const ft_rotation_plan *const *plans, int kind, int direction,
                             int precision, void *const *dA, const int *M, int count, int L, int W, int engine,
                             int part) {
    size_t c = (size_t)count;
    g_batch.plans = malloc(sizeof(*g_batch.plans) * c);
    g_batch.serial = malloc(sizeof(*g_batch.serial) * c);
    g_batch.n = malloc(sizeof(int) * c);
    g_batch.dA = malloc(sizeof(void *) * c);
    g_batch.M = malloc(sizeof(int) * c);
    if (!g_batch.plans || !g_batch.serial || !g_batch.n || !g_batch.dA || !g_batch.M) {
        g_batch.sched = S;
        batch_cache_clear_locked();
        return FTAMD_ENOMEM;
    }
    for (int i = 0; i < count; i++) {
        g_batch.plans[i] = plans[i];
        g_batch.serial[i] = plans[i]->serial;
        g_batch.n[i] = plans[i]->n;
        g_batch.dA[i] = dA[i];
        g_batch.M[i] = M[i];
    }
    g_batch.sched = S;
    g_batch.count = count;
    g_batch.kind = kind;
    g_batch.dir = direction;
    g_batch.precision = precision;
    g_batch.L = L;
    g_batch.W = W;
    g_batch.part = part;
    g_batch.engine = engine;
    return FTAMD_OK;
}
/* Validates a batch and uploads the largest plan's table; *big gets that plan. */
/* Validates a batch: plans of one family and one set of parameters (their
 * tables are one closed form), each of the kind's geometry with a valid
 * width.  Sets *geo and *layout of the kind and *big, the largest plan. */
static int batch_check(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                       void *const *dA, const int *M, int count, int *geo, int *layout, ft_rotation_plan **big) {
    if (!plans || !dA || !M || count < 0 || (direction != 0 && direction != 1) || (precision != 0 && precision != 1))
        return FT_GPU_EINVAL;
    *big = NULL;
    *geo = kind_geometry(kind, layout);
    if (*geo < 0) return FT_GPU_EINVAL;
    for (int i = 0; i < count; i++) {
        const ft_rotation_plan *P = plans[i];
        if (!valid_call(P, kind, M[i], geo, layout) || !dA[i]) return FT_GPU_EINVAL;
        if (P->family != plans[0]->family || P->alpha != plans[0]->alpha || P->beta != plans[0]->beta ||
            P->gamma != plans[0]->gamma || P->rho != plans[0]->rho)
            return FT_GPU_EINVAL;
        if (!*big || P->n > (*big)->n) *big = (ft_rotation_plan *)P;
    }
    if (count == 0) return FT_GPU_OK;
    int rc = ftamd_hip_device(NULL, NULL);
    if (rc != FTAMD_OK) return rc;
    pthread_mutex_lock(&(*big)->lock);
    rc = ftamd_plan_upload(*big, precision);
    pthread_mutex_unlock(&(*big)->lock);
    return rc;
}

/* Band-systolic schedule of a validated batch, or NULL (rc FTAMD_OK) where
 * some transform does not fit the engine or the transforms would need
 * different band widths. */
static int batch_build_sys(const ft_rotation_plan *const *plans, int kind, int layout, int direction, int precision,
                           void *const *dA, const int *M, int count, int maxpart, ft_schedule **out) {
    *out = NULL;
    int S0 = 0;
    for (int i = 0; i < count; i++) {
        int Sq = ftamd_sys_choose_S(kind, plans[i]->n, precision);
        if (Sq == 0 || (S0 && Sq != S0) || (long long)plans[i]->n * M[i] > 0x7fffffffLL) return FTAMD_OK;
        S0 = Sq;
    }
    ft_sys_tables **tabs = malloc(sizeof(*tabs) * (size_t)count);
    if (!tabs) return FTAMD_ENOMEM;
    int rc = FTAMD_OK;
    for (int i = 0; rc == FTAMD_OK && i < count; i++) {
        ft_rotation_plan *RP = (ft_rotation_plan *)plans[i];
        pthread_mutex_lock(&RP->lock);
        if (!RP->sys[direction][precision])
            rc = ftamd_sys_tables_build(RP, direction, precision, S0, &RP->sys[direction][precision]);
        tabs[i] = RP->sys[direction][precision];
        pthread_mutex_unlock(&RP->lock);
    }
    FtamdSysBand *bands = NULL;
    FtamdSysChain *ch = NULL;
    int nb = 0, nch = 0, nprog = 0;
    const void *tab_base = NULL;
    if (rc == FTAMD_OK)
        rc = ftamd_sys_batch(plans, tabs, dA, M, count, layout, maxpart, &bands, &nb, &ch, &nch, &nprog, &tab_base);
    free(tabs);
    if (rc != FTAMD_OK) {
        free(bands);
        free(ch);
        return rc;
    }
    ft_schedule *S = calloc(1, sizeof(ft_schedule));
    if (!S) rc = FTAMD_ENOMEM;
    void *dbands = NULL;
    if (rc == FTAMD_OK)
        rc = sys_upload(S, ch, nch, nprog, bands, bands, sizeof(FtamdSysBand) * (size_t)nb, &dbands);
    free(bands);
    free(ch);
    if (rc != FTAMD_OK) {
        free(S);
        return rc;
    }
    S->sys = 1;
    S->layout = layout;
    S->dir = direction;
    S->precision = precision;
    S->maxpart = maxpart;
    S->L = S0;
    FtamdSysLaunch *a = &S->sys_launch;
    a->bands = (const FtamdSysBand *)dbands;
    a->rowmask = sys_rowmask(kind, layout);
    a->tab = tab_base;
    a->A = dA[0];
    *out = S;
    return FTAMD_OK;
}

/* Builds the schedule of a validated batch: the pipelined engine, or the
 * band-systolic one when forced (it keeps one chain per CU, and the many
 * short chains of a batch of small transforms leave most of each CU idle:
 * 20.2 ms against 10.4 ms for 512 sph transforms of degree 32..1018). */
static int batch_build(const ft_rotation_plan *const *plans, int kind, int layout, int direction, int precision,
                       void *const *dA, const int *M, int count, ft_rotation_plan *big, const ftamd_choice *ch,
                       ft_schedule **out) {
    if (ch->engine == FT_GPU_ENGINE_SYSTOLIC) {
        int rs = batch_build_sys(plans, kind, layout, direction, precision, dA, M, count, ch->maxpart, out);
        if (rs != FTAMD_OK) return rs;
        if (*out) return FTAMD_OK;
        if (ch->forced) return FTAMD_EINVAL;
        /* preferred but not applicable (no systolic kernel for the precision) */
    }
    const int L = ch->L, W = ch->W;
    long total = 0;
    for (int i = 0; i < count; i++) total += M[i];
    ft_col *cols = malloc(sizeof(ft_col) * (size_t)(total > 0 ? total : 1));
    int ncols = 0;
    int rc = cols ? FTAMD_OK : FTAMD_ENOMEM;
    for (int i = 0; rc == FTAMD_OK && i < count; i++) {
        ft_col *one = NULL;
        int k = 0;
        rc = matrix_columns(kind, layout, plans[i]->n, M[i], dA[i], &one, &k);
        if (rc == FTAMD_OK) memcpy(cols + ncols, one, sizeof(ft_col) * (size_t)k);
        ncols += k;
        free(one);
    }
    ft_schedule *S = NULL;
    if (rc == FTAMD_OK) {
        S = calloc(1, sizeof(ft_schedule));
        rc = S ? FTAMD_OK : FTAMD_ENOMEM;
    }
    if (rc == FTAMD_OK) {
        S->layout = layout;
        S->dir = direction;
        S->precision = precision;
        S->L = L;
        S->W = W;
        S->wg_per_cu = ch->wg_per_cu;
        S->batch = 1;
        rc = build_from_columns(S, kind, direction, L, W, cols, ncols, ftamd_plan_table(big, precision), big->n);
    }
    free(cols);
    if (rc != FTAMD_OK) {
        if (S) ftamd_schedule_free_device(S);
        free(S);
        return rc;
    }
    *out = S;
    return FTAMD_OK;
}

int ft_gpu_execute_batch(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                         void *const *dA, const int *M, int count, void *stream) {
    ft_rotation_plan *big = NULL;
    int geo, layout;
    int rc = batch_check(plans, kind, direction, precision, dA, M, count, &geo, &layout, &big);
    if (rc != FTAMD_OK || count == 0) return rc;
    ftamd_choice ch;
    ftamd_choose(geo, precision, direction, 1, big->n, &ch);
    /* the cache key holds every choice the schedule was built with */
    const int L = ch.L, W = ch.W, engine = ch.engine * 2 + ch.forced, part = ch.maxpart * 16 + ch.wg_per_cu;
    pthread_mutex_lock(&g_batch.lock);
    ft_schedule *S = g_batch.sched;
    if (!batch_cache_hit(plans, kind, direction, precision, dA, M, count, L, W, engine, part)) {
        batch_cache_clear_locked();
        S = NULL;
        rc = batch_build(plans, geo, layout, direction, precision, dA, M, count, big, &ch, &S);
        if (rc == FTAMD_OK)
            rc = batch_cache_store(S, plans, kind, direction, precision, dA, M, count, L, W, engine, part);
        if (rc != FTAMD_OK) {
            pthread_mutex_unlock(&g_batch.lock);
            return rc;
        }
    }
    t_last_engine = S->sys ? FT_GPU_ENGINE_SYSTOLIC : FT_GPU_ENGINE_PIPELINE;
    rc = launch(S, precision, geo, direction, NULL, stream);
    if (rc == FTAMD_OK) pending_add(S, stream);
    pthread_mutex_unlock(&g_batch.lock);
    return rc;
}

/* ------------------------------------------------------- batch plans -- */

struct ft_gpu_batch_s {
    ft_schedule *S;
    int geo, dir, precision, count; /* geo: the table geometry of its kind */
};

ft_gpu_batch *ft_gpu_plan_batch(const ft_rotation_plan *const *plans, int kind, int direction, int precision,
                                void *const *dA, const int *M, int count) {
    ft_rotation_plan *big = NULL;
    int geo = -1, layout = 0;
    int rc = batch_check(plans, kind, direction, precision, dA, M, count, &geo, &layout, &big);
    if (rc == FTAMD_OK && count == 0) rc = FTAMD_EINVAL;
    ft_gpu_batch *B = NULL;
    if (rc == FTAMD_OK) {
        B = calloc(1, sizeof(ft_gpu_batch));
        rc = B ? FTAMD_OK : FTAMD_ENOMEM;
    }
    if (rc == FTAMD_OK) {
        ftamd_choice ch;
        ftamd_choose(geo, precision, direction, 1, big->n, &ch);
        rc = batch_build(plans, geo, layout, direction, precision, dA, M, count, big, &ch, &B->S);
    }
    if (rc != FTAMD_OK) {
        free(B);
        ftamd_set_error(rc);
        return NULL;
    }
    B->geo = geo;
    B->dir = direction;
    B->precision = precision;
    B->count = count;
    return B;
}

int ft_gpu_execute_batch_plan(const ft_gpu_batch *B, void *stream) {
    if (!B) return FT_GPU_EINVAL;
    int rc = launch(B->S, B->precision, B->geo, B->dir, NULL, stream);
    if (rc == FTAMD_OK) pending_add(B->S, stream);
    return rc;
}

void ft_gpu_destroy_batch(ft_gpu_batch *B) {
    if (!B) return;
    ftamd_hip_sync(NULL);
    ftamd_schedule_release(B->S);
    free(B);
}
