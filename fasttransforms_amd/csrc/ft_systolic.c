/*
 * ft_systolic.c -- host side of the band-systolic engine (ft_systolic.hpp):
 * the classes of sequences of a plan, their bands and step-major tables
 * (built once per plan, direction and precision), and the chains of one
 * n x M matrix.
 *
 * The columns and sequences are those of the reference's executors
 * (src/drivers.c:58-110, 314-326, 414-430) and kernels
 * (src/rotations/rotations_default.c:31-71): a column of order m takes
 * sequences j = m-2, m-4, .. m%2 (sph, disk) or m-1, .. 0 (tri), rotating
 * rows (k, k+2) (sph) or (k, k+1) (tri, disk) for k = top .. 0 with the
 * plan's s(k,j), c(k,j) (packed index rotations.c:113-114, 170-171); lo2hi
 * takes the same sequences in reverse with the transposed rotation.
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ft_internal.h"

static long idx_tri(int n, int l, int m) { return l + (long)m * (2 * n + 1 - m) / 2; }
static long idx_disk(int n, int l, int m) { return l + (long)m * n - (long)(m / 2) * ((m + 1) / 2); }

/* One class: sequences k = 0 .. K-1 of j = jtop - jstep*k (hi2lo order). */
typedef struct {
    int jtop, jstep, K;
    int rpar, rstep;   /* class row e is matrix row rstep*e + rpar */
    int jpar;          /* order parity of its columns (sph, disk); -1 for tri */
} sys_class;

static int class_top(int kind, int n, const sys_class *cl, int j) {
    if (kind == FTAMD_SPH) {
        int x = n - 3 - j - cl->rpar;
        return x >= 0 ? x / 2 : -1;
    }
    if (kind == FTAMD_TRI) return n - 2 - j;
    return n - 2 - (j + 1) / 2;
}

static long class_index(int kind, int n, const sys_class *cl, int e, int j) {
    int row = cl->rstep * e + cl->rpar;
    return kind == FTAMD_DISK ? idx_disk(n, row, j) : idx_tri(n, row, j);
}

static int make_classes(int kind, int n, sys_class *cl) {
    int nc = 0;
    if (kind == FTAMD_TRI) {
        if (n >= 2) {
            sys_class c = {n - 2, 1, n - 1, 0, 1, -1};
            cl[nc++] = c;
        }
        return nc;
    }
    int jmax = kind == FTAMD_SPH ? n - 3 : 2 * n - 4;
    for (int jpar = 0; jpar < 2; jpar++) {
        int jtop = jmax - ((jmax - jpar) % 2 + 2) % 2;
        if (jtop < jpar) continue;
        int K = (jtop - jpar) / 2 + 1;
        for (int rpar = 0; rpar < (kind == FTAMD_SPH ? 2 : 1); rpar++) {
            sys_class c = {jtop, 2, K, rpar, kind == FTAMD_SPH ? 2 : 1, jpar};
            cl[nc++] = c;
        }
    }
    return nc;
}

static int round_up(int x, int u) { return (x + u - 1) / u * u; }

/* Step at which stream position 0 reaches sequence q = S*i + s of a band as
 * its fresh row: within a lane a sequence takes the row the sequence before
 * it emits in the same step, from lane to lane a row moves one step later
 * (DPP of the previous step's registers), so lane i adds S+1 steps. */
static int sys_delay(int S, int q) { return (S + 1) * (q / S) + q % S; }

/* Tables of every band of every class for one direction and precision. */
struct ft_sys_tables {
    int S, dir, precision, ncls;
    sys_class cls[4];
    int band0[4], nb[4];     /* bands of class c: band0[c] .. band0[c]+nb[c]-1 */
    FtamdSysBand *bands;     /* host copy */
    int nbands;
    void *d_bands;
    void *d_tab;
    size_t tab_bytes;
};

int ftamd_sys_tables_S(const ft_sys_tables *T) { return T->S; }
const void *ftamd_sys_tables_dev_bands(const ft_sys_tables *T) { return T->d_bands; }
const void *ftamd_sys_tables_dev_tab(const ft_sys_tables *T) { return T->d_tab; }
const FtamdSysBand *ftamd_sys_tables_host_bands(const ft_sys_tables *T) { return T->bands; }

void ftamd_sys_tables_free(ft_sys_tables *T) {
    if (!T) return;
    ftamd_hip_free(T->d_bands);
    ftamd_hip_free(T->d_tab);
    free(T->bands);
    free(T);
}

/* Sequences per band lane of the plan's tables: 2, or 4 where the build has
 * no S = 2 engine for the precision (0 if neither). */
int ftamd_sys_choose_S(int kind, int n, int precision) {
    (void)kind;
    (void)n;
    if (ftamd_hip_sys_supported(2, precision)) return 2;
    if (ftamd_hip_sys_supported(4, precision)) return 4;
    return 0;
}

/* Bands of every class and their tables, on the host (T->bands, *host). */
static int tables_host(const ft_rotation_plan *RP, int dir, int precision, int S, ft_sys_tables *T, void **host,
                       long long *nentries) {
    const int kind = RP->kind, n = RP->n, B = FTAMD_SYS_LANES * S, U = FTAMD_SYS_UNROLL, PF = FTAMD_SYS_PREFETCH;
    *host = NULL;
    T->S = S;
    T->dir = dir;
    T->precision = precision;
    T->ncls = make_classes(kind, n, T->cls);
    long long entries = 0;
    for (int c = 0; c < T->ncls; c++) {
        T->band0[c] = T->nbands;
        T->nb[c] = (T->cls[c].K + B - 1) / B;
        T->nbands += T->nb[c];
    }
    T->bands = calloc((size_t)(T->nbands > 0 ? T->nbands : 1), sizeof(FtamdSysBand));
    if (!T->bands) return FTAMD_ENOMEM;
    for (int c = 0; c < T->ncls; c++) {
        const sys_class *cl = &T->cls[c];
        for (int b = 0; b < T->nb[c]; b++) {
            FtamdSysBand *bd = &T->bands[T->band0[c] + b];
            int g0 = B * b, nseq = cl->K - g0 < B ? cl->K - g0 : B;
            /* the stream's top row: one above the longest sequence's top */
            int klong = dir == 0 ? g0 + nseq - 1 : cl->K - 1 - g0;
            bd->R = class_top(kind, n, cl, cl->jtop - cl->jstep * klong) + 1;
            if (bd->R < 0) bd->R = 0;
            /* the last sequence emits position R at step R + delay + 1 */
            bd->steps = round_up(bd->R + sys_delay(S, B - 1) + 2, U);
            bd->gbase = g0;
            bd->nseq = nseq;
            bd->tab = entries;
            entries += (long long)(bd->steps + PF) * FTAMD_SYS_LANES * S;
        }
    }
    size_t esz = precision == 0 ? sizeof(double) : sizeof(float);
    void *h = malloc((size_t)(entries > 0 ? entries : 1) * 2 * esz);
    if (!h) return FTAMD_ENOMEM;
    for (int c = 0; c < T->ncls; c++) {
        const sys_class *cl = &T->cls[c];
        for (int b = 0; b < T->nb[c]; b++) {
            const FtamdSysBand *bd = &T->bands[T->band0[c] + b];
            for (int t = 0; t < bd->steps + PF; t++)
                for (int q = 0; q < B; q++) {
                    /* lane i holds sequences q = S*i .. S*i+S-1: entry (t, i, s),
                     * stored slot-major in FP64 */
                    int i_lane = q / S, s_slot = q % S;
                    long long at = bd->tab + (long long)t * B +
                                   (FTAMD_SYS_SLOT_MAJOR(precision) ? s_slot * FTAMD_SYS_LANES + i_lane : q);
                    double sv = 0.0, cv = 1.0;
                    if (q < bd->nseq) {
                        int g = bd->gbase + q;
                        int k = dir == 0 ? g : cl->K - 1 - g;
                        int j = cl->jtop - cl->jstep * k;
                        /* class row whose rotation sequence q does at step t */
                        int d = sys_delay(S, q);
                        int e = dir == 0 ? bd->R + d - t : t - d - 1;
                        if (e >= 0 && e <= class_top(kind, n, cl, j)) {
                            long ix = class_index(kind, n, cl, e, j);
                            sv = RP->s[ix];
                            cv = RP->c[ix];
                        }
                    }
                    if (precision == 0) {
                        ((double *)h)[2 * at] = sv;
                        ((double *)h)[2 * at + 1] = cv;
                    } else {
                        ((float *)h)[2 * at] = (float)sv;
                        ((float *)h)[2 * at + 1] = (float)cv;
                    }
                }
        }
    }
    *host = h;
    *nentries = entries;
    return FTAMD_OK;
}

int ftamd_sys_tables_build(const ft_rotation_plan *RP, int dir, int precision, int S, ft_sys_tables **out) {
    *out = NULL;
    ft_sys_tables *T = calloc(1, sizeof(ft_sys_tables));
    if (!T) return FTAMD_ENOMEM;
    void *host = NULL;
    long long entries = 0;
    int rc = tables_host(RP, dir, precision, S, T, &host, &entries);
    size_t esz = precision == 0 ? sizeof(double) : sizeof(float);
    T->tab_bytes = (size_t)entries * 2 * esz;
    /* the kernel addresses the tables with 32-bit byte offsets */
    if (rc == FTAMD_OK && T->tab_bytes >= 0x7fffffffULL) rc = FTAMD_EINVAL;
    if (rc == FTAMD_OK) rc = ftamd_hip_malloc(&T->d_tab, (size_t)(entries > 0 ? entries : 1) * 2 * esz);
    if (rc == FTAMD_OK)
        rc = ftamd_hip_malloc(&T->d_bands, sizeof(FtamdSysBand) * (size_t)(T->nbands > 0 ? T->nbands : 1));
    if (rc == FTAMD_OK && entries > 0) rc = ftamd_hip_h2d(T->d_tab, host, (size_t)entries * 2 * esz, NULL);
    if (rc == FTAMD_OK && T->nbands > 0)
        rc = ftamd_hip_h2d(T->d_bands, T->bands, sizeof(FtamdSysBand) * (size_t)T->nbands, NULL);
    if (rc == FTAMD_OK) rc = ftamd_hip_sync(NULL);
    free(host);
    if (rc != FTAMD_OK) {
        ftamd_sys_tables_free(T);
        return rc;
    }
    *out = T;
    return FTAMD_OK;
}

static int cmp_nb_desc(const void *x, const void *y) {
    const FtamdSysChain *a = x, *b = y;
    if (a->nb != b->nb) return b->nb - a->nb;
    return a->band0 - b->band0;
}

/* Longest chains first, in place and without allocating (Shell sort). */
static void sort_chains(FtamdSysChain *c, int n) {
    static const int gaps[] = {701, 301, 132, 57, 23, 10, 4, 1};
    for (int g = 0; g < 8; g++) {
        int gap = gaps[g];
        for (int i = gap; i < n; i++) {
            FtamdSysChain v = c[i];
            int j = i;
            while (j >= gap && cmp_nb_desc(&c[j - gap], &v) > 0) {
                c[j] = c[j - gap];
                j -= gap;
            }
            c[j] = v;
        }
    }
}

/* The chain record of class c over up to FTAMD_SYS_COLS columns: column
 * col[q] of stride lda from aoff (-1: none) applies the class's sequences from
 * j = jfirst[q] down (hi2lo order; lo2hi up to it), the class's row e at
 * element rmul * (rstep*e + rpar) of the column.  0 if no column starts at a
 * sequence of the class. */
static int chain_record(const ft_sys_tables *T, int c, const int *col, const int *jfirst, int lda, int rmul,
                        long long aoff, FtamdSysChain *out) {
    const int B = FTAMD_SYS_LANES * T->S;
    const sys_class *cl = &T->cls[c];
    FtamdSysChain x;
    memset(&x, 0, sizeof(x));
    int kmin = INT_MAX, any = 0;
    for (int q = 0; q < FTAMD_SYS_COLS; q++) {
        const int ok = col[q] >= 0 && jfirst[q] <= cl->jtop && jfirst[q] >= cl->jtop - cl->jstep * (cl->K - 1);
        x.col[q] = ok ? col[q] : -1;
        x.glo[q] = INT_MIN;
        x.ghi[q] = INT_MAX;
        if (!ok) continue;
        int ks = (cl->jtop - jfirst[q]) / cl->jstep;
        if (T->dir == 0) x.glo[q] = ks;
        else x.ghi[q] = cl->K - 1 - ks;
        if (ks < kmin) kmin = ks;
        any = 1;
    }
    if (!any) return 0;
    int b_first, b_last;
    if (T->dir == 0) {
        b_first = kmin / B;
        b_last = T->nb[c] - 1;
    } else {
        b_first = 0;
        b_last = (cl->K - 1 - kmin) / B;
    }
    x.band0 = T->band0[c] + b_first;
    x.nb = b_last - b_first + 1;
    x.dir = T->dir;
    x.rstep = rmul * cl->rstep;
    x.rpar = rmul * cl->rpar;
    x.lda = lda;
    x.aoff = aoff;
    x.split_in = x.split_out = -1;
    *out = x;
    return 1;
}

/* Whole chains of an n x M matrix: per class, the columns in groups of 8
 * (four orders of the class's parity x their two columns for sph, sphv and
 * disk, eight orders of the parity for the rectangle-disk, eight consecutive
 * orders for tri), each with the bands its columns need, written to buf (at
 * most cap records, longest first) without allocating.  `layout` is the
 * column layout (FTAMD_LAYOUT_*) of a sphere-geometry matrix.  The records
 * are not split yet (ftamd_sys_split).  FTAMD_ENOMEM if more than cap chains
 * are needed (*count then holds the number needed). */
int ftamd_sys_chains_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int M, int layout, FtamdSysChain *buf,
                          int cap, int *count) {
    const int kind = RP->kind;
    /* orders per chain step and the highest order of the layout
     * (drivers.c:94-110, 204-220, 314-326, 414-430, 496-508) */
    const int one_col = kind == FTAMD_TRI || layout == FTAMD_LAYOUT_RECTDISK;
    const int mtop = kind == FTAMD_TRI || layout == FTAMD_LAYOUT_RECTDISK ? M - 1
                     : layout == FTAMD_LAYOUT_SPHV                        ? M / 2 - 1
                                                                          : (M - 1) / 2;
    int nch = 0;
    for (int c = 0; c < T->ncls; c++) {
        const sys_class *cl = &T->cls[c];
        /* orders of the class, highest first */
        int mhi = mtop;
        if (kind != FTAMD_TRI && mhi % 2 != cl->jpar) mhi--;
        int mlo = kind == FTAMD_TRI ? 1 : 2;
        for (int m0 = mhi; m0 >= mlo; m0 -= kind == FTAMD_TRI ? 8 : 8 * (one_col ? 2 : 1)) {
            int cols[FTAMD_SYS_COLS], jfirst[FTAMD_SYS_COLS];
            for (int q = 0; q < FTAMD_SYS_COLS; q++) {
                int m, col;
                if (kind == FTAMD_TRI) {
                    m = m0 - q;
                    col = m;
                } else if (one_col) {
                    m = m0 - 2 * q;
                    col = m;
                } else {
                    m = m0 - 2 * (q / 2);
                    col = (layout == FTAMD_LAYOUT_SPHV ? 2 * m + 1 : 2 * m - 1) + (q % 2);
                }
                jfirst[q] = kind == FTAMD_TRI ? m - 1 : m - 2;
                cols[q] = m >= mlo && col < M ? col : -1;
            }
            FtamdSysChain x;
            if (!chain_record(T, c, cols, jfirst, RP->n, 1, 0, &x)) continue;
            if (nch < cap) buf[nch] = x;
            nch++;
        }
    }
    *count = nch;
    if (nch > cap) return FTAMD_ENOMEM;
    sort_chains(buf, nch);
    return FTAMD_OK;
}

/* Whole chains of one of the tetrahedron's two rotation stages (drivers.c:
 * 510-533) on `count` consecutive slabs of an N x L x M array, slab s (order
 * m = m0 + s) being the N x L block at element N*L*s, into buf as
 * ftamd_sys_chains_fill does:
 *   part 0, the triangle's rotations on the slab columns (RP: the triangle
 *   plan of size N): column l of slab m is a triangle column of order l + m
 *   (rotations.c:163 on A + N*(l + L*m)); chains of eight consecutive columns
 *   of one slab;
 *   part 1, the block rotations (rotations.c:297-321; RP: the triangle plan
 *   of size L, whose (s, c) are those of the size-N plan for every (l, j)
 *   with l + j <= L - 2): row k of slab m is a vector of L entries at stride
 *   N rotated as a triangle column of order m; chains of eight consecutive
 *   rows. */
int ftamd_sys_chains_tet_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int part, int N, int L, int m0,
                              int count, FtamdSysChain *buf, int cap, int *nout) {
    int nch = 0;
    for (int c = 0; c < T->ncls; c++) {
        const int jtop = T->cls[c].jtop;
        for (int sl = 0; sl < count; sl++) {
            const int m = m0 + sl;
            const long long aoff = (long long)N * L * sl;
            int cols[FTAMD_SYS_COLS], jfirst[FTAMD_SYS_COLS];
            FtamdSysChain x;
            if (part == 0) {
                for (int l0 = L - m - 1; l0 >= 0; l0 -= FTAMD_SYS_COLS) {
                    for (int q = 0; q < FTAMD_SYS_COLS; q++) {
                        cols[q] = l0 - q;
                        jfirst[q] = l0 - q + m - 1;
                    }
                    if (!chain_record(T, c, cols, jfirst, N, 1, aoff, &x)) continue;
                    if (nch < cap) buf[nch] = x;
                    nch++;
                }
            } else {
                const int jf = m - 1 < jtop ? m - 1 : jtop;
                for (int k0 = 0; k0 < N; k0 += FTAMD_SYS_COLS) {
                    for (int q = 0; q < FTAMD_SYS_COLS; q++) {
                        cols[q] = k0 + q < N ? k0 + q : -1;
                        jfirst[q] = jf;
                    }
                    if (!chain_record(T, c, cols, jfirst, 1, N, aoff, &x)) continue;
                    if (nch < cap) buf[nch] = x;
                    nch++;
                }
            }
        }
    }
    *nout = nch;
    if (nch > cap) return FTAMD_ENOMEM;
    sort_chains(buf, nch);
    return FTAMD_OK;
}

/* Parts of a chain of nb bands with parts of at most maxpart bands. */
static int split_parts(int nb, int maxpart) { return (nb + maxpart - 1) / maxpart; }

/* Splits whole chains (buf[0 .. nfull-1], longest first) in place into
 * parts of at most maxpart (<= FTAMD_SYS_WAVES) bands, as equal as
 * possible: the parts of a chain stay together, in chain order, linked by
 * progress counters 0 .. *nprog-1.  Sets each part's issue priority from its
 * whole chain's length.  FTAMD_ENOMEM if the parts need more than cap records
 * (*count then holds the number needed). */
int ftamd_sys_split(FtamdSysChain *buf, int nfull, int cap, int maxpart, int *count, int *nprog) {
    if (maxpart < 1) maxpart = 1;
    if (maxpart > FTAMD_SYS_WAVES) maxpart = FTAMD_SYS_WAVES;
    int total = 0, links = 0;
    for (int c = 0; c < nfull; c++) {
        int np = split_parts(buf[c].nb, maxpart);
        total += np;
        links += np - 1;
    }
    *count = total;
    *nprog = links;
    if (total > cap) return FTAMD_ENOMEM;
    /* from the back, so that no record is overwritten before it is read */
    int at = total, link = links;
    for (int c = nfull - 1; c >= 0; c--) {
        const FtamdSysChain x = buf[c];
        const int np = split_parts(x.nb, maxpart);
        const int prio = x.nb >= 12 ? 3 : x.nb >= 8 ? 2 : x.nb >= 4 ? 1 : 0;
        at -= np;
        link -= np - 1;
        int b = 0;
        for (int k = 0; k < np; k++) {
            FtamdSysChain y = x;
            int nbk = x.nb / np + (k < x.nb % np ? 1 : 0);
            y.band0 = x.band0 + b;
            y.nb = nbk;
            y.prio = prio;
            y.split_in = k > 0 ? link + k - 1 : -1;
            y.split_out = k < np - 1 ? link + k : -1;
            buf[at + k] = y;
            b += nbk;
        }
    }
    return FTAMD_OK;
}

/* Deals chain parts to slots: nxcd XCDs of nsx slots each (slot s of XCD x
 * is x*nsx + s; the kernel runs an XCD's slots in order, each in a
 * workgroup of that XCD).  Chain by chain (longest first): the XCD with the
 * least work so far, then part by part the slot with the least work among
 * those after the previous part's (a part waits only on parts in earlier
 * slots of its XCD), then the wave offset that keeps that slot's busiest wave
 * least busy (its waves stream their bands of successive parts one after
 * another).  Work = steps of all bands (the longest-processing-time rule).
 * Writes items[nch] grouped by slot, slot_items[nxcd*nsx+1] and
 * xcd_slot[nxcd+1]; ws holds nxcd*nsx*(1+FTAMD_SYS_WAVES) + nxcd longs and
 * 2*nch ints of workspace.  FTAMD_EINVAL if a chain has more parts than an
 * XCD has slots.  No allocation. */
int ftamd_sys_assign(const FtamdSysChain *ch, int nch, const FtamdSysBand *bands, int nxcd, int nsx,
                     FtamdSysItem *items, int *slot_items, int *xcd_slot, long *ws_load, int *ws_chain) {
    const int nslots = nxcd * nsx;
    long *slot_load = ws_load, *wave_load = ws_load + nslots, *xcd_load = wave_load + (long)nslots * FTAMD_SYS_WAVES;
    int *chain_slot = ws_chain, *chain_w0 = ws_chain + nch;
    for (long w = 0; w < (long)nslots * (1 + FTAMD_SYS_WAVES) + nxcd; w++) ws_load[w] = 0;
    for (int c = 0; c < nch;) {
        /* the parts of one chain: c .. e-1 */
        int e = c + 1;
        while (e < nch && ch[e].split_in >= 0 && ch[e].split_in == ch[e - 1].split_out) e++;
        const int np = e - c;
        if (np > nsx) return FTAMD_EINVAL;
        long chain_work = 0;
        for (int p = c; p < e; p++)
            for (int k = 0; k < ch[p].nb; k++) chain_work += bands[ch[p].band0 + k].steps;
        int x = 0;
        for (int y = 1; y < nxcd; y++)
            if (xcd_load[y] < xcd_load[x]) x = y;
        xcd_load[x] += chain_work;
        int lo = 0;
        for (int p = c; p < e; p++) {
            const int nb = ch[p].nb;
            const FtamdSysBand *bd = bands + ch[p].band0;
            long work = 0;
            for (int k = 0; k < nb; k++) work += bd[k].steps;
            /* leave a later slot for each part still to come */
            const int hi = nsx - (e - 1 - p);
            int best = lo;
            for (int s = lo + 1; s < hi; s++)
                if (slot_load[x * nsx + s] < slot_load[x * nsx + best]) best = s;
            const int slot = x * nsx + best;
            long *wl = wave_load + (long)slot * FTAMD_SYS_WAVES;
            int o_best = FTAMD_SYS_WAVES - nb;
            long m_best = -1;
            for (int o = FTAMD_SYS_WAVES - nb; o >= 0; o--) {
                long m = 0;
                for (int k = 0; k < nb; k++)
                    if (wl[o + k] + bd[k].steps > m) m = wl[o + k] + bd[k].steps;
                if (m_best < 0 || m < m_best) {
                    m_best = m;
                    o_best = o;
                }
            }
            for (int k = 0; k < nb; k++) wl[o_best + k] += bd[k].steps;
            slot_load[slot] += work;
            chain_slot[p] = slot;
            chain_w0[p] = o_best;
            lo = best + 1;
        }
        c = e;
    }
    for (int s = 0; s <= nslots; s++) slot_items[s] = 0;
    for (int c = 0; c < nch; c++) slot_items[chain_slot[c] + 1]++;
    for (int s = 0; s < nslots; s++) slot_items[s + 1] += slot_items[s];
    for (int s = 0; s < nslots; s++) ws_load[s] = slot_items[s]; /* next free item of each slot */
    for (int c = 0; c < nch; c++) {
        long at = ws_load[chain_slot[c]]++;
        items[at].chain = c;
        items[at].wave0 = chain_w0[c];
    }
    for (int y = 0; y <= nxcd; y++) xcd_slot[y] = y * nsx;
    return FTAMD_OK;
}

/* Whole chains in a malloc'd array the caller frees. */
static int chains_full(const ft_rotation_plan *RP, const ft_sys_tables *T, int M, int layout, FtamdSysChain **out,
                       int *count) {
    *out = NULL;
    *count = 0;
    int need = 0;
    ftamd_sys_chains_fill(RP, T, M, layout, NULL, 0, &need);
    FtamdSysChain *ch = malloc(sizeof(FtamdSysChain) * (size_t)(need > 0 ? need : 1));
    if (!ch) return FTAMD_ENOMEM;
    int rc = ftamd_sys_chains_fill(RP, T, M, layout, ch, need, count);
    if (rc != FTAMD_OK) {
        free(ch);
        *count = 0;
        return rc;
    }
    *out = ch;
    return FTAMD_OK;
}

/* Splits the malloc'd whole chains *ch (n of them) into parts, growing the
 * array as needed. */
static int split_alloc(FtamdSysChain **ch, int n, int maxpart, int *count, int *nprog) {
    int need = 0;
    if (ftamd_sys_split(*ch, n, 0, maxpart, &need, nprog) == FTAMD_OK && need == 0) {
        *count = 0;
        return FTAMD_OK;
    }
    FtamdSysChain *grown = realloc(*ch, sizeof(FtamdSysChain) * (size_t)(need > n ? need : n));
    if (!grown) return FTAMD_ENOMEM;
    *ch = grown;
    return ftamd_sys_split(*ch, n, need, maxpart, count, nprog);
}

/* The chain parts of an n x M matrix (longest chains first, parts of at most
 * maxpart bands) in a malloc'd array the caller frees; *nprog progress
 * counters link them. */
int ftamd_sys_chains(const ft_rotation_plan *RP, const ft_sys_tables *T, int M, int layout, int maxpart,
                     FtamdSysChain **out, int *count, int *nprog) {
    int n = 0;
    *nprog = 0;
    int rc = chains_full(RP, T, M, layout, out, &n);
    if (rc == FTAMD_OK) rc = split_alloc(out, n, maxpart, count, nprog);
    if (rc != FTAMD_OK) {
        free(*out);
        *out = NULL;
        *count = 0;
    }
    return rc;
}

/* Whole chains of the spin-weighted sphere's second stage (rotations_default.c:
 * 99-101, 105-107) on the n x M complex matrix A (interleaved re, im; column
 * c of order m = 0, -1, 1, -2, 2, ... as drivers.c:687-705 lays them out) for
 * spin s, on the sphere plan RP of size n, into buf as ftamd_sys_chains_fill
 * does: complex column c takes the sphere's sequences of order d = |am - as|
 * (j = d-2, d-4, ... >= d%2, capped at the class's highest); its real and
 * imaginary parts are the two columns of a lane row (doubles 2n*c and 2n*c+1,
 * rows at stride 2), four complex columns of one order parity per chain,
 * highest orders first. */
int ftamd_sys_chains_spin_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int s, int M, FtamdSysChain *buf,
                               int cap, int *nout) {
    const int n = RP->n, as = s < 0 ? -s : s;
    int nch = 0;
    for (int c = 0; c < T->ncls; c++) {
        const sys_class *cl = &T->cls[c];
        /* columns of the class's order parity, by d descending: d's of one
         * column are at most max(am, as) <= M/2 + as, visited top down */
        const int dmax = M / 2 + as;
        int cols[FTAMD_SYS_COLS], jfirst[FTAMD_SYS_COLS], q = 0;
        for (int d = dmax - ((dmax - cl->jpar) % 2 + 2) % 2; d >= 2; d -= 2) {
            /* the orders with |am - as| = d: am = as + d and am = as - d, each as -am and +am */
            int ams[2] = {as + d, as - d}, na = as - d >= 0 && d != 0 ? 2 : 1;
            for (int a = 0; a < na; a++) {
                const int am = ams[a];
                for (int sg = -1; sg <= 1; sg += 2) {
                    const int col = am == 0 ? (sg < 0 ? 0 : -1) : (sg < 0 ? 2 * am - 1 : 2 * am);
                    if (col < 0 || col >= M) continue;
                    const int jf = d - 2 < cl->jtop ? d - 2 : cl->jtop;
                    cols[q] = 2 * n * col;
                    cols[q + 1] = 2 * n * col + 1;
                    jfirst[q] = jfirst[q + 1] = jf;
                    q += 2;
                    if (q == FTAMD_SYS_COLS) {
                        FtamdSysChain x;
                        if (chain_record(T, c, cols, jfirst, 1, 2, 0, &x)) {
                            if (nch < cap) buf[nch] = x;
                            nch++;
                        }
                        q = 0;
                    }
                }
            }
        }
        if (q > 0) {
            for (; q < FTAMD_SYS_COLS; q++) cols[q] = -1, jfirst[q] = 0;
            FtamdSysChain x;
            if (chain_record(T, c, cols, jfirst, 1, 2, 0, &x)) {
                if (nch < cap) buf[nch] = x;
                nch++;
            }
        }
    }
    *nout = nch;
    if (nch > cap) return FTAMD_ENOMEM;
    sort_chains(buf, nch);
    return FTAMD_OK;
}

/* The chain parts of the spin-weighted sphere's second stage in a malloc'd
 * array the caller frees, as ftamd_sys_chains. */
int ftamd_sys_chains_spin(const ft_rotation_plan *RP, const ft_sys_tables *T, int s, int M, int maxpart,
                          FtamdSysChain **out, int *nout, int *nprog) {
    *out = NULL;
    *nout = 0;
    *nprog = 0;
    int need = 0, k = 0;
    ftamd_sys_chains_spin_fill(RP, T, s, M, NULL, 0, &need);
    FtamdSysChain *ch = malloc(sizeof(FtamdSysChain) * (size_t)(need > 0 ? need : 1));
    if (!ch) return FTAMD_ENOMEM;
    int rc = ftamd_sys_chains_spin_fill(RP, T, s, M, ch, need, &k);
    if (rc == FTAMD_OK) rc = split_alloc(&ch, k, maxpart, nout, nprog);
    if (rc != FTAMD_OK) {
        free(ch);
        *nout = 0;
        return rc;
    }
    *out = ch;
    return FTAMD_OK;
}

/* The chain parts of one of the tetrahedron's stages (ftamd_sys_chains_tet_fill)
 * in a malloc'd array the caller frees, as ftamd_sys_chains. */
int ftamd_sys_chains_tet(const ft_rotation_plan *RP, const ft_sys_tables *T, int part, int N, int L, int m0, int count,
                         int maxpart, FtamdSysChain **out, int *nout, int *nprog) {
    *out = NULL;
    *nout = 0;
    *nprog = 0;
    int need = 0, n = 0;
    ftamd_sys_chains_tet_fill(RP, T, part, N, L, m0, count, NULL, 0, &need);
    FtamdSysChain *ch = malloc(sizeof(FtamdSysChain) * (size_t)(need > 0 ? need : 1));
    if (!ch) return FTAMD_ENOMEM;
    int rc = ftamd_sys_chains_tet_fill(RP, T, part, N, L, m0, count, ch, need, &n);
    if (rc == FTAMD_OK) rc = split_alloc(&ch, n, maxpart, nout, nprog);
    if (rc != FTAMD_OK) {
        free(ch);
        *nout = 0;
        return rc;
    }
    *out = ch;
    return FTAMD_OK;
}

/* Bands and chains of a batch of transforms of one kind, direction,
 * precision and S (every plan's tables built): the bands of each distinct
 * plan once, their table offsets taken from the first plan's tables, the
 * chains of every transform with its bands, matrix offset from dA[0] and
 * column stride, longest first.  FTAMD_EINVAL if a chain does not fit. */
int ftamd_sys_batch(const ft_rotation_plan *const *plans, ft_sys_tables *const *tabs, void *const *dA,
                    const int *M, int count, int layout, int maxpart, FtamdSysBand **bands_out, int *nbands_out,
                    FtamdSysChain **chains_out, int *nchains_out, int *nprog_out, const void **tab_base) {
    *bands_out = NULL;
    *chains_out = NULL;
    *nbands_out = *nchains_out = *nprog_out = 0;
    *tab_base = count > 0 ? tabs[0]->d_tab : NULL;
    if (count <= 0) return FTAMD_OK;
    const long long pair = tabs[0]->precision == 0 ? 2 * (long long)sizeof(double) : 2 * (long long)sizeof(float);
    const long long esz = pair / 2;
    /* distinct tables and the first band of each in the merged array */
    const ft_sys_tables **seen = malloc(sizeof(*seen) * (size_t)count);
    int *first = malloc(sizeof(int) * (size_t)count);
    int nseen = 0, nb = 0, rc = (seen && first) ? FTAMD_OK : FTAMD_ENOMEM;
    for (int i = 0; rc == FTAMD_OK && i < count; i++) {
        int k = 0;
        while (k < nseen && seen[k] != tabs[i]) k++;
        if (k == nseen) {
            seen[nseen] = tabs[i];
            first[nseen++] = nb;
            nb += tabs[i]->nbands;
        }
    }
    FtamdSysBand *bands = rc == FTAMD_OK ? malloc(sizeof(FtamdSysBand) * (size_t)(nb > 0 ? nb : 1)) : NULL;
    if (rc == FTAMD_OK && !bands) rc = FTAMD_ENOMEM;
    for (int k = 0; rc == FTAMD_OK && k < nseen; k++) {
        /* tables are separate allocations (aligned far beyond one pair) */
        const long long shift = ((long long)(uintptr_t)seen[k]->d_tab - (long long)(uintptr_t)*tab_base) / pair;
        for (int b = 0; b < seen[k]->nbands; b++) {
            bands[first[k] + b] = seen[k]->bands[b];
            bands[first[k] + b].tab += shift;
        }
    }
    FtamdSysChain *all = NULL;
    int nall = 0, cap = 0;
    for (int i = 0; rc == FTAMD_OK && i < count; i++) {
        int k = 0;
        while (seen[k] != tabs[i]) k++;
        FtamdSysChain *ch = NULL;
        int nch = 0;
        rc = chains_full(plans[i], tabs[i], M[i], layout, &ch, &nch);
        if (rc == FTAMD_OK && nall + nch > cap) {
            cap = 2 * (nall + nch) + 16;
            FtamdSysChain *grown = realloc(all, sizeof(FtamdSysChain) * (size_t)cap);
            if (!grown) rc = FTAMD_ENOMEM;
            else all = grown;
        }
        for (int c = 0; rc == FTAMD_OK && c < nch; c++) {
            ch[c].band0 += first[k];
            ch[c].aoff = ((long long)(uintptr_t)dA[i] - (long long)(uintptr_t)dA[0]) / esz;
            all[nall++] = ch[c];
        }
        free(ch);
    }
    free(seen);
    free(first);
    if (rc != FTAMD_OK) {
        free(bands);
        free(all);
        return rc;
    }
    sort_chains(all, nall);
    int nparts = 0, nprog = 0;
    rc = all ? split_alloc(&all, nall, maxpart, &nparts, &nprog) : FTAMD_OK;
    if (rc != FTAMD_OK) {
        free(bands);
        free(all);
        return rc;
    }
    *bands_out = bands;
    *nbands_out = nb;
    *chains_out = all;
    *nchains_out = nparts;
    *nprog_out = nprog;
    return FTAMD_OK;
}

/* Diagnostics (no device work) of the spin-weighted sphere's second stage:
 * the FP64 bands and table of the sphere plan RP and the chain parts
 * ftamd_sys_chains_spin builds, as ftamd_sys_debug_host. */
int ftamd_sys_debug_spin(const ft_rotation_plan *RP, int dir, int S, int s, int M, int maxpart, FtamdSysBand **bands,
                         int *nbands, double **tab, long long *nentries, FtamdSysChain **chains, int *nchains) {
    ft_sys_tables T;
    memset(&T, 0, sizeof(T));
    void *host = NULL;
    int rc = tables_host(RP, dir, 0, S, &T, &host, nentries);
    int nprog = 0;
    if (rc == FTAMD_OK) rc = ftamd_sys_chains_spin(RP, &T, s, M, maxpart, chains, nchains, &nprog);
    if (rc != FTAMD_OK) {
        free(T.bands);
        free(host);
        return rc;
    }
    *bands = T.bands;
    *nbands = T.nbands;
    *tab = host;
    return FTAMD_OK;
}

/* Diagnostics (no device work) of a tetrahedron stage: the FP64 bands and
 * table of RP and the chain parts ftamd_sys_chains_tet builds, as
 * ftamd_sys_debug_host. */
int ftamd_sys_debug_tet(const ft_rotation_plan *RP, int dir, int S, int part, int N, int L, int m0, int count,
                        int maxpart, FtamdSysBand **bands, int *nbands, double **tab, long long *nentries,
                        FtamdSysChain **chains, int *nchains) {
    ft_sys_tables T;
    memset(&T, 0, sizeof(T));
    void *host = NULL;
    int rc = tables_host(RP, dir, 0, S, &T, &host, nentries);
    int nprog = 0;
    if (rc == FTAMD_OK) rc = ftamd_sys_chains_tet(RP, &T, part, N, L, m0, count, maxpart, chains, nchains, &nprog);
    if (rc != FTAMD_OK) {
        free(T.bands);
        free(host);
        return rc;
    }
    *bands = T.bands;
    *nbands = T.nbands;
    *tab = host;
    return FTAMD_OK;
}

/* Diagnostics (no device work): the FP64 bands, table and chain parts (of
 * at most maxpart bands) the engine would run for an n x M matrix, in malloc'd host arrays the caller frees.
 * Lets the host-side construction be checked on a machine without a GPU. */
int ftamd_sys_debug_host(const ft_rotation_plan *RP, int dir, int S, int M, int layout, int maxpart,
                         FtamdSysBand **bands, int *nbands, double **tab, long long *nentries, FtamdSysChain **chains,
                         int *nchains) {
    ft_sys_tables T;
    memset(&T, 0, sizeof(T));
    void *host = NULL;
    int rc = tables_host(RP, dir, 0, S, &T, &host, nentries);
    int nprog = 0;
    if (rc == FTAMD_OK) rc = ftamd_sys_chains(RP, &T, M, layout, maxpart, chains, nchains, &nprog);
    if (rc != FTAMD_OK) {
        free(T.bands);
        free(host);
        return rc;
    }
    *bands = T.bands;
    *nbands = T.nbands;
    *tab = host;
    return FTAMD_OK;
}
