/*
 * ft_hip_abi.h -- the thin C ABI between the C host code (ft_plans.c,
 * ft_drivers.c) and the HIP kernels (ft_rotations.hip).  Plain C types only.
 */
#ifndef FT_HIP_ABI_H
#define FT_HIP_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FTAMD_SPH = 0, FTAMD_TRI = 1, FTAMD_DISK = 2 };
/* Column layouts of a matrix whose geometry is the sphere's: the sphere
 * (orders 2..(M-1)/2, columns 2m-1 and 2m), vector spherical harmonics
 * (orders 2..M/2-1, columns 2m+1 and 2m+2), the rectangle-disk (orders
 * 2..M-1, column m).  Triangle and disk matrices have one layout each. */
enum { FTAMD_LAYOUT_STD = 0, FTAMD_LAYOUT_SPHV = 1, FTAMD_LAYOUT_RECTDISK = 2 };
/* The tetrahedron's two stages on triangle plans: the slab columns, the slab
 * rows; the spin-weighted sphere's second stage on a sphere plan (complex
 * columns, real and imaginary parts in one lane row). */
enum { FTAMD_LAYOUT_TET = 3, FTAMD_LAYOUT_TETBLOCK = 4, FTAMD_LAYOUT_SPIN = 5 };
enum { FTAMD_OK = 0, FTAMD_EINVAL = 1, FTAMD_ENOMEM = 2, FTAMD_EDEVICE = 3, FTAMD_ETIMEOUT = 4, FTAMD_EPTR = 5,
       FTAMD_EINCOMPLETE = 6 };
enum { FTAMD_PTR_HOST = 0, FTAMD_PTR_DEVICE = 1 };
enum { FTAMD_ERR_TIMEOUT = 1, FTAMD_ERR_INCOMPLETE = 2 };

/* Largest number of XCDs (each with its own L2) a device may report. */
#define FTAMD_MAX_XCD 16

/* "No row is final yet" value of a progress slot (rows >= value are final). */
#define FTAMD_PROG_NONE 0x7fffffff

/* Spare rows on both sides of every boundary buffer (rows handed from one
 * task to the next), so that an edge tile's emission stores all of its rows. */
#define FTAMD_BND_PAD 16

/* Padding (table entries) kept before and after every device table so that
 * edge tiles may load a few entries beyond a column without leaving the
 * allocation. */
#define FTAMD_TABLE_PAD 256

/* One group: up to 64 columns of one transform that need the same rotation
 * sequences.  Sequences are numbered q = 0 .. Q-1 in application order;
 * j(q) = j_first -/+ step*q for hi2lo/lo2hi. */
typedef struct {
    void *A;          /* the transform's matrix (device) */
    int n;            /* its rows */
    int rs;           /* row stride in elements (1 for a column-major matrix) */
    int j_first;      /* j of sequence 0 */
    int Q;            /* number of sequences */
    int nstages;      /* ceil(Q / L) */
    int npass;        /* ceil(nstages / (W-2)): a task's workgroup has W-2 stage waves and 2 I/O waves */
    int jfull;        /* every column of the group needs every sequence with j <= jfull */
    int prog_base;    /* first progress slot (one per pass) */
    int rpar;         /* sphere: row parity of this sub-problem (rows rpar, rpar+2, ...); -1 otherwise */
    int prio;         /* wave issue priority 0..3 of its tasks (longest chains highest) */
    long long bnd;    /* element offset of its boundary buffers: npass-1 of [rows + 2*FTAMD_BND_PAD][64] */
} FtamdGroup;

typedef struct {
    const FtamdGroup *groups;
    const int *lane_off;        /* [groups*64] element offset of the lane's column in A, -1 if none */
    const int *lane_jmax;       /* [groups*64] largest j the lane's column needs, -1 if none */
    const int *tasks;           /* [2*ntasks] (group, pass): per XCD, producers before consumers */
    const int *xcd_off;         /* [nxcd+1] first task of each XCD's list */
    int ntasks;
    int nxcd;
    const void *sc;             /* interleaved (s, c) table, entry 0 (padding on both sides) */
    void *bnd;                  /* boundary buffers handing rows from one task to the next */
    void *A;                    /* matrix of every group when non-NULL (single-matrix launches; passed
                                   by value, so a captured launch keeps the matrix it was given) */
    int ntab;                   /* n the table was built for (index function) */
    unsigned *ticket;           /* [nxcd] zero at launch; reset by the kernel */
    unsigned *done;             /* zero at launch; reset by the kernel */
    unsigned *completed;        /* zero at launch; reset by the kernel */
    int *prog;                  /* FTAMD_PROG_NONE at launch; reset by the kernel */
    int nprog;
    int *err;                   /* FTAMD_ERR_* bits */
    unsigned long long *trace;  /* diagnostic timeline (instrumented builds only), else NULL */
} FtamdLaunch;

/* ---- band-systolic engine (ft_systolic.hpp) ----
 * Lanes are rotation sequences.  A class is the list of sequences one set of
 * columns shares (sph: one order parity and one row parity; tri: all; disk:
 * one order parity), numbered k = 0 .. K-1 in application order.  A band is
 * FTAMD_SYS_LANES * S consecutive sequences of a class; a chain is the bands
 * one group of FTAMD_SYS_COLS columns needs.  A chain runs in parts of at
 * most FTAMD_SYS_WAVES bands, one wave per band; consecutive parts of a
 * chain run in different workgroups of one XCD and hand the rows over
 * through the matrix itself (a progress counter per hand-off). */
#define FTAMD_SYS_LANES 16   /* lanes (sequence slots) of a band row */
#define FTAMD_SYS_COLS 8     /* columns of a chain: 4 rows of 16 lanes x 2 per lane */
#define FTAMD_SYS_WAVES 16   /* most bands of a chain part (waves of its workgroup) */
#define FTAMD_SYS_UNROLL 8   /* steps per loop iteration (a band's steps are a multiple) */
#define FTAMD_SYS_PREFETCH 16 /* most steps of table entries fetched ahead; tables carry that many spare steps */

/* Within a step of a band table the (s, c) pairs of the LANES x S sequences
 * are slot-major ([S][LANES]) in FP64 and lane-major ([LANES][S]) in FP32
 * (lane i's slot s is sequence S*i + s of the band). */
#define FTAMD_SYS_SLOT_MAJOR(precision) ((precision) == 0)

typedef struct {
    long long tab;    /* entry offset of the band's table: [steps + PREFETCH] steps of LANES*S (s, c) pairs */
    int R;            /* the band's row stream is rows 0 .. R of the class */
    int steps;        /* steps the band runs: R + (S+1)*(LANES-1) + S + 1 rounded up to UNROLL */
    int gbase;        /* position of the band's first sequence in the class, in application order */
    int nseq;         /* sequences of the class in the band (the rest are identities) */
} FtamdSysBand;

typedef struct {
    int band0;        /* its first band in the band array */
    int nb;           /* bands, <= FTAMD_SYS_WAVES */
    int dir;          /* 0 hi2lo: stream rows R .. 0; 1 lo2hi: rows 0 .. R */
    int rstep, rpar;  /* matrix row of class row e: rstep*e + rpar */
    int prio;         /* wave issue priority 0..3 (longest chains highest) */
    int lda;          /* column stride of its matrix in elements */
    int split_in;     /* progress counter of the part before it (its first band reads that part's
                         rows from A as they are published), -1 for a chain's first part */
    int split_out;    /* progress counter its last band publishes (it stores every row it emits
                         to A), -1 for a chain's last part */
    int col[FTAMD_SYS_COLS];  /* [row][c]: column of A, -1 if none */
    int glo[FTAMD_SYS_COLS];  /* the column applies the sequences at positions glo .. ghi */
    int ghi[FTAMD_SYS_COLS];
    long long aoff;   /* element offset of the chain's matrix from the launch's A */
} FtamdSysChain;

/* One chain part of a slot's list: its bands run on waves wave0 .. wave0+nb-1.
 * A slot's workgroup streams the parts of its list one after another through
 * its waves (FTAMD_SYS_WAVES of them). */
typedef struct {
    int chain;        /* index in the chain array */
    int wave0;        /* wave of its first band */
} FtamdSysItem;

/* Control words of a launch, in one 256-byte area (zero at upload; the
 * kernel's last workgroup resets all but the error word): */
#define FTAMD_SYS_CTL_BYTES 256
#define FTAMD_SYS_CTL_ERR 0        /* int: FTAMD_ERR_* bits */
#define FTAMD_SYS_CTL_DONE 1       /* unsigned: workgroups finished */
#define FTAMD_SYS_CTL_COMPLETED 2  /* unsigned: slots run */
#define FTAMD_SYS_CTL_TICKET 16    /* unsigned[FTAMD_MAX_XCD]: next slot of each XCD */

typedef struct {
    const FtamdSysBand *bands;
    const FtamdSysChain *chains;
    int nchains;
    const FtamdSysItem *items;  /* every slot's list of chain parts, one after another */
    const int *slot_items;      /* [nslots+1]: slot s runs items slot_items[s] .. slot_items[s+1]-1 */
    const int *xcd_slot;        /* [nxcd+1]: the slots of XCD x are xcd_slot[x] .. xcd_slot[x+1]-1; a
                                   part that waits on another runs in a later slot of the same XCD */
    int nslots;
    int nxcd;
    int grid;                   /* workgroups: nxcd * (most slots of one XCD) */
    int rowmask;      /* both columns of every lane row apply the same sequences (sph, disk) */
    const void *tab;  /* base of the band tables (bands address them by entry offsets from it) */
    void *A;          /* base of the matrices (chains address them by element offsets from it) */
    int *ctl;         /* control words (FTAMD_SYS_CTL_*) */
    int *prog;        /* [nprog] rows published by each split part, zero at launch; reset by the kernel */
    int nprog;
    int spin_limit;   /* polls before a wait gives up (FTAMD_ERR_TIMEOUT) */
    unsigned long long *trace;  /* diagnostic per-wave timeline (instrumented builds only), else NULL */
} FtamdSysLaunch;

int ftamd_hip_sys_launch(const FtamdSysLaunch *a, int precision, int S, void *stream);

/* Triangular products of the harmonic transforms (drivers.c:833-1021), out
 * of place from X (n x M, column-major) into Y (same shape, another buffer).
 * Left (right == 0): for each class c < ncls, the columns j = c,
 * c + cstride, ... < M are multiplied by op(U[c]) (n x n upper triangular,
 * leading dimension n; op = identity, or the transpose when trans);
 * U[c] == NULL leaves class c out.  Right (right == 1, ncls 1, cstride 1):
 * Y = X op(U[0]) with U[0]'s leading M x M block (leading dimension n).
 * Output tiles of FTAMD_TRMM_ROWS x FTAMD_TRMM_COLS; ctiles[c] column tiles
 * per class. */
#define FTAMD_TRMM_ROWS 64
#define FTAMD_TRMM_COLS 32
typedef struct {
    const double *U[4];
    const double *X;
    double *Y;
    int n, M, cstride, ncls, trans, right;
    int ctiles[4];
} FtamdTrmmLaunch;
int ftamd_hip_trmm_launch(const FtamdTrmmLaunch *a, void *stream);
/* Synthesis / analysis on the sphere's grid (ft_grid.hip): hipFFT plans and
 * work buffers for N x M (M odd).  type3: a synthesis plan (DCT-III / DST-III
 * along theta, c2r along phi), else an analysis plan (DCT-II / DST-II, r2c);
 * plan_mode: 0 the sphere's column classes, 1 the vector field's, 2 the
 * disk's (DCT-IV on the sine class in both kinds), 3 the annulus's (one
 * class; w: its N radial weights, applied on columns 1 and 2 mod 4 after the
 * theta transform of a synthesis plan or before that of an analysis plan,
 * multiplied (wmode 1) or divided (wmode 2)); exec_mode the executor's
 * classes for the fix-ups, in the same coding.  ftamd_grid_run: with a
 * synthesis plan, fix-ups (f0 on row 0 of the executor's cosine columns, f1
 * on the last row of its sine columns), the theta transforms, scalings (a on
 * every entry, b on column 0 again) and the phi transform; with an analysis
 * plan the phi transform, the scalings, the theta transforms and the
 * fix-ups.  X (device, N x M) is overwritten in place. */
typedef struct FtamdGrid FtamdGrid;
int ftamd_grid_create(int N, int M, int type3, int plan_mode, const double *w, FtamdGrid **out);
void ftamd_grid_destroy(FtamdGrid *g);
int ftamd_grid_run(FtamdGrid *g, double *X, int exec_mode, double f0, double f1, double a, double b, int wmode,
                   void *stream);

/* The triangle's two-dimensional DCTs (ft_grid.hip): DCT-III (type3) or
 * DCT-II along both axes of the buffer as FFTW's r2r_2d(N, M) reads it (an
 * N x M row-major array).  Before the first axis: if pre_fix, x2 where the
 * buffer index e < N and again where e % N == 0, then x pre_mul and
 * / pre_div (0: none); after the second axis: if post_fix, x0.5 at the same
 * places, then / post_div (0: none). */
typedef struct FtamdTriGrid FtamdTriGrid;
int ftamd_trigrid_create(int N, int M, int type3, FtamdTriGrid **out);
void ftamd_trigrid_destroy(FtamdTriGrid *g);
int ftamd_trigrid_run(FtamdTriGrid *g, double *X, int pre_fix, double pre_mul, double pre_div, int post_fix,
                      double post_div, void *stream);
/* The same for FFTW's r2r_3d(N, L, M) (dims in FFTW's order, rank 2 or 3);
 * the face scalings act per zero coordinate of the column-major view
 * (e = i + N (j + L k)); post_mul scales after the transform. */
int ftamd_boxgrid_create(int rank, const int *dims, int type3, FtamdTriGrid **out);
/* The rectangle-disk's grid (ft_grid.hip): passes of DCT-III (type3) or
 * DCT-II along the columns (axis 0; DST on the odd columns) or the rows
 * (axis 1) of the N x M column-major array, with a factor on a set of entries
 * before (pre_mask, pre_f) and after (post_mask, post_f) it -- 1: row 0 of
 * the even columns, 2: the last row of the odd columns, 3: column 0 -- then
 * x post_mul and / post_div (0: none). */
/* The spin-weighted sphere's grid (ft_grid.hip): complex N x M (M odd);
 * type3: a synthesis plan (DCT-III / DST-III, backward row DFT), else an
 * analysis plan; odd_spin: the vector field's column classes.  A run with a
 * synthesis plan: fix-ups (f0 on row 0 of the cosine columns, f1 on the last
 * row of the sine columns), theta, x a, colswap, row DFT; with an analysis
 * plan: row DFT, colswap's inverse, x a, theta, fix-ups. */
typedef struct FtamdSpinGrid FtamdSpinGrid;
int ftamd_spingrid_create(int N, int M, int type3, int odd_spin, FtamdSpinGrid **out);
void ftamd_spingrid_destroy(FtamdSpinGrid *g);
int ftamd_spingrid_run(FtamdSpinGrid *g, double *X, double f0, double f1, double a, void *stream);
typedef struct FtamdRectGrid FtamdRectGrid;
int ftamd_rectgrid_create(int N, int M, int type3, FtamdRectGrid **out);
void ftamd_rectgrid_destroy(FtamdRectGrid *g);
int ftamd_rectgrid_pass(FtamdRectGrid *g, double *X, int axis, int pre_mask, double pre_f, int post_mask, double post_f,
                        double post_mul, double post_div, void *stream);
int ftamd_boxgrid_run(FtamdTriGrid *g, double *X, int pre_fix, double pre_mul, double pre_div, int post_fix,
                      double post_mul, double post_div, void *stream);

/* The spin-weighted sphere's first stage (ft_k_spin.hip) on the complex
 * columns c0 .. c0+count-1 of A (interleaved re, im): column c at element
 * 2*(col_stride*(c - c0)) with rows at complex stride row_stride, of order
 * m(c) (the executors' layout: 0, -1, 1, -2, 2, ...) or of order fixed_m for
 * every column if use_fixed.  tab2: the plan's (s2, c2) pairs. */
int ftamd_hip_spin_stage1(const double *tab2, int n, int s, double *A, int c0, int count, long long col_stride,
                          int row_stride, int use_fixed, int fixed_m, int dir, void *stream);
/* A[i + j n] *= (j == 0 ? s1 : 1) * (i == 0 ? s1 : 1) * s2, in that order,
 * for the n x M matrix A (the 2D Chebyshev normalization, drivers.c:6-24). */
int ftamd_hip_cheb_scale(double *A, int n, int M, double s1, double s2, void *stream);
/* On the columns j = 1, 2 mod 4 of the n x M matrix A: x <- R^{-1} x
 * (solve, trans 0), R^{-T} x (solve, trans 1), R x or R^T x (solve 0), R the
 * upper bidiagonal device matrix {diagonal[n], superdiagonal[n-1]}. */
int ftamd_hip_bidiag(double *A, int n, int M, const double *R, int solve, int trans, void *stream);
int ftamd_hip_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* The allocation-free entry point's device scratch: one fixed array per
 * device, FTAMD_NOALLOC_SCRATCH_BYTES long, whose first 256 bytes are the
 * error word's area and the rest the chain records of the launch. */
#define FTAMD_NOALLOC_SCRATCH_BYTES (256 * 1024)
#define FTAMD_MAX_DEVICES 64
int ftamd_hip_scratch(void **ptr, size_t *bytes);
int ftamd_hip_sys_supported(int S, int precision);

int ftamd_hip_device(int *device, int *num_cu);
int ftamd_hip_xcd_count(int *nxcd);
int ftamd_hip_arch(char *name, int cap);
int ftamd_hip_pointer_kind(const void *ptr, int *kind);
int ftamd_hip_malloc(void **ptr, size_t bytes);
int ftamd_hip_free(void *ptr);
int ftamd_hip_memset(void *ptr, int value, size_t bytes);
int ftamd_hip_h2d(void *dst, const void *src, size_t bytes, void *stream);
int ftamd_hip_d2h(void *dst, const void *src, size_t bytes, void *stream);
int ftamd_hip_sync(void *stream);
int ftamd_hip_device_sync(void);
int ftamd_hip_last_runtime_error(void);
int ftamd_hip_launch(const FtamdLaunch *a, int precision, int kind, int dir, int L, int W, int grid, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FT_HIP_ABI_H */
