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

int ftamd_hip_device(int *device, int *num_cu);
int ftamd_hip_xcd_count(int *nxcd);
int ftamd_hip_pointer_kind(const void *ptr, int *kind);
int ftamd_hip_malloc(void **ptr, size_t bytes);
int ftamd_hip_free(void *ptr);
int ftamd_hip_memset(void *ptr, int value, size_t bytes);
int ftamd_hip_h2d(void *dst, const void *src, size_t bytes, void *stream);
int ftamd_hip_d2h(void *dst, const void *src, size_t bytes, void *stream);
int ftamd_hip_sync(void *stream);
int ftamd_hip_last_runtime_error(void);
int ftamd_hip_launch(const FtamdLaunch *a, int precision, int kind, int dir, int L, int W, int grid, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FT_HIP_ABI_H */
