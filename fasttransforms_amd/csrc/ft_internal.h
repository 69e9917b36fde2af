/*
 * ft_internal.h -- private definitions of the C host layer.
 */
#ifndef FT_INTERNAL_H
#define FT_INTERNAL_H

#include <pthread.h>

#include "fasttransforms_amd.h"
#include "ft_hip_abi.h"

/* A launch-ready schedule: groups, lane maps and the task order on the
 * device, plus the kernel's ticket/progress workspace. */
typedef struct ft_schedule {
    struct ft_schedule *next;
    /* key (single-vector schedules: m1, m2 = first and last sequence j, rs = stride) */
    int M, dir, precision, L, W, rs, m1, m2, single;
    int layout;              /* FTAMD_LAYOUT_* of its columns (part of the key) */
    int slab, order0;        /* tetrahedron stages: columns per slab and the first slab's order
                                (with M slabs and rs = rows per slab, part of the key) */
    /* launched asynchronously since its error word was last read (ft_gpu_synchronize) */
    int pending;
    void *pending_stream;
    /* device arrays (one allocation), and the boundary buffers rows pass
     * through from one task of a group to the next */
    void *d_mem;
    void *d_bnd;
    FtamdLaunch launch;
    int ngroups, ntasks;
    int xcd_tasks;           /* most tasks any XCD's list holds */
    /* band-systolic engine (sys != 0): chains on the device, launch record */
    int sys;
    int batch;         /* a batch schedule (grid capped at 2 workgroups per CU by default) */
    int maxpart;       /* bands per chain part it was built with (part of its key) */
    int wg_per_cu;     /* pipelined: workgroups per CU it was built for, 0: no cap */
    FtamdSysLaunch sys_launch;
    void *d_sys;
    double flops;            /* algorithmic flops (6 per rotation) */
} ft_schedule;

/* The reference's layout (ftinternal.h:370-374) first, then device state. */
struct ft_rotation_plan_s {
    double *s;
    double *c;
    int n;
    /* extension */
    int kind;                /* geometry of the tables: FTAMD_SPH, _TRI or _DISK */
    int family;              /* its constructor: FT_GPU_SPH, _TRI, _DISK, _RECTDISK or FTAMD_FAMILY_ANNULUS
                                (a batch takes plans of one family and one parameter set) */
    double alpha, beta, gamma, rho;
    long nsc;                /* entries per table */
    int device;              /* HIP device of the uploads, -1 before upload */
    void *d_tab[2];          /* interleaved (s,c) with padding: [0] FP64, [1] FP32 */
    ft_schedule *scheds;
    struct ft_sys_tables *sys[2][2];  /* band-systolic tables [direction][precision] */
    void *d_stage;           /* staging buffer of host-pointer calls (used under exec_lock) */
    size_t stage_bytes;
    unsigned long long serial; /* distinct for every plan the process creates */
    /* triangle plans of smaller sizes with this plan's parameters (the
     * tetrahedron's block rotations on slabs of fewer columns), freed with it */
    struct ft_rotation_plan_s *subs, *sub_next;
    pthread_mutex_t lock;      /* device state: uploads, schedule lists */
    pthread_mutex_t exec_lock; /* held for the whole of a synchronous executor or kernel call */
};

typedef struct ft_sys_tables ft_sys_tables;

/* Family of ft_plan_rotannulus plans (disk tables and kernels). */
#define FTAMD_FAMILY_ANNULUS 5

/* ft_plans.c */
ft_rotation_plan *ftamd_plan_alloc(int kind, int family, int n, double alpha, double beta, double gamma);
int ftamd_plan_upload(ft_rotation_plan *RP, int precision);
const void *ftamd_plan_table(const ft_rotation_plan *RP, int precision);

/* ft_connection.c: connection matrices (transforms.c:50-93), malloc'd n x n
 * column-major upper triangular, NULL if out of memory */
double *ftamd_plan_leg2cheb(int normleg, int normcheb, int n);
double *ftamd_plan_cheb2leg(int normcheb, int normleg, int n);
double *ftamd_plan_ultra2ultra(int norm1, int norm2, int n, double lambda, double mu);
double *ftamd_plan_jac2jac(int norm1, int norm2, int n, double alpha, double beta, double gamma, double delta);
double *ftamd_plan_jac2cheb(int normjac, int normcheb, int n, double alpha, double beta);
double *ftamd_plan_cheb2jac(int normcheb, int normjac, int n, double alpha, double beta);
double *ftamd_plan_ultra2cheb(int normultra, int normcheb, int n, double lambda);
double *ftamd_plan_cheb2ultra(int normcheb, int normultra, int n, double lambda);

/* ft_drivers.c */
int ftamd_valid_M(int kind, int n, int M); /* widths a whole-matrix executor of public kind `kind` accepts */
void ftamd_set_error(int code);
int ftamd_tuning_grid(const ft_schedule *S);
void ftamd_schedule_free_device(ft_schedule *S);
void ftamd_schedule_release(ft_schedule *S);
void ftamd_batch_forget_plan(const ft_rotation_plan *RP);
int ftamd_engine(void);

/* The engine and shape a launch uses: the knobs (ft_gpu_set_*) where set,
 * else the autotuning table's row, else the built-in choice. */
typedef struct {
    int engine;     /* FT_GPU_ENGINE_PIPELINE or FT_GPU_ENGINE_SYSTOLIC (preferred; forced if the knob is) */
    int forced;     /* the engine knob is set */
    int maxpart;    /* band-systolic bands per chain part */
    int L, W;       /* pipelined sequences per wave, stage waves per workgroup */
    int wg_per_cu;  /* pipelined workgroups per CU, 0: no cap */
} ftamd_choice;
void ftamd_choose(int kind, int precision, int dir, int batch, int n, ftamd_choice *c);

/* ft_tuning.c */
int ftamd_tuned_lookup(int kind, int precision, int dir, int batch, int n, int *engine, int *maxpart, int *L, int *W,
                       int *wg_per_cu);

/* The reference's layout (ftinternal.h:425-432) first, then device state.
 * The sphere tables s1, c1 are those of the sphere plan (rotations.c:116-133
 * and 443-455 evaluate one closed form), which runs the second stage on the
 * band-systolic engine; s2, c2 go to the device interleaved for the first. */
struct ft_spin_rotation_plan_s {
    double *s1;
    double *c1;
    double *s2;
    double *c2;
    int n;
    int s;
    /* extension */
    ft_rotation_plan *sph;   /* sphere plan of size n: the second stage's tables, schedules,
                                staging buffer and execution lock */
    int device;              /* HIP device of the upload, -1 before */
    void *d_tab2;            /* (s2, c2) pairs, |s| n^2 of them (NULL if s = 0) */
    pthread_mutex_t lock;    /* the upload */
};

int ftamd_spin_upload(ft_spin_rotation_plan *P);

/* ft_systolic.c */
int ftamd_sys_choose_S(int kind, int n, int precision);
int ftamd_sys_tables_build(const ft_rotation_plan *RP, int dir, int precision, int S, ft_sys_tables **out);
void ftamd_sys_tables_free(ft_sys_tables *T);
int ftamd_sys_tables_S(const ft_sys_tables *T);
const void *ftamd_sys_tables_dev_bands(const ft_sys_tables *T);
const void *ftamd_sys_tables_dev_tab(const ft_sys_tables *T);
int ftamd_sys_chains(const ft_rotation_plan *RP, const ft_sys_tables *T, int M, int layout, int maxpart,
                     FtamdSysChain **out, int *count, int *nprog);
int ftamd_sys_assign(const FtamdSysChain *ch, int nch, const FtamdSysBand *bands, int nxcd, int nsx,
                     FtamdSysItem *items, int *slot_items, int *xcd_slot, long *ws_load, int *ws_chain);
const FtamdSysBand *ftamd_sys_tables_host_bands(const ft_sys_tables *T);
int ftamd_sys_chains_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int M, int layout, FtamdSysChain *buf,
                          int cap, int *count);
int ftamd_sys_split(FtamdSysChain *buf, int nfull, int cap, int maxpart, int *count, int *nprog);
int ftamd_sys_chains_tet_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int part, int N, int L, int m0,
                              int count, FtamdSysChain *buf, int cap, int *nout);
int ftamd_sys_chains_tet(const ft_rotation_plan *RP, const ft_sys_tables *T, int part, int N, int L, int m0, int count,
                         int maxpart, FtamdSysChain **out, int *nout, int *nprog);
int ftamd_sys_chains_spin_fill(const ft_rotation_plan *RP, const ft_sys_tables *T, int s, int M, FtamdSysChain *buf,
                               int cap, int *nout);
int ftamd_sys_chains_spin(const ft_rotation_plan *RP, const ft_sys_tables *T, int s, int M, int maxpart,
                          FtamdSysChain **out, int *nout, int *nprog);
int ftamd_sys_debug_tet(const ft_rotation_plan *RP, int dir, int S, int part, int N, int L, int m0, int count,
                        int maxpart, FtamdSysBand **bands, int *nbands, double **tab, long long *nentries,
                        FtamdSysChain **chains, int *nchains);
int ftamd_sys_batch(const ft_rotation_plan *const *plans, ft_sys_tables *const *tabs, void *const *dA,
                    const int *M, int count, int layout, int maxpart, FtamdSysBand **bands, int *nbands,
                    FtamdSysChain **chains,
                    int *nchains, int *nprog, const void **tab_base);

#endif /* FT_INTERNAL_H */
