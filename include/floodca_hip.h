/*
 * floodca_hip.h -- C-ABI of the MI355X (gfx950) backend for the FloodCA
 * weighted-cellular-automata step.
 *
 * Two groups of declarations live here:
 *
 *  1. The data-structure ABI the reference's backends share (FcaPhys,
 *     FcaRunCfg, FcaGrid, FcaAcc, FcaTimers). A caller that already includes
 *     the reference's src/common/fca_types.h gets its definitions (the guard
 *     FCA_TYPES_H is shared), so this library links beside serial/omp/cuda with
 *     no shim. Built on its own, the same layouts are restated below; field
 *     order, types and array sizes must stay identical to
 *     src/common/fca_types.h:34-159, which tests/test_abi.py checks with
 *     sizeof/offsetof probes in both precisions.
 *
 *  2. The backend entry points. fca_hip_* mirror src/cuda/cuda_backend.h:12-32
 *     one to one (same arguments, same contract), except that device-API
 *     failures return nonzero instead of calling exit(1). The fca_hip_sim_*
 *     session and fca_hip_ens_counts_u32 have no reference counterpart; they
 *     exist for callers that keep a grid resident in HBM across calls
 *     (benchmarks, parity tests, screening loops over >65535 scenarios).
 *
 * `real` is float unless FCA_DOUBLE is defined, as in the reference. The
 * library is built once per precision: libfloodca_hip.so (float) and
 * libfloodca_hip_f64.so (double); each exports the same symbol names.
 */
#ifndef FLOODCA_HIP_H
#define FLOODCA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifndef FCA_TYPES_H
#define FCA_TYPES_H

#ifdef FCA_DOUBLE
typedef double real;
#else
typedef float real;
#endif

/* Cell classes of mask.bin (src/common/fca_types.h:15-19). */
enum { FCA_LAND = 0, FCA_SEA = 1, FCA_WALL = 2 };
/* Domain-edge treatment (fca_types.h:21-24). */
enum { FCA_BND_OPEN = 0, FCA_BND_CLOSED = 1 };
/* Unit hyetograph shape (fca_types.h:26-29). */
enum { FCA_HYETO_TRIANGLE = 0, FCA_HYETO_UNIFORM = 1 };

#define FCA_MAX_FORCE_DAYS 64
#define FCA_DAY_S 86400.0

/* Physical constants and scheme switches (fca_types.h:34-49). */
typedef struct FcaPhys {
    real dx;       /* cell size (m) */
    real A;        /* cell area dx*dx (m^2) */
    real g;        /* gravity */
    real alpha;    /* CFL-like factor of the adaptive step */
    real tau_h;    /* depth below which a face or cell is dry */
    real tau_l;    /* level difference below which a face carries nothing */
    real z_sea;    /* fixed sea level of SEA cells */
    real dt_max;
    real dt_min;
    real S_bnd;    /* slope of the open-boundary virtual neighbour */
    real h_flood;
    int U;         /* recompute dt every U steps */
    int K;         /* 4 or 8 neighbours */
    int boundary;  /* FCA_BND_OPEN or FCA_BND_CLOSED */
} FcaPhys;

/* Run configuration (fca_types.h:51-86). The long char arrays are part of the
 * ABI even where this backend does not read them. */
typedef struct FcaRunCfg {
    FcaPhys phys;
    real T_sim;
    real out_every;
    int diag_every;
    int progress_every;
    int hyeto;
    real D_h;
    real phi;
    real lambda;
    int steps;
    int no_output;
    int neigh;
    char config_path[512];
    char grid_dir[512];
    char scenario[128];
    char scenario_file[512];
    char warmstart[512];
    char rain_file[512];
    char out_dir[512];
    char csv_path[512];
    char exp_id[32];
    int repeat;
    int threads;
    int ranks;
    int gpus;
    int block_x;
    int block_y;
    char halo_mode[16];
    char reduce_impl[16];
    char gpu_model[64];
    char backend[32];
    char forcing_csv[512];
    int n_force_days;
    char force_files[FCA_MAX_FORCE_DAYS][768];
} FcaRunCfg;

struct FcaGrid;

/* Distributed-memory hooks (fca_types.h:91-98). Always NULL for this backend. */
typedef struct FcaComm {
    void *ctx;
    int rank;
    void (*halo)(struct FcaGrid *g, real *arr, void *ctx);
    real (*min_real)(real v, void *ctx);
    double (*sum_double)(double v, void *ctx);
    int (*load_rday)(struct FcaGrid *g, const char *path, void *ctx);
} FcaComm;

/* Host-side grid (fca_types.h:100-126): row-major, a = r*ncols + c. Only
 * z, h, n, C, mask, Rday, hmax and tmax are read or written by this backend;
 * s, out and the face arrays are the CPU backends' scratch. */
typedef struct FcaGrid {
    int nrows;
    int ncols;
    int r0, r1, c0, c1;
    int edge_n, edge_s, edge_w, edge_e;
    FcaComm *comm;
    real dx;
    real x0;
    real y0;
    real *z;
    real *h;
    real *n;
    real *C;
    uint8_t *mask;
    real *Rday;
    real *hmax;
    real *tmax;
    real *s;
    real *out;
    real *Qx;
    real *Qy;
    real *Qd1;
    real *Qd2;
    real *Qopen;
} FcaGrid;

/* Run accumulators (fca_types.h:128-142). */
typedef struct FcaAcc {
    double V;
    double V0;
    double V_rain;
    double V_sea;
    double V_bnd;
    double clamp_err;
    long cap_active;
    long cap_total;
    long dt_min_hits;
    double t;
    int step;
    real dt_base;
    int force_day;
} FcaAcc;

/* Phase timers (fca_types.h:144-159); negative = does not apply. */
typedef struct FcaTimers {
    double wall_total;
    double t_init;
    double t_flux;
    double t_limiter;
    double t_update;
    double t_dt;
    double t_io;
    double t_h2d;
    double t_d2h;
    double t_halo;
    double t_allreduce;
    double bytes_per_step;
    long active_cells;
} FcaTimers;

#endif /* FCA_TYPES_H */

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Entry points with a reference counterpart ------------------------- */

/* Replaces fca_cuda_run (src/cuda/cuda_backend.h:12-13, kernels.cu:617-776).
 * Runs the whole simulation on the current device: cfg->steps > 0 steps, else
 * until acc->t reaches cfg->T_sim. The caller has run fca_acc_init (V0 set).
 * On return g->h, g->hmax and g->tmax hold the device results; acc gets
 * V_rain, clamp_err, V_sea, V_bnd, cap_active and cap_total added, and t, step,
 * dt_base, dt_min_hits, force_day and the final V set; tm gets wall_total,
 * t_init, t_h2d, t_d2h, t_dt, t_io and bytes_per_step (phases that are fused
 * into one kernel are reported as -1). cfg->reduce_impl is set to "wave" when
 * empty. Returns 0, -1 on a configuration error, -2 on a device error (the
 * message is in fca_hip_last_error()). */
int fca_hip_run(FcaGrid *g, FcaRunCfg *cfg, FcaAcc *acc, FcaTimers *tm);

/* Replaces fca_cuda_device_name (cuda_backend.h:14). */
void fca_hip_device_name(char *name, size_t n);

/* Replaces fca_cuda_select_device (cuda_backend.h:16-17): binds the calling
 * thread to device local_rank % device_count and returns that index, or -1. */
int fca_hip_select_device(int local_rank);

/* Ensemble session; replaces fca_cuda_ens_open/run/counts/close
 * (cuda_backend.h:19-32, kernels.cu:836-1043). The static fields (z, n, C,
 * mask) stay on the device for all scenarios; each fca_hip_ens_run uploads h,
 * hmax, tmax and Rday from g, runs like fca_hip_run, and adds hmax >= thr[j]
 * on LAND into per-cell counts kept on the device. flooded[j] receives the
 * scenario's LAND cell count at each threshold and *max_depth its largest
 * LAND hmax (at least 0). g->h, g->hmax and g->tmax are not downloaded. */
typedef struct FcaHipEns FcaHipEns;
FcaHipEns *fca_hip_ens_open(const FcaGrid *g, FcaRunCfg *cfg, int nthr, const double *thr);
int fca_hip_ens_run(FcaHipEns *e, FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc,
                    FcaTimers *tm, long *flooded, double *max_depth, void (*poll)(void *),
                    void *pctx, int poll_every);
/* uint16 counts, nthr * N values, threshold-major (kernels.cu:1026-1031).
 * Counts above 65535 saturate. */
int fca_hip_ens_counts(FcaHipEns *e, uint16_t *host);
void fca_hip_ens_close(FcaHipEns *e);

/* ---- Entry points with no reference counterpart ------------------------ */

/* uint32 variant of fca_hip_ens_counts for ensembles beyond 65535 members. */
int fca_hip_ens_counts_u32(FcaHipEns *e, uint32_t *host);

/* Message of the last failed call on this thread ("" if none). */
const char *fca_hip_last_error(void);

/* Step kernel families. FCA_HIP_FUSED is the default: one launch per step
 * reads each field once and writes h once. FCA_HIP_SPLIT runs the step as
 * separate flux / limiter / update / time-step launches (the reference CUDA
 * backend's phase order with the time step kept on the device); it is kept
 * as an independent second implementation for cross-checks. */
enum { FCA_HIP_FUSED = 0, FCA_HIP_SPLIT = 1 };

/* Per-session tuning; zero fields take the defaults. */
typedef struct FcaHipOpts {
    int kernel;        /* FCA_HIP_FUSED or FCA_HIP_SPLIT */
    int strip_rows;    /* fused: rows walked by one wavefront (default 96) */
    int graph_batch;   /* launches between controller read-backs, replayed as one
                        * captured hipGraph when a multiple of 6 (default 30;
                        * 1 = every launch issued on its own) */
    int record_steps;  /* > 0: keep a per-step log of the first record_steps steps */
    int max_blocks;    /* fused: cap on workgroups, each then walking several strips
                        * (default 0: one strip per wavefront, hardware-dispatched) */
    int flags;         /* FCA_HIP_NO_LAND_FORM: the fused walk never takes its LAND
                        * form (the same results; for comparisons) */
    int tiles_per_wave; /* fused: tiles each wavefront walks per launch (default 1):
                         * k > 1 launches a k-th of the workgroups, each wavefront
                         * walking k strips in turn (more cells per lane) */
    int cols_per_lane;  /* fused: adjacent columns each work-item walks, 1 or 2
                         * (default 1): tiles of 58 or 122 written columns (the
                         * same results; 2 shares each row's work over two
                         * columns at fewer wavefronts per SIMD) */
} FcaHipOpts;

/* flags: FCA_HIP_NO_LAND_FORM as above; FCA_HIP_NO_SKIP walks every tile
 * every step (by default a tile that is dry, rain-free and unchanged with
 * its neighbours is skipped, which leaves every result as it is). */
enum { FCA_HIP_NO_LAND_FORM = 1, FCA_HIP_NO_SKIP = 2 };

/* One executed step, as the controller saw it (fca_step, model.c:363-382). */
typedef struct FcaHipStepRec {
    double t;        /* time at the start of the step */
    double F;        /* hyetograph integral over [t, t+dt] */
    real dt_base;    /* held step after the dt_min/dt_max clamp */
    real dt;         /* step actually taken (event-clamped) */
    int step;        /* index of this step (acc->step before it) */
    int hit_event;
} FcaHipStepRec;

/* A grid resident in HBM. open uploads every field of g once and computes
 * the first time step; acc supplies V0, t, step, force_day and dt_base. */
typedef struct FcaHipSim FcaHipSim;
FcaHipSim *fca_hip_sim_open(const FcaGrid *g, const FcaRunCfg *cfg, const FcaAcc *acc,
                            const FcaHipOpts *opts);
/* Enqueues up to nsteps step launches (fewer if the run ends at T_sim when
 * cfg->steps <= 0). Asynchronous: returns once the launches are queued.
 * Returns the number of launches queued, or < 0 on error. */
int fca_hip_sim_advance(FcaHipSim *s, int nsteps);
/* Waits for the queued work. Returns 0 or < 0. */
int fca_hip_sim_sync(FcaHipSim *s);
/* Waits, then copies the state back: any of h, hmax, tmax may be NULL; acc
 * (may be NULL) receives the accumulators as fca_hip_run would set them. */
int fca_hip_sim_read(FcaHipSim *s, real *h, real *hmax, real *tmax, FcaAcc *acc);
/* Copies up to cap step records into out; returns how many were recorded. */
int fca_hip_sim_steplog(FcaHipSim *s, FcaHipStepRec *out, int cap);
/* Runs until the reference loop would stop (fca_run, model.c:607-643), with
 * multi-day forcing reloads at day boundaries. Returns 0 or < 0. */
int fca_hip_sim_run(FcaHipSim *s);
/* HIP stream the session launches on (a hipStream_t), for callers that time
 * it with their own events. */
void *fca_hip_sim_stream(FcaHipSim *s);
void fca_hip_sim_close(FcaHipSim *s);

/* Batched scenarios: n grids (sizes may differ, K may not) resident at once,
 * each with its own configuration, time step and accumulators; one launch
 * advances every unfinished scenario by one step, so small grids share the
 * GPU instead of each filling a fraction of it. Results are those of
 * running each scenario alone, bit for bit. acc is an array of n (V0, t,
 * step, force_day, dt_base as for fca_hip_sim_open); multi-day forcing is not
 * supported. No counterpart in the reference (its ensemble runs scenarios one
 * after another, kernels.cu:930-997). */
typedef struct FcaHipBatch FcaHipBatch;
FcaHipBatch *fca_hip_batch_open(int n, const FcaGrid *const *g, const FcaRunCfg *const *cfg,
                                const FcaAcc *acc, const FcaHipOpts *opts);
int fca_hip_batch_nscen(const FcaHipBatch *b);
/* Runs every scenario to its end (T_sim, or cfg->steps). Returns 0 or < 0. */
int fca_hip_batch_run(FcaHipBatch *b);
/* As fca_hip_sim_read / fca_hip_sim_steplog for scenario i. */
int fca_hip_batch_read(FcaHipBatch *b, int i, real *h, real *hmax, real *tmax, FcaAcc *acc);
int fca_hip_batch_steplog(FcaHipBatch *b, int i, FcaHipStepRec *out, int cap);
/* Scenario i's LAND cells with hmax >= thr[j] (nthr <= 4) and its largest
 * LAND hmax, as the ensemble reports them (ens_common.c:166-188). */
int fca_hip_batch_flooded(FcaHipBatch *b, int i, int nthr, const double *thr, long *flooded,
                          double *max_depth);
/* Adds, for every scenario of the batch (all of one shape, nthr <= 4), the
 * LAND cells with hmax >= thr[j] into host[j*N + a] (uint32, threshold-major
 * as fca_hip_ens_counts_u32), so the batches of one ensemble accumulate into
 * one raster; the count rasters of src/hybrid/ens_common.c:166-221. */
int fca_hip_batch_add_counts_u32(FcaHipBatch *b, int nthr, const double *thr, uint32_t *host);
/* Scenario i's steps taken and simulated time. */
int fca_hip_batch_status(FcaHipBatch *b, int i, int *step, double *t);
/* Seconds spent in fca_hip_batch_run so far. */
double fca_hip_batch_wall(const FcaHipBatch *b);
/* Device addresses of scenario i's hmax (real, nr x nc row-major) and mask
 * (uint8), valid until fca_hip_batch_close and complete once
 * fca_hip_batch_run has returned: for consumers that reduce the product on
 * the device without a copy to the host (no counterpart in the reference,
 * whose workers copy hmax back, src/hybrid/ens_common.c:166-188). */
int fca_hip_batch_device_fields(FcaHipBatch *b, int i, void **hmax, void **mask, int *nr,
                                int *nc);
void fca_hip_batch_close(FcaHipBatch *b);

/* Crop scenarios built on the device. A base grid (bed, mask and a land-cover
 * class per cell) is uploaded once and stays resident; each scenario is a
 * window of it with its own roughness per class and a uniform Rday (C = 1),
 * starting dry (h = hmax = tmax = 0; accumulators as fca_acc_init gives them
 * for a dry grid, force_day = -1). Opening a batch of such scenarios copies
 * nothing from the host but the scenario table, so the cost of assembling
 * and uploading every crop is gone. The scenarios are those an FcaGrid
 * holding the same values would give to fca_hip_batch_open, bit for bit. No
 * counterpart in the reference, whose ensemble keeps one grid and varies the
 * forcing (src/hybrid/main_ens_mpi_cuda.c:83-96). */
#define FCA_HIP_MAX_CLASSES 16
typedef struct FcaHipBase FcaHipBase;
/* g supplies z and mask (its other fields are not read); cls holds a class
 * in [0, FCA_HIP_MAX_CLASSES) per cell. NULL on error (fca_hip_last_error). */
FcaHipBase *fca_hip_base_open(const FcaGrid *g, const uint8_t *cls);
void fca_hip_base_close(FcaHipBase *b);
typedef struct FcaHipCrop {
    int r0, c0, nrows, ncols;              /* window of the base grid */
    real rday;                             /* Rday of every cell (C = 1) */
    real n_class[FCA_HIP_MAX_CLASSES];     /* Manning n of each class */
} FcaHipCrop;
/* n scenarios from crops[i] with configuration cfg[i] (same K for all). */
FcaHipBatch *fca_hip_batch_open_crops(const FcaHipBase *base, int n, const FcaHipCrop *crops,
                                      const FcaRunCfg *const *cfg, const FcaHipOpts *opts);

/* ---- Entry point that allocates nothing ---------------------------------- */

/* Largest grid, in cells, that fca_hip_run_noalloc accepts (2048 x 2048). */
#define FCA_HIP_NOALLOC_MAX_CELLS (1 << 22)

/* No counterpart in the reference. fca_hip_run for callers that cannot
 * allocate device memory: tight serving loops and harnesses that forbid
 * hipMalloc. g's z, h, n, C, mask, Rday, hmax and tmax are DEVICE pointers
 * owned by the caller (h, hmax, tmax are read and receive the results, the
 * others are only read); the run is fca_hip_run's, with the same kernels,
 * results bit for bit, accumulators and timers (nothing crosses PCIe: t_h2d
 * adds nothing, t_d2h times the final read-back of the accumulators).
 * stream is a hipStream_t (NULL: the default stream); the call returns once
 * the run is complete. It calls no hipMalloc, hipFree, stream or graph
 * creation, and no host allocation: its workspace is one array per device,
 * reserved with the library's code object (about 2 * 4 * 2^22 B + 4 MiB in
 * float, twice the per-cell part in double) and held until the process ends.
 * Calls that overlap on one device share that array and race. Rejected with
 * -1 (message in fca_hip_last_error()): more than FCA_HIP_NOALLOC_MAX_CELLS
 * cells, a NULL field, K other than 4 or 8, alpha*dx <= 0, multi-day
 * forcing (its day files are read on the host). -2 on a device error. */
int fca_hip_run_noalloc(FcaGrid *g, FcaRunCfg *cfg, FcaAcc *acc, FcaTimers *tm, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FLOODCA_HIP_H */
