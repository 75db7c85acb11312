/*
 * hip_backend.c -- C host of the MI355X backend: the entry points of
 * include/floodca_hip.h. It owns device memory, uploads and downloads, the
 * launch loop with its occasional controller read-back (multi-day forcing
 * reloads, end of run), and the ensemble session. All arithmetic of the step
 * runs in fca_kernels.hip.
 *
 * Built once per precision (real = float, or double with -DFCA_DOUBLE).
 */
#include "floodca_hip.h"
#include "fca_launch.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static __thread char g_err[512];

const char *fca_hip_last_error(void)
{
    return g_err;
}

static int fail(int rc, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return rc;
}

#define HIP_OR(expr, rc)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return fail((rc), "%s failed at %s:%d: %s", #expr, __FILE__, __LINE__,        \
                        hipGetErrorString(e_));                                           \
        }                                                                                 \
    } while (0)

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- resident session ------------------------------------------------------ */

struct FcaHipSim {
    FcaDev d;
    hipStream_t st;
    size_t N;
    long long k;           /* launches enqueued so far */
    FcaAcc acc0;           /* accumulators at upload time */
    long faces_per_step;   /* faces whose cap flag the reference counts per step */
    int batch;             /* launches between controller read-backs */
    FcaDevCtrl *h_ctrl;    /* host copy of a controller record */
    double *h_red;         /* host copy of reduction partials */
    real *h_stage;         /* pinned staging for forcing reloads */
    FcaRunCfg *cfg;        /* copy of the run configuration */
    double t_h2d, t_io;
    int n_ok;              /* every LAND roughness in the LAND form's range */
    int no_land_form;      /* FCA_HIP_NO_LAND_FORM was asked for */
    int own_stream;        /* st was created for this session */
    int skip_wanted;       /* quiet-tile skipping not switched off */
    int noalloc;           /* fca_hip_run_noalloc: no allocation of any kind */
    hipGraphExec_t gexec;  /* glen step launches, captured once */
    int glen;
};

/* Launch indices enter a launch only through k & 1 and k % 3, so a capture of
 * launches 0..n-1 with n a multiple of 6 replays as launches k..k+n-1 for
 * any k that is a multiple of 6. */
#define FCA_GRAPH_PERIOD 6

static int capture_steps(hipStream_t st, hipGraphExec_t *exec, int n,
                         int (*launch)(void *ctx, long long k), void *ctx)
{
    hipGraph_t gr = NULL;
    int j, rc = 0;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        return fail(-2, "hipStreamBeginCapture failed");
    }
    for (j = 0; j < n && rc == 0; j++) {
        rc = launch(ctx, j);
    }
    if (hipStreamEndCapture(st, &gr) != hipSuccess || rc != 0 || !gr) {
        if (gr) {
            hipGraphDestroy(gr);
        }
        return fail(-2, "step capture failed");
    }
    if (hipGraphInstantiate(exec, gr, NULL, NULL, 0) != hipSuccess) {
        hipGraphDestroy(gr);
        *exec = NULL;
        return fail(-2, "hipGraphInstantiate failed");
    }
    hipGraphDestroy(gr);
    return 0;
}

static int launch_sim_step(void *ctx, long long k)
{
    FcaHipSim *s = (FcaHipSim *)ctx;
    return fca_k_step(&s->d, k, s->st);
}

static void dev_free(FcaDev *d)
{
    if (d->arena) {
        hipFree(d->arena);
    }
    memset(d, 0, sizeof(*d));
}

/* Carves a session's device buffers out of one allocation: a batch opens
 * hundreds of sessions, and one hipMalloc/hipFree each instead of one per
 * buffer takes most of the cost out of opening and closing them. Every
 * buffer starts on a 256-byte boundary. */
typedef struct ArenaReq {
    void **ptr;
    size_t bytes;
} ArenaReq;

static int dev_arena(FcaDev *d, const ArenaReq *req, int n)
{
    size_t off = 0;
    int i;
    char *base;
    for (i = 0; i < n; i++) {
        off += (req[i].bytes + 255) & ~(size_t)255;
    }
    HIP_OR(hipMalloc(&d->arena, off ? off : 256), -2);
    base = (char *)d->arena;
    off = 0;
    for (i = 0; i < n; i++) {
        *req[i].ptr = base + off;
        off += (req[i].bytes + 255) & ~(size_t)255;
    }
    return 0;
}

void fca_hip_sim_close(FcaHipSim *s)
{
    if (!s) {
        return;
    }
    if (s->st) {
        hipStreamSynchronize(s->st);
    }
    dev_free(&s->d);
    free(s->h_ctrl);
    free(s->h_red);
    if (s->h_stage) {
        hipHostFree(s->h_stage);
    }
    if (s->gexec) {
        hipGraphExecDestroy(s->gexec);
    }
    if (s->st && s->own_stream) {
        hipStreamDestroy(s->st);
    }
    free(s->cfg);
    free(s);
}

/* Faces between cells (K=4: east and south; K=8 adds two diagonals). */
static long count_inner_faces(long nr, long nc, const FcaPhys *p)
{
    return nr * (nc - 1) + (nr - 1) * nc + (p->K == 8 ? 2 * (nr - 1) * (nc - 1) : 0);
}

static long count_faces(const FcaGrid *g, const FcaPhys *p)
{
    long nr = g->nrows, nc = g->ncols, f, r, c;
    f = count_inner_faces(nr, nc, p);
    if (p->boundary == FCA_BND_OPEN) {
        for (r = 0; r < nr; r++) {
            for (c = 0; c < nc; c++) {
                if (g->mask[r * nc + c] == FCA_LAND) {
                    f += (c == 0) + (c == nc - 1) + (r == 0) + (r == nr - 1);
                }
            }
        }
    }
    return f;
}

#define NRED 4096

/* Whether q0 = a*r, q = fma(fma(-q0, d, a), r, q0) with r = RN(1/d) equals the
 * correctly rounded a/d for every float a. Division, multiply and fma scale
 * exactly with a power of two while no intermediate leaves the normal range,
 * so one binade of mantissas decides it (the device applies the shortcut only
 * for 2^-80 <= |a| <= 2^80). Float precision only. */
static int div_const_exact(double dd)
{
#ifdef FCA_DOUBLE
    (void)dd;
    return 0;
#else
    const float d = (float)dd, r = 1.0f / d;
    uint32_t m;
    if (!(d > 0.0f) || d != d || d > 1e30f || d < 1e-30f) {
        return 0;
    }
    for (m = 0; m < (1u << 23); m++) {
        uint32_t bits = 0x3f800000u | m;
        float a, q0, q;
        memcpy(&a, &bits, 4);
        q0 = a * r;
        q = fmaf(fmaf(-q0, d, a), r, q0);
        if (q != a / d) {
            return 0;
        }
    }
    return 1;
#endif
}

/* div_const_exact with the answers remembered: a check walks 2^23 mantissas,
 * and a batch opens hundreds of sessions over the same few divisors. */
static int div_const_exact_cached(double dd)
{
    static double keys[16];
    static int vals[16], used;
    int i;
    for (i = 0; i < used; i++) {
        if (keys[i] == dd) {
            return vals[i];
        }
    }
    i = used < 16 ? used++ : 15;
    keys[i] = dd;
    vals[i] = div_const_exact(dd);
    return vals[i];
}

/* The divisors of the step: dx, (real)sqrt(2.0)*dx (physics.h:45-48), A;
 * and whether the fused walk may take its LAND form (n_ok: every LAND
 * roughness in range, from the static upload). */
static void set_div_flags(FcaDev *d, int n_ok, int no_land_form)
{
    const real dx = d->phys.dx, ddiag = (real)sqrt(2.0) * dx;
    d->fdiv[0] = div_const_exact_cached((double)dx);
    d->fdiv[1] = d->phys.K == 8 ? div_const_exact_cached((double)ddiag) : 0;
    d->fdiv[2] = div_const_exact_cached((double)d->phys.A);
    d->land_form = !no_land_form && n_ok && fca_land_phys_ok(&d->phys);
#ifndef FCA_DOUBLE
    d->land_form = d->land_form && d->fdiv[0] && d->fdiv[2] && (d->phys.K != 8 || d->fdiv[1]);
#endif
}

/* Default strip height: about three tiles per resident wavefront slot
 * (4096 at this kernel's occupancy), so no launch ends in a long tail,
 * between 16 rows and 96 (where the 5 halo rows a strip re-walks cost 5%).
 * A 16384^2 grid gets 96 rows, 4096^2 gets 23 (measured at 4096^2: 0.295
 * ms/step at 16 or 24 rows, 0.418 at 96). */
static int default_strip_rows(int nrows, int ncols, int cpl)
{
    const long groups = (ncols + FCA_TILE_COLS(cpl) - 1) / FCA_TILE_COLS(cpl);
    const long rows = (long)nrows * groups / (3 * 4096);
    return (int)(rows < 16 ? 16 : (rows > 96 ? 96 : rows));
}

/* The run's physical constants and the controller's parameters. */
static void set_run_consts(FcaDev *d, const FcaRunCfg *cfg)
{
    d->phys = cfg->phys;
    d->ctl.T_sim = (double)cfg->T_sim;
    d->ctl.out_every = (double)cfg->out_every;
    d->ctl.D_h = (double)cfg->D_h;
    d->ctl.phi = (double)cfg->phi;
    d->ctl.hyeto = cfg->hyeto;
    d->ctl.steps = cfg->steps;
    d->ctl.U = cfg->phys.U;
    d->ctl.n_force_days = cfg->n_force_days;
}

static int sim_alloc(FcaHipSim *s, const FcaGrid *g, const FcaRunCfg *cfg,
                     const FcaHipOpts *o, hipStream_t shared)
{
    FcaDev *d = &s->d;
    const size_t N = (size_t)g->nrows * (size_t)g->ncols, B = N * sizeof(real);
    d->nr = g->nrows;
    d->nc = g->ncols;
    d->K = cfg->phys.K;
    d->kernel = o ? o->kernel : FCA_HIP_FUSED;
    d->cols_per_lane = (o && o->cols_per_lane == 2) ? 2 : 1;
    d->strip_rows = (o && o->strip_rows > 0)
                        ? o->strip_rows
                        : default_strip_rows(g->nrows, g->ncols, d->cols_per_lane);
    d->rec_cap = (o && o->record_steps > 0) ? o->record_steps : 0;
    set_run_consts(d, cfg);
    s->no_land_form = o && (o->flags & FCA_HIP_NO_LAND_FORM);
    set_div_flags(d, 0, s->no_land_form);
    d->max_blocks = (o && o->max_blocks > 0) ? o->max_blocks : 0;
    d->tiles_per_wave = (o && o->tiles_per_wave > 1) ? o->tiles_per_wave : 1;
    fca_k_geometry(d);
    s->N = N;
    s->batch = (o && o->graph_batch > 0) ? o->graph_batch : 30;
    /* replays of glen launches; graph_batch 1 launches every step itself */
    s->glen = (s->batch / FCA_GRAPH_PERIOD) * FCA_GRAPH_PERIOD;
    if (shared) {
        s->st = shared;
        s->own_stream = 0;
    } else {
        HIP_OR(hipStreamCreate(&s->st), -2);
        s->own_stream = 1;
    }
    /* one balance slot per workgroup: the fused walk's, or one per block of
     * the split kernels' cell-parallel launches */
    d->nslots = d->kernel == FCA_HIP_FUSED ? d->nblocks : (int)((N + 255) / 256);
    d->nred = NRED * 5;
    {
        ArenaReq req[24];
        int nq = 0;
#define FCA_REQ(p_, bytes_) (req[nq].ptr = (void **)&(p_), req[nq].bytes = (bytes_), nq++)
        FCA_REQ(d->z, B);
        FCA_REQ(d->n, B);
        FCA_REQ(d->C, B);
        FCA_REQ(d->Rday, B);
        FCA_REQ(d->cr, B);
        FCA_REQ(d->h[0], B);
        FCA_REQ(d->h[1], B);
        FCA_REQ(d->hmax, B);
        FCA_REQ(d->tmax, B);
        FCA_REQ(d->mask, N);
        FCA_REQ(d->ctrl, 2 * sizeof(FcaDevCtrl));
        FCA_REQ(d->cand, 3 * sizeof(double));
        FCA_REQ(d->slots, (size_t)d->nslots * sizeof(FcaAccSlot));
        FCA_REQ(d->rec, (size_t)(d->rec_cap > 0 ? d->rec_cap : 1) * sizeof(FcaHipStepRec));
        FCA_REQ(d->red, (size_t)d->nred * sizeof(double));
        if (d->kernel == FCA_HIP_FUSED) {
            FCA_REQ(d->tflags, (size_t)2 * d->ntiles);
            FCA_REQ(d->tnorain, (size_t)d->ntiles);
        }
        if (d->kernel == FCA_HIP_SPLIT) {
            size_t nqx = (size_t)d->nr * (size_t)(d->nc - 1), nqy = (size_t)(d->nr - 1) * d->nc;
            size_t nqd = (size_t)(d->nr - 1) * (size_t)(d->nc - 1);
            FCA_REQ(d->qx, (nqx ? nqx : 1) * sizeof(real));
            FCA_REQ(d->qy, (nqy ? nqy : 1) * sizeof(real));
            FCA_REQ(d->qd1, (nqd ? nqd : 1) * sizeof(real));
            FCA_REQ(d->qd2, (nqd ? nqd : 1) * sizeof(real));
            FCA_REQ(d->qo, 4 * B);
            FCA_REQ(d->s, B);
            FCA_REQ(d->sv, 256);
        }
#undef FCA_REQ
        if (dev_arena(d, req, nq) != 0) {
            return -2;
        }
    }
    if (d->kernel == FCA_HIP_FUSED) {
        s->skip_wanted = !(o && (o->flags & FCA_HIP_NO_SKIP));
        d->skip = 0;
    }
    if (d->kernel == FCA_HIP_SPLIT) {
        HIP_OR(hipMemset(d->qo, 0, 4 * B), -2);
    }
    /* small read-backs go to ordinary memory: pinning costs more than it
     * saves at these sizes, and a batch opens hundreds of sessions */
    s->h_ctrl = (FcaDevCtrl *)malloc(2 * sizeof(FcaDevCtrl));
    s->h_red = (double *)malloc((size_t)d->nred * sizeof(double));
    if (!s->h_ctrl || !s->h_red) {
        return fail(-2, "out of host memory");
    }
    return 0;
}

static int sim_upload_static(FcaHipSim *s, const FcaGrid *g)
{
    FcaDev *d = &s->d;
    const size_t B = s->N * sizeof(real);
    double t0 = now_s();
    HIP_OR(hipMemcpy(d->z, g->z, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->n, g->n, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->C, g->C, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->mask, g->mask, s->N, hipMemcpyHostToDevice), -2);
    s->t_h2d += now_s() - t0;
    s->n_ok = fca_land_n_ok(g->n, g->mask, s->N);
    set_div_flags(d, s->n_ok, s->no_land_form);
    return 0;
}

static int sim_start(FcaHipSim *s, const FcaAcc *acc);

/* Uploads the per-run state (h, hmax, tmax, Rday), then sim_start. */
static int sim_upload_state(FcaHipSim *s, const FcaGrid *g, const FcaAcc *acc)
{
    FcaDev *d = &s->d;
    const size_t B = s->N * sizeof(real);
    double t0 = now_s();
    HIP_OR(hipMemcpy(d->h[0], g->h, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->h[1], d->h[0], B, hipMemcpyDeviceToDevice), -2);
    HIP_OR(hipMemcpy(d->hmax, g->hmax, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->tmax, g->tmax, B, hipMemcpyHostToDevice), -2);
    HIP_OR(hipMemcpy(d->Rday, g->Rday, B, hipMemcpyHostToDevice), -2);
    s->t_h2d += now_s() - t0;
    return sim_start(s, acc);
}

/* With the fields on the device: resets the controller from acc, clears the
 * slots and tile flags, computes cr = C*Rday, the rain-free tiles and the
 * first time-step candidate. */
static int sim_start(FcaHipSim *s, const FcaAcc *acc)
{
    FcaDev *d = &s->d;
    FcaDevCtrl c;
    memset(&c, 0, sizeof(c));
    c.t = acc->t;
    c.step = acc->step;
    c.force_day = acc->force_day;
    c.state = FCA_CTL_RUN;
    c.hbuf = 0;
    c.dt_base = (double)acc->dt_base;
    HIP_OR(hipMemcpyAsync(d->ctrl, &c, sizeof(c), hipMemcpyHostToDevice, s->st), -2);
    if (fca_k_zero(d->slots, (size_t)d->nslots * sizeof(FcaAccSlot), s->st) != 0 ||
        (d->tflags && fca_k_zero(d->tflags, (size_t)2 * d->ntiles, s->st) != 0)) {
        return fail(-2, "clearing the slots failed"); /* no tile is known quiet yet */
    }
    if (d->tflags && s->skip_wanted) {
        /* skipping can only help where a LAND tile gets no rain or the
         * hyetograph can fall to zero; otherwise the kernel without the
         * quiet-tile bookkeeping runs (the no-alloc path does not look for
         * rain-free tiles: either kernel gives the same results) */
        int can = s->cfg->hyeto != FCA_HYETO_UNIFORM || s->cfg->n_force_days > 0;
        if (!can && !s->noalloc) {
            uint8_t *tn = (uint8_t *)malloc((size_t)d->ntiles);
            int t;
            if (!tn) {
                return fail(-2, "out of host memory");
            }
            if (fca_k_cr(d, s->st) != 0 || fca_k_tile_norain(d, s->st) != 0 ||
                hipMemcpyAsync(tn, d->tnorain, (size_t)d->ntiles, hipMemcpyDeviceToHost,
                               s->st) != hipSuccess ||
                hipStreamSynchronize(s->st) != hipSuccess) {
                free(tn);
                return fail(-2, "rain-free tile map failed");
            }
            for (t = 0; t < d->ntiles && !can; t++) {
                can = (tn[t] & 3) == 3;
            }
            free(tn);
        }
        d->skip = can;
    }
    if (fca_k_cr(d, s->st) != 0 || fca_k_tile_norain(d, s->st) != 0 ||
        fca_k_cand_init(d, 0, 0, s->st) != 0 ||
        fca_k_cand_reset(d, 1, s->st) != 0 || fca_k_cand_reset(d, 2, s->st) != 0) {
        return fail(-2, "kernel launch failed during upload");
    }
    HIP_OR(hipStreamSynchronize(s->st), -2);
    s->k = 0;
    s->acc0 = *acc;
    return 0;
}

static int check_cfg(const FcaGrid *g, const FcaRunCfg *cfg, const FcaHipOpts *o)
{
    if (!g || !cfg || g->nrows <= 0 || g->ncols <= 0) {
        return fail(-1, "empty grid");
    }
    if ((long long)g->nrows * (long long)g->ncols > (1ll << 31) - 1) {
        return fail(-1, "grid of %d x %d cells exceeds 2^31-1 cells", g->nrows, g->ncols);
    }
    if (cfg->phys.K != 4 && cfg->phys.K != 8) {
        return fail(-1, "K must be 4 or 8, not %d", cfg->phys.K);
    }
    if (!(cfg->phys.alpha * cfg->phys.dx > 0)) {
        return fail(-1, "alpha*dx must be positive");
    }
    if (o && o->kernel != FCA_HIP_FUSED && o->kernel != FCA_HIP_SPLIT) {
        return fail(-1, "unknown kernel family %d", o->kernel);
    }
    if (o && o->strip_rows < 0) {
        return fail(-1, "strip_rows must be >= 0");
    }
    if (o && (o->cols_per_lane < 0 || o->cols_per_lane > 2)) {
        return fail(-1, "cols_per_lane must be 0, 1 or 2, not %d", o->cols_per_lane);
    }
    if (o && o->cols_per_lane == 2 && !FCA_CPL2_BUILT) {
        return fail(-1, "cols_per_lane 2 is built for the float library only");
    }
    if (cfg->n_force_days < 0 || cfg->n_force_days > FCA_MAX_FORCE_DAYS) {
        return fail(-1, "n_force_days out of range");
    }
    return 0;
}

static FcaHipSim *sim_create(const FcaGrid *g, const FcaRunCfg *cfg, const FcaAcc *acc,
                             const FcaHipOpts *opts, hipStream_t shared)
{
    FcaHipSim *s;
    g_err[0] = 0;
    if (check_cfg(g, cfg, opts) != 0 || !acc) {
        if (!acc) {
            fail(-1, "acc is NULL");
        }
        return NULL;
    }
    s = (FcaHipSim *)calloc(1, sizeof(*s));
    if (!s) {
        fail(-2, "out of host memory");
        return NULL;
    }
    s->cfg = (FcaRunCfg *)malloc(sizeof(FcaRunCfg));
    if (!s->cfg) {
        free(s);
        fail(-2, "out of host memory");
        return NULL;
    }
    memcpy(s->cfg, cfg, sizeof(FcaRunCfg));
    s->faces_per_step = count_faces(g, &cfg->phys);
    if (sim_alloc(s, g, cfg, opts, shared) != 0 || sim_upload_static(s, g) != 0 ||
        sim_upload_state(s, g, acc) != 0) {
        char keep[512];
        memcpy(keep, g_err, sizeof(keep));
        fca_hip_sim_close(s);
        memcpy(g_err, keep, sizeof(keep));
        return NULL;
    }
    return s;
}

/* ---- crop scenarios built on the device ---------------------------------- */

struct FcaHipBase {
    int nr, nc;
    real *z;
    uint8_t *mask, *cls;
    uint8_t *h_mask;    /* host copy, for the face counts of a window */
    unsigned classes;   /* bit k: class k occurs somewhere in the base */
};

FcaHipBase *fca_hip_base_open(const FcaGrid *g, const uint8_t *cls)
{
    FcaHipBase *b;
    size_t N, i;
    g_err[0] = 0;
    if (!g || !cls || !g->z || !g->mask || g->nrows <= 0 || g->ncols <= 0) {
        fail(-1, "bad arguments");
        return NULL;
    }
    N = (size_t)g->nrows * (size_t)g->ncols;
    b = (FcaHipBase *)calloc(1, sizeof(*b));
    if (!b) {
        fail(-2, "out of host memory");
        return NULL;
    }
    b->nr = g->nrows;
    b->nc = g->ncols;
    for (i = 0; i < N; i++) {
        if (cls[i] >= FCA_HIP_MAX_CLASSES) {
            fail(-1, "class %d at cell %zu is not below %d", cls[i], i, FCA_HIP_MAX_CLASSES);
            free(b);
            return NULL;
        }
        b->classes |= 1u << cls[i];
    }
    b->h_mask = (uint8_t *)malloc(N);
    if (!b->h_mask || hipMalloc((void **)&b->z, N * sizeof(real)) != hipSuccess ||
        hipMalloc((void **)&b->mask, N) != hipSuccess ||
        hipMalloc((void **)&b->cls, N) != hipSuccess ||
        hipMemcpy(b->z, g->z, N * sizeof(real), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->mask, g->mask, N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->cls, cls, N, hipMemcpyHostToDevice) != hipSuccess) {
        fail(-2, "base grid upload failed");
        fca_hip_base_close(b);
        return NULL;
    }
    memcpy(b->h_mask, g->mask, N);
    return b;
}

void fca_hip_base_close(FcaHipBase *b)
{
    if (!b) {
        return;
    }
    if (b->z) {
        hipFree(b->z);
    }
    if (b->mask) {
        hipFree(b->mask);
    }
    if (b->cls) {
        hipFree(b->cls);
    }
    free(b->h_mask);
    free(b);
}

/* count_faces for a window of the base (its mask read from the host copy). */
static long count_faces_window(const FcaHipBase *b, const FcaHipCrop *w, const FcaPhys *p)
{
    long nr = w->nrows, nc = w->ncols, f, r, c;
    f = nr * (nc - 1) + (nr - 1) * nc + (p->K == 8 ? 2 * (nr - 1) * (nc - 1) : 0);
    if (p->boundary == FCA_BND_OPEN) {
        /* open faces of LAND cells on the window's edges, as count_faces */
        for (r = 0; r < nr; r++) {
            const uint8_t *row = b->h_mask + (size_t)(w->r0 + r) * (size_t)b->nc + w->c0;
            const long er = (r == 0) + (r == nr - 1);
            if (er) {
                for (c = 0; c < nc; c++) {
                    if (row[c] == FCA_LAND) {
                        f += er + (c == 0) + (c == nc - 1);
                    }
                }
            } else {
                if (row[0] == FCA_LAND) {
                    f += 1 + (nc == 1);
                }
                if (nc > 1 && row[nc - 1] == FCA_LAND) {
                    f += 1;
                }
            }
        }
    }
    return f;
}

static FcaHipSim *sim_create_crop(const FcaHipBase *b, const FcaHipCrop *w,
                                  const FcaRunCfg *cfg, const FcaHipOpts *opts,
                                  hipStream_t shared)
{
    FcaHipSim *s;
    FcaGrid gh;
    FcaAcc acc;
    unsigned k;
    g_err[0] = 0;
    memset(&gh, 0, sizeof(gh));
    gh.nrows = w->nrows;
    gh.ncols = w->ncols;
    if (check_cfg(&gh, cfg, opts) != 0) {
        return NULL;
    }
    if (w->r0 < 0 || w->c0 < 0 || w->r0 + w->nrows > b->nr || w->c0 + w->ncols > b->nc) {
        fail(-1, "crop window outside the base grid");
        return NULL;
    }
    s = (FcaHipSim *)calloc(1, sizeof(*s));
    if (!s) {
        fail(-2, "out of host memory");
        return NULL;
    }
    s->cfg = (FcaRunCfg *)malloc(sizeof(FcaRunCfg));
    if (!s->cfg) {
        free(s);
        fail(-2, "out of host memory");
        return NULL;
    }
    memcpy(s->cfg, cfg, sizeof(FcaRunCfg));
    s->faces_per_step = count_faces_window(b, w, &cfg->phys);
    /* dry start: fca_acc_init of h = 0 (model.c:102-109) */
    memset(&acc, 0, sizeof(acc));
    acc.force_day = -1;
    if (sim_alloc(s, &gh, cfg, opts, shared) != 0 ||
        fca_k_crop(&s->d, b->z, b->mask, b->cls, b->nc, w, s->st) != 0) {
        char keep[512];
        memcpy(keep, g_err, sizeof(keep));
        fca_hip_sim_close(s);
        memcpy(g_err, keep, sizeof(keep));
        if (!g_err[0]) {
            fail(-2, "crop launch failed");
        }
        return NULL;
    }
    /* the roughness of every class that occurs in the base, hence of every
     * LAND cell of the window, is in the LAND form's range (fca_land_n_ok) */
    s->n_ok = 1;
    for (k = 0; k < FCA_HIP_MAX_CLASSES; k++) {
        if ((b->classes >> k) & 1u) {
            s->n_ok = s->n_ok && w->n_class[k] >= (real)1e-12 && w->n_class[k] <= (real)0x1p100;
        }
    }
    set_div_flags(&s->d, s->n_ok, s->no_land_form);
    if (sim_start(s, &acc) != 0) {
        char keep[512];
        memcpy(keep, g_err, sizeof(keep));
        fca_hip_sim_close(s);
        memcpy(g_err, keep, sizeof(keep));
        return NULL;
    }
    return s;
}

FcaHipSim *fca_hip_sim_open(const FcaGrid *g, const FcaRunCfg *cfg, const FcaAcc *acc,
                            const FcaHipOpts *opts)
{
    return sim_create(g, cfg, acc, opts, NULL);
}

void *fca_hip_sim_stream(FcaHipSim *s)
{
    return s ? (void *)s->st : NULL;
}

int fca_hip_sim_advance(FcaHipSim *s, int nsteps)
{
    int left;
    if (!s || nsteps < 0) {
        return fail(-1, "bad arguments");
    }
    left = nsteps;
    while (left > 0) {
        if (s->glen > 0 && s->k % FCA_GRAPH_PERIOD == 0 && left >= s->glen) {
            if (!s->gexec && capture_steps(s->st, &s->gexec, s->glen, launch_sim_step, s) != 0) {
                return -2;
            }
            if (hipGraphLaunch(s->gexec, s->st) != hipSuccess) {
                return fail(-2, "graph launch failed");
            }
            s->k += s->glen;
            left -= s->glen;
            continue;
        }
        if (fca_k_step(&s->d, s->k, s->st) != 0) {
            return fail(-2, "step launch failed: %s", hipGetErrorString(hipGetLastError()));
        }
        s->k++;
        left--;
    }
    return nsteps;
}

int fca_hip_sim_sync(FcaHipSim *s)
{
    HIP_OR(hipStreamSynchronize(s->st), -2);
    return 0;
}

/* Current controller record (the one the next launch reads). */
static int read_ctrl(FcaHipSim *s, FcaDevCtrl *c)
{
    HIP_OR(hipMemcpyAsync(s->h_ctrl, &s->d.ctrl[s->k & 1], sizeof(FcaDevCtrl),
                          hipMemcpyDeviceToHost, s->st), -2);
    HIP_OR(hipStreamSynchronize(s->st), -2);
    *c = *s->h_ctrl;
    return 0;
}

static int write_ctrl(FcaHipSim *s, const FcaDevCtrl *c)
{
    *s->h_ctrl = *c;
    HIP_OR(hipMemcpyAsync(&s->d.ctrl[s->k & 1], s->h_ctrl, sizeof(FcaDevCtrl),
                          hipMemcpyHostToDevice, s->st), -2);
    HIP_OR(hipStreamSynchronize(s->st), -2);
    return 0;
}

/* Completes hmax and tmax on the device for the state the record c
 * describes (the fused walk stores them only when a cell stops rising);
 * nothing to do before the first step of the session. Enqueued on the
 * session's stream. */
static int sim_finalize(FcaHipSim *s, const FcaDevCtrl *c)
{
    if (c->armed && fca_k_hmax_finalize(&s->d, c->hbuf, c->tmax_last, s->st) != 0) {
        return fail(-2, "hmax finalise launch failed");
    }
    return 0;
}

/* Loads forcing day `day` (float32 raster) into Rday, scaled by lambda
 * (model.c:338-358, grid.c:239-254). Returns 0, or -1 if unreadable. */
static int reload_day(FcaHipSim *s, int day)
{
    FcaDev *d = &s->d;
    const FcaRunCfg *cfg = s->cfg;
    FILE *fp;
    float *tmp;
    size_t i, got;
    double t0 = now_s();
    if (!s->h_stage &&
        hipHostMalloc((void **)&s->h_stage, s->N * sizeof(real), hipHostMallocDefault) !=
            hipSuccess) {
        return fail(-2, "pinned staging allocation failed");
    }
    fp = fopen(cfg->force_files[day], "rb");
    if (!fp) {
        return -1;
    }
    tmp = (float *)malloc(s->N * sizeof(float));
    if (!tmp) {
        fclose(fp);
        return fail(-2, "out of host memory");
    }
    got = fread(tmp, sizeof(float), s->N, fp);
    fclose(fp);
    if (got != s->N) {
        free(tmp);
        return -1;
    }
    for (i = 0; i < s->N; i++) {
        s->h_stage[i] = (real)tmp[i];
        if (cfg->lambda != (real)1) {
            s->h_stage[i] *= cfg->lambda;
        }
    }
    free(tmp);
    s->t_io += now_s() - t0;
    t0 = now_s();
    HIP_OR(hipMemcpyAsync(d->Rday, s->h_stage, s->N * sizeof(real), hipMemcpyHostToDevice,
                          s->st), -2);
    if (fca_k_cr(d, s->st) != 0 || fca_k_tile_norain(d, s->st) != 0) {
        return fail(-2, "cr launch failed");
    }
    HIP_OR(hipStreamSynchronize(s->st), -2);
    s->t_h2d += now_s() - t0;
    return 0;
}

typedef struct RunHooks {
    void (*poll)(void *);
    void *pctx;
    int poll_every;
} RunHooks;

/* fca_run's loop (model.c:607-643): launches in batches, reading the
 * controller record back between batches to stop at T_sim, to reload the
 * forcing day, or to stop when a fixed-step run cannot advance. */
static int sim_run_hooked(FcaHipSim *s, const RunHooks *hk)
{
    const FcaRunCfg *cfg = s->cfg;
    FcaDevCtrl c;
    int last_poll;
    if (read_ctrl(s, &c) != 0) {
        return -2;
    }
    last_poll = c.step;
    for (;;) {
        int n = s->batch, before = c.step;
        if (c.state == FCA_CTL_DONE ||
            (cfg->steps <= 0 && c.t >= (double)cfg->T_sim - 1e-15)) {
            break;
        }
        if (c.state == FCA_CTL_PAUSE) {
            int rc = reload_day(s, c.want_day);
            if (rc < -1) {
                return rc;
            }
            if (rc == -1) { /* missing day file: the run ends at T_sim */
                c.t = (double)cfg->T_sim;
                c.state = FCA_CTL_DONE;
            } else {
                c.force_day = c.want_day;
                c.state = FCA_CTL_RUN;
            }
            if (write_ctrl(s, &c) != 0) {
                return -2;
            }
            continue;
        }
        if (cfg->steps > 0) {
            if (c.step >= cfg->steps) {
                break;
            }
            if (n > cfg->steps - c.step) {
                n = cfg->steps - c.step;
            }
        }
        if (hk && hk->poll && hk->poll_every > 0 && n > hk->poll_every - c.step % hk->poll_every) {
            /* end the batch at the next multiple of poll_every, so that poll
             * runs after every poll_every steps as the reference's loop
             * calls it (kernels.cu:967-968) */
            n = hk->poll_every - c.step % hk->poll_every;
        }
        if (fca_hip_sim_advance(s, n) < 0 || read_ctrl(s, &c) != 0) {
            return -2;
        }
        if (hk && hk->poll && hk->poll_every > 0 &&
            c.step / hk->poll_every != last_poll / hk->poll_every) {
            hk->poll(hk->pctx);
            last_poll = c.step;
        }
        if (cfg->progress_every > 0 &&
            c.step / cfg->progress_every != before / cfg->progress_every) {
            printf("progress step=%d t=%.1f dt_base=%.4g\n", c.step, c.t, c.dt_base);
            fflush(stdout);
        }
        if (c.state == FCA_CTL_RUN && c.step == before && cfg->steps > 0) {
            break; /* no step possible (the reference would spin here) */
        }
    }
    return 0;
}

int fca_hip_sim_run(FcaHipSim *s)
{
    if (!s) {
        return fail(-1, "NULL session");
    }
    return sim_run_hooked(s, NULL);
}

/* Sums the per-workgroup balance slots in a fixed order (read through a
 * buffer on the stack: nothing is allocated). */
static int read_slots(FcaHipSim *s, double sums[4], long long *cap)
{
    FcaDev *d = &s->d;
    FcaAccSlot buf[1024];
    int i0, i, n;
    sums[0] = sums[1] = sums[2] = sums[3] = 0.0;
    *cap = 0;
    for (i0 = 0; i0 < d->nslots; i0 += n) {
        n = d->nslots - i0 < 1024 ? d->nslots - i0 : 1024;
        HIP_OR(hipMemcpyAsync(buf, d->slots + i0, (size_t)n * sizeof(FcaAccSlot),
                              hipMemcpyDeviceToHost, s->st), -2);
        HIP_OR(hipStreamSynchronize(s->st), -2);
        for (i = 0; i < n; i++) {
            sums[0] += buf[i].v_rain;
            sums[1] += buf[i].clamp;
            sums[2] += buf[i].v_sea;
            sums[3] += buf[i].v_bnd;
            *cap += buf[i].cap;
        }
    }
    return 0;
}

static int device_volume(FcaHipSim *s, int hbuf, double *V)
{
    int nb = fca_k_volume(&s->d, hbuf, s->st), i;
    if (nb < 0) {
        return fail(-2, "volume launch failed");
    }
    HIP_OR(hipMemcpyAsync(s->h_red, s->d.red, (size_t)nb * sizeof(double),
                          hipMemcpyDeviceToHost, s->st), -2);
    HIP_OR(hipStreamSynchronize(s->st), -2);
    *V = 0.0;
    for (i = 0; i < nb; i++) {
        *V += s->h_red[i];
    }
    return 0;
}

static int sim_acc(FcaHipSim *s, const FcaDevCtrl *c, FcaAcc *acc)
{
    double sums[4] = {0, 0, 0, 0}, V = 0.0;
    long long cap = 0;
    if (read_slots(s, sums, &cap) != 0 || device_volume(s, c->hbuf, &V) != 0) {
        return -2;
    }
    *acc = s->acc0;
    acc->V_rain += sums[0];
    acc->clamp_err += sums[1];
    acc->V_sea += sums[2];
    acc->V_bnd += sums[3];
    acc->cap_active += (long)cap;
    acc->cap_total += s->faces_per_step * (long)(c->step - s->acc0.step);
    acc->dt_min_hits += (long)c->dt_min_hits;
    acc->t = c->t;
    acc->step = c->step;
    acc->dt_base = (real)c->dt_base;
    acc->force_day = c->force_day;
    acc->V = V;
    return 0;
}

int fca_hip_sim_read(FcaHipSim *s, real *h, real *hmax, real *tmax, FcaAcc *acc)
{
    FcaDevCtrl c;
    const size_t B = s->N * sizeof(real);
    if (!s) {
        return fail(-1, "NULL session");
    }
    if (read_ctrl(s, &c) != 0 || sim_finalize(s, &c) != 0) {
        return -2;
    }
    HIP_OR(hipStreamSynchronize(s->st), -2);
    if (h) {
        HIP_OR(hipMemcpy(h, s->d.h[c.hbuf], B, hipMemcpyDeviceToHost), -2);
    }
    if (hmax) {
        HIP_OR(hipMemcpy(hmax, s->d.hmax, B, hipMemcpyDeviceToHost), -2);
    }
    if (tmax) {
        HIP_OR(hipMemcpy(tmax, s->d.tmax, B, hipMemcpyDeviceToHost), -2);
    }
    if (acc && sim_acc(s, &c, acc) != 0) {
        return -2;
    }
    return 0;
}

int fca_hip_sim_steplog(FcaHipSim *s, FcaHipStepRec *out, int cap)
{
    FcaDevCtrl c;
    int n;
    if (!s || (cap > 0 && !out)) {
        return fail(-1, "bad arguments");
    }
    if (read_ctrl(s, &c) != 0) {
        return -2;
    }
    n = c.nrec < s->d.rec_cap ? c.nrec : s->d.rec_cap;
    if (n > cap) {
        n = cap;
    }
    if (n > 0) {
        HIP_OR(hipMemcpy(out, s->d.rec, (size_t)n * sizeof(FcaHipStepRec),
                         hipMemcpyDeviceToHost), -2);
    }
    return n;
}

/* ---- entry points with a reference counterpart ---------------------------- */

int fca_hip_run(FcaGrid *g, FcaRunCfg *cfg, FcaAcc *acc, FcaTimers *tm)
{
    FcaHipSim *s;
    double t0;
    int rc;
    if (!cfg || !acc) {
        return fail(-1, "NULL cfg or acc");
    }
    if (!cfg->reduce_impl[0]) {
        snprintf(cfg->reduce_impl, sizeof(cfg->reduce_impl), "wave");
    }
    t0 = now_s();
    s = fca_hip_sim_open(g, cfg, acc, NULL);
    if (!s) {
        return g_err[0] && strstr(g_err, "failed") ? -2 : -1;
    }
    if (tm) {
        tm->t_init += now_s() - t0 - s->t_h2d;
        tm->t_h2d = (tm->t_h2d > 0 ? tm->t_h2d : 0) + s->t_h2d;
    }
    s->t_h2d = 0;
    t0 = now_s();
    rc = fca_hip_sim_run(s);
    if (rc == 0) {
        rc = fca_hip_sim_sync(s);
    }
    if (tm) {
        tm->wall_total += now_s() - t0;
        tm->t_io += s->t_io;
        tm->t_h2d += s->t_h2d;
        tm->t_flux = tm->t_limiter = tm->t_update = tm->t_dt = -1;
    }
    if (rc == 0) {
        t0 = now_s();
        rc = fca_hip_sim_read(s, g->h, g->hmax, g->tmax, acc);
        if (tm) {
            tm->t_d2h = (tm->t_d2h > 0 ? tm->t_d2h : 0) + (now_s() - t0);
        }
    }
    fca_hip_sim_close(s);
    return rc;
}

/* fca_hip_run on caller-owned device fields with no allocation: the
 * session lives on this stack frame, its device buffers are carved from the
 * library's per-device workspace (fca_k_noalloc_scratch), controller
 * read-backs land in stack variables, and the steps are launched one by one
 * (instantiating a graph would allocate). */
int fca_hip_run_noalloc(FcaGrid *g, FcaRunCfg *cfg, FcaAcc *acc, FcaTimers *tm, void *stream)
{
    FcaHipSim s;
    FcaDev *d = &s.d;
    FcaDevCtrl hctrl[2], c;
    double hred[FCA_NOALLOC_NRED];
    unsigned long long scan[2], *dscan;
    char *base = NULL;
    size_t bytes = 0, off = 0, N, B;
    double t0;
    int rc;

    g_err[0] = 0;
    if (!g || !cfg || !acc) {
        return fail(-1, "NULL grid, cfg or acc");
    }
    if (check_cfg(g, cfg, NULL) != 0) {
        return -1;
    }
    N = (size_t)g->nrows * (size_t)g->ncols;
    if (N > (size_t)FCA_HIP_NOALLOC_MAX_CELLS) {
        return fail(-1, "grid of %d x %d cells exceeds the %d cells fca_hip_run_noalloc accepts",
                    g->nrows, g->ncols, FCA_HIP_NOALLOC_MAX_CELLS);
    }
    if (!g->z || !g->h || !g->n || !g->C || !g->mask || !g->Rday || !g->hmax || !g->tmax) {
        return fail(-1, "a field of the grid is NULL");
    }
    if (cfg->n_force_days > 0) {
        return fail(-1, "multi-day forcing needs host staging; fca_hip_run_noalloc rejects it");
    }
    if (!cfg->reduce_impl[0]) {
        snprintf(cfg->reduce_impl, sizeof(cfg->reduce_impl), "wave");
    }
    t0 = now_s();
    if (fca_k_noalloc_scratch((void **)&base, &bytes) != 0 || !base) {
        return fail(-2, "workspace address unavailable: %s", hipGetErrorString(hipGetLastError()));
    }
    memset(&s, 0, sizeof(s));
    s.noalloc = 1;
    s.cfg = cfg;
    s.h_ctrl = hctrl;
    s.h_red = hred;
    s.st = (hipStream_t)stream;
    s.N = N;
    s.batch = 30;
    s.glen = 0;
    s.skip_wanted = 1;
    B = N * sizeof(real);
    d->nr = g->nrows;
    d->nc = g->ncols;
    d->K = cfg->phys.K;
    d->kernel = FCA_HIP_FUSED;
    d->cols_per_lane = 1;
    d->strip_rows = default_strip_rows(g->nrows, g->ncols, 1);
    set_run_consts(d, cfg);
    fca_k_geometry(d);
    d->nslots = d->nblocks;
    d->nred = FCA_NOALLOC_NRED;
#define FCA_CARVE(p_, n_) ((p_) = (void *)(base + off), off += ((size_t)(n_) + 255) & ~(size_t)255)
    {
        void *p;
        FCA_CARVE(p, B);
        d->h[1] = (real *)p;
        FCA_CARVE(p, B);
        d->cr = (real *)p;
        FCA_CARVE(p, 2 * sizeof(FcaDevCtrl));
        d->ctrl = (FcaDevCtrl *)p;
        FCA_CARVE(d->cand, 3 * sizeof(double));
        FCA_CARVE(p, (size_t)d->nslots * sizeof(FcaAccSlot));
        d->slots = (FcaAccSlot *)p;
        FCA_CARVE(p, sizeof(FcaHipStepRec));
        d->rec = (FcaHipStepRec *)p;
        FCA_CARVE(p, (size_t)d->nred * sizeof(double));
        d->red = (double *)p;
        FCA_CARVE(p, (size_t)2 * d->ntiles);
        d->tflags = (uint8_t *)p;
        FCA_CARVE(p, (size_t)d->ntiles);
        d->tnorain = (uint8_t *)p;
        FCA_CARVE(p, 2 * sizeof(unsigned long long));
        dscan = (unsigned long long *)p;
    }
#undef FCA_CARVE
    if (off > bytes) {
        return fail(-1, "a %d x %d grid needs %zu B of workspace, more than the %zu B reserved",
                    g->nrows, g->ncols, off, bytes);
    }
    d->z = g->z, d->n = g->n, d->C = g->C, d->Rday = g->Rday, d->mask = g->mask;
    d->h[0] = g->h, d->hmax = g->hmax, d->tmax = g->tmax;
    /* what fca_hip_run learns from host fields: the LAND roughness range and
     * the open faces counted into cap_total */
    if (fca_k_grid_scan(d, dscan, s.st) != 0) {
        return fail(-2, "grid scan launch failed");
    }
    HIP_OR(hipMemcpyAsync(scan, dscan, sizeof(scan), hipMemcpyDeviceToHost, s.st), -2);
    HIP_OR(hipMemcpyAsync(d->h[1], d->h[0], B, hipMemcpyDeviceToDevice, s.st), -2);
    HIP_OR(hipStreamSynchronize(s.st), -2);
    s.n_ok = scan[0] == 0;
    set_div_flags(d, s.n_ok, 0);
    s.faces_per_step = count_inner_faces(g->nrows, g->ncols, &cfg->phys) + (long)scan[1];
    if (sim_start(&s, acc) != 0) {
        return -2;
    }
    if (tm) {
        tm->t_init += now_s() - t0;
        tm->t_h2d = tm->t_h2d > 0 ? tm->t_h2d : 0;
    }
    t0 = now_s();
    rc = sim_run_hooked(&s, NULL);
    if (rc == 0) {
        rc = fca_hip_sim_sync(&s);
    }
    if (tm) {
        tm->wall_total += now_s() - t0;
        tm->t_io += s.t_io;
        tm->t_flux = tm->t_limiter = tm->t_update = tm->t_dt = -1;
    }
    if (rc == 0) {
        t0 = now_s();
        rc = read_ctrl(&s, &c);
        if (rc == 0) {
            rc = sim_finalize(&s, &c);
        }
        if (rc == 0 && c.hbuf == 1) { /* the last step wrote the workspace's buffer */
            if (hipMemcpyAsync(g->h, d->h[1], B, hipMemcpyDeviceToDevice, s.st) != hipSuccess ||
                hipStreamSynchronize(s.st) != hipSuccess) {
                rc = fail(-2, "depth copy-back failed");
            }
        }
        if (rc == 0) {
            rc = sim_acc(&s, &c, acc);
        }
        if (tm) {
            tm->t_d2h = (tm->t_d2h > 0 ? tm->t_d2h : 0) + (now_s() - t0);
        }
    }
    return rc;
}

void fca_hip_device_name(char *name, size_t n)
{
    char buf[256];
    char arch[64];
    hipDeviceProp_t prop;
    int dev = 0;
    if (!name || n == 0) {
        return;
    }
    buf[0] = arch[0] = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetName(buf, (int)sizeof(buf), dev) != hipSuccess) {
        buf[0] = 0;
    }
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        snprintf(arch, sizeof(arch), "%.63s", prop.gcnArchName);
        if (!buf[0]) {
            snprintf(buf, sizeof(buf), "%.255s", prop.name);
        }
    }
    if (!buf[0]) {
        snprintf(buf, sizeof(buf), "AMD GPU");
    }
    if (arch[0]) {
        snprintf(name, n, "%.255s (%.63s)", buf, arch);
    } else {
        snprintf(name, n, "%s", buf);
    }
}

int fca_hip_select_device(int local_rank)
{
    int n = 0, dev;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        return fail(-1, "no HIP device");
    }
    dev = (local_rank < 0 ? -local_rank : local_rank) % n;
    if (hipSetDevice(dev) != hipSuccess) {
        return fail(-1, "hipSetDevice(%d) failed", dev);
    }
    return dev;
}

/* ---- ensemble session ------------------------------------------------------ */

#define ENS_MAXTHR 4

struct FcaHipEns {
    FcaHipSim *s;
    int nthr;
    double thr[ENS_MAXTHR];
    uint32_t *counts; /* device, nthr * N */
};

FcaHipEns *fca_hip_ens_open(const FcaGrid *g, FcaRunCfg *cfg, int nthr, const double *thr)
{
    FcaHipEns *e;
    FcaAcc acc;
    int j;
    g_err[0] = 0;
    if (nthr < 1 || nthr > ENS_MAXTHR || !thr) {
        fail(-1, "fca_hip_ens_open: %d thresholds, at most %d", nthr, ENS_MAXTHR);
        return NULL;
    }
    if (check_cfg(g, cfg, NULL) != 0) {
        return NULL;
    }
    if (!cfg->reduce_impl[0]) {
        snprintf(cfg->reduce_impl, sizeof(cfg->reduce_impl), "wave");
    }
    e = (FcaHipEns *)calloc(1, sizeof(*e));
    if (!e) {
        fail(-2, "out of host memory");
        return NULL;
    }
    e->nthr = nthr;
    for (j = 0; j < nthr; j++) {
        e->thr[j] = thr[j];
    }
    memset(&acc, 0, sizeof(acc));
    acc.force_day = -1;
    e->s = fca_hip_sim_open(g, cfg, &acc, NULL);
    if (!e->s) {
        free(e);
        return NULL;
    }
    if (hipMalloc((void **)&e->counts, (size_t)nthr * e->s->N * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(e->counts, 0, (size_t)nthr * e->s->N * sizeof(uint32_t)) != hipSuccess) {
        fail(-2, "count raster allocation failed");
        fca_hip_ens_close(e);
        return NULL;
    }
    return e;
}

int fca_hip_ens_run(FcaHipEns *e, FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc,
                    FcaTimers *tm, long *flooded, double *max_depth, void (*poll)(void *),
                    void *pctx, int poll_every)
{
    FcaHipSim *s;
    RunHooks hk;
    FcaDevCtrl c;
    double t0;
    int nb, b, j, rc;
    if (!e || !g || !cfg || !acc) {
        return fail(-1, "bad arguments");
    }
    s = e->s;
    if (g->nrows != s->d.nr || g->ncols != s->d.nc || cfg->phys.K != s->d.K) {
        return fail(-1, "scenario grid or K differs from the session's");
    }
    /* per-scenario configuration: forcing shape, horizon, step count */
    memcpy(s->cfg, cfg, sizeof(FcaRunCfg));
    s->d.phys = cfg->phys;
    set_div_flags(&s->d, s->n_ok, s->no_land_form);
    if (s->gexec) { /* captured with the previous scenario's parameters */
        hipGraphExecDestroy(s->gexec);
        s->gexec = NULL;
    }
    s->d.ctl.T_sim = (double)cfg->T_sim;
    s->d.ctl.out_every = (double)cfg->out_every;
    s->d.ctl.D_h = (double)cfg->D_h;
    s->d.ctl.phi = (double)cfg->phi;
    s->d.ctl.hyeto = cfg->hyeto;
    s->d.ctl.steps = cfg->steps;
    s->d.ctl.U = cfg->phys.U;
    s->d.ctl.n_force_days = cfg->n_force_days;
    s->t_h2d = s->t_io = 0;
    if (sim_upload_state(s, g, acc) != 0) {
        return -2;
    }
    if (tm) {
        tm->t_h2d = (tm->t_h2d > 0 ? tm->t_h2d : 0) + s->t_h2d;
    }
    hk.poll = poll;
    hk.pctx = pctx;
    hk.poll_every = poll_every;
    t0 = now_s();
    rc = sim_run_hooked(s, &hk);
    if (rc == 0) {
        rc = fca_hip_sim_sync(s);
    }
    if (tm) {
        tm->wall_total += now_s() - t0;
        tm->t_flux = tm->t_limiter = tm->t_update = tm->t_dt = -1;
    }
    if (rc != 0 || read_ctrl(s, &c) != 0 || sim_finalize(s, &c) != 0 ||
        sim_acc(s, &c, acc) != 0) {
        return -2;
    }
    t0 = now_s();
    nb = fca_k_ens_count(&s->d, e->counts, e->nthr, e->thr, s->st);
    if (nb < 0) {
        return fail(-2, "count launch failed");
    }
    HIP_OR(hipMemcpyAsync(s->h_red, s->d.red, (size_t)nb * 5 * sizeof(double),
                          hipMemcpyDeviceToHost, s->st), -2);
    HIP_OR(hipStreamSynchronize(s->st), -2);
    for (j = 0; j < e->nthr; j++) {
        double f = 0.0;
        for (b = 0; b < nb; b++) {
            f += s->h_red[(size_t)b * 5 + j];
        }
        flooded[j] = (long)f;
    }
    *max_depth = 0.0;
    for (b = 0; b < nb; b++) {
        double m = s->h_red[(size_t)b * 5 + 4];
        *max_depth = m > *max_depth ? m : *max_depth;
    }
    if (tm) {
        tm->t_d2h = (tm->t_d2h > 0 ? tm->t_d2h : 0) + (now_s() - t0);
    }
    return 0;
}

int fca_hip_ens_counts_u32(FcaHipEns *e, uint32_t *host)
{
    if (!e || !host) {
        return fail(-1, "bad arguments");
    }
    HIP_OR(hipMemcpy(host, e->counts, (size_t)e->nthr * e->s->N * sizeof(uint32_t),
                     hipMemcpyDeviceToHost), -2);
    return 0;
}

int fca_hip_ens_counts(FcaHipEns *e, uint16_t *host)
{
    size_t i, n;
    uint32_t *tmp;
    if (!e || !host) {
        return fail(-1, "bad arguments");
    }
    n = (size_t)e->nthr * e->s->N;
    tmp = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!tmp) {
        return fail(-2, "out of host memory");
    }
    if (fca_hip_ens_counts_u32(e, tmp) != 0) {
        free(tmp);
        return -2;
    }
    for (i = 0; i < n; i++) {
        host[i] = (uint16_t)(tmp[i] > 65535u ? 65535u : tmp[i]);
    }
    free(tmp);
    return 0;
}

void fca_hip_ens_close(FcaHipEns *e)
{
    if (!e) {
        return;
    }
    if (e->counts) {
        hipFree(e->counts);
    }
    fca_hip_sim_close(e->s);
    free(e);
}

/* ---- batched session ------------------------------------------------------- */

struct FcaHipBatch {
    int n, K;
    int cpl; /* columns per work-item of every scenario's walk */
    FcaHipSim **sims;
    hipStream_t st;
    FcaDevCtrl *ctrl_pool; /* device: 2 records per scenario, contiguous */
    FcaDevCtrl *h_ctrl;    /* pinned copy of the pool */
    void *tab;             /* device scenario table */
    int *wgs_dev, *wgs_host;
    int nwg, cap_wg;
    char *active;
    int *last_step;
    long long *k_done; /* launch count when a scenario left the list: its
                        * current controller record is ctrl[k_done & 1] */
    long long k;
    int batch;
    double wall;
    hipGraphExec_t gexec; /* batch launches over the current workgroup list */
};

static int launch_batch_step(void *ctx, long long k)
{
    FcaHipBatch *b = (FcaHipBatch *)ctx;
    return fca_k_step_batch(b->tab, b->wgs_dev, b->nwg, b->K, b->cpl, k, b->st);
}

/* Workgroup list of the active scenarios: {scenario, workgroup} pairs. */
static int batch_list(FcaHipBatch *b)
{
    int i, w, m = 0;
    for (i = 0; i < b->n; i++) {
        if (!b->active[i]) {
            continue;
        }
        for (w = 0; w < b->sims[i]->d.nblocks; w++) {
            b->wgs_host[2 * m] = i;
            b->wgs_host[2 * m + 1] = w;
            m++;
        }
    }
    b->nwg = m;
    if (b->gexec) { /* captured over the previous list */
        hipGraphExecDestroy(b->gexec);
        b->gexec = NULL;
    }
    if (m > 0) {
        HIP_OR(hipMemcpyAsync(b->wgs_dev, b->wgs_host, (size_t)m * 2 * sizeof(int),
                              hipMemcpyHostToDevice, b->st), -2);
        HIP_OR(hipStreamSynchronize(b->st), -2);
    }
    return 0;
}

void fca_hip_batch_close(FcaHipBatch *b)
{
    int i;
    if (!b) {
        return;
    }
    if (b->st) {
        hipStreamSynchronize(b->st);
    }
    for (i = 0; i < b->n; i++) {
        if (b->sims && b->sims[i]) {
            b->sims[i]->d.ctrl = NULL; /* a slice of the pool */
            fca_hip_sim_close(b->sims[i]);
        }
    }
    if (b->gexec) {
        hipGraphExecDestroy(b->gexec);
    }
    if (b->ctrl_pool) {
        hipFree(b->ctrl_pool);
    }
    if (b->h_ctrl) {
        hipHostFree(b->h_ctrl);
    }
    if (b->tab) {
        hipFree(b->tab);
    }
    if (b->wgs_dev) {
        hipFree(b->wgs_dev);
    }
    if (b->wgs_host) {
        hipHostFree(b->wgs_host);
    }
    if (b->st) {
        hipStreamDestroy(b->st);
    }
    free(b->sims);
    free(b->active);
    free(b->last_step);
    free(b->k_done);
    free(b);
}

/* An empty batch of n scenarios with step kernel K (NULL on failure). */
static FcaHipBatch *batch_new(int n, int K, const FcaHipOpts *opts)
{
    FcaHipBatch *b = (FcaHipBatch *)calloc(1, sizeof(*b));
    if (!b) {
        fail(-2, "out of host memory");
        return NULL;
    }
    b->n = n;
    b->K = K;
    b->cpl = (opts && opts->cols_per_lane == 2) ? 2 : 1;
    b->batch = (opts && opts->graph_batch > 0) ? opts->graph_batch : 30;
    b->sims = (FcaHipSim **)calloc((size_t)n, sizeof(FcaHipSim *));
    b->active = (char *)calloc((size_t)n, 1);
    b->last_step = (int *)calloc((size_t)n, sizeof(int));
    b->k_done = (long long *)calloc((size_t)n, sizeof(long long));
    if (!b->sims || !b->active || !b->last_step || !b->k_done) {
        fail(-2, "out of host memory");
        fca_hip_batch_close(b);
        return NULL;
    }
    if (hipStreamCreate(&b->st) != hipSuccess) {
        fail(-2, "hipStreamCreate failed");
        fca_hip_batch_close(b);
        return NULL;
    }
    return b;
}

/* Scenario i's session is in place: count its workgroups. */
static void batch_add(FcaHipBatch *b, int i, int step0)
{
    b->cap_wg += b->sims[i]->d.nblocks;
    b->active[i] = 1;
    b->last_step[i] = step0;
}

/* With every session open: one pool for every controller record (a single
 * copy reads them all), the scenario table and the workgroup list. Closes
 * the batch and returns NULL on failure. */
static FcaHipBatch *batch_finish(FcaHipBatch *b)
{
    FcaDev **devs = (FcaDev **)calloc((size_t)b->n, sizeof(FcaDev *));
    int i;
    if (!devs) {
        fail(-2, "out of host memory");
        goto fail_out;
    }
    if (hipMalloc((void **)&b->ctrl_pool, (size_t)b->n * 2 * sizeof(FcaDevCtrl)) != hipSuccess ||
        hipHostMalloc((void **)&b->h_ctrl, (size_t)b->n * 2 * sizeof(FcaDevCtrl),
                      hipHostMallocDefault) != hipSuccess) {
        fail(-2, "controller pool allocation failed");
        goto fail_out;
    }
    for (i = 0; i < b->n; i++) {
        FcaDev *d = &b->sims[i]->d;
        if (hipMemcpyAsync(&b->ctrl_pool[2 * i], d->ctrl, 2 * sizeof(FcaDevCtrl),
                           hipMemcpyDeviceToDevice, b->st) != hipSuccess) {
            fail(-2, "controller pool copy failed");
            goto fail_out;
        }
        d->ctrl = &b->ctrl_pool[2 * i]; /* the session's own pair stays unused in its arena */
        devs[i] = d;
    }
    if (hipStreamSynchronize(b->st) != hipSuccess) {
        fail(-2, "controller pool copy failed");
        goto fail_out;
    }
    b->tab = fca_k_batch_table((const FcaDev *const *)devs, b->n);
    if (!b->tab) {
        fail(-2, "scenario table upload failed");
        goto fail_out;
    }
    if (hipMalloc((void **)&b->wgs_dev, (size_t)b->cap_wg * 2 * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&b->wgs_host, (size_t)b->cap_wg * 2 * sizeof(int),
                      hipHostMallocDefault) != hipSuccess) {
        fail(-2, "workgroup list allocation failed");
        goto fail_out;
    }
    if (batch_list(b) != 0) {
        goto fail_out;
    }
    free(devs);
    return b;
fail_out:
    free(devs);
    fca_hip_batch_close(b);
    return NULL;
}

/* Options of a batch's sessions: with no strip height given, every scenario
 * takes 96 rows (a batch fills the GPU with its many scenarios, so the
 * single-grid default, which shortens strips to fill it, would only add halo
 * rows; measured on the surrogate's 4096-run dataset: 35.3 s at 96 rows,
 * 37.9 s at 32). */
static const FcaHipOpts *batch_opts(const FcaHipOpts *opts, FcaHipOpts *bo)
{
    if (opts && opts->strip_rows > 0) {
        return opts;
    }
    if (opts) {
        *bo = *opts;
    } else {
        memset(bo, 0, sizeof(*bo));
    }
    bo->strip_rows = 96;
    return bo;
}

/* Prefixes the current error with the scenario index. */
static void scenario_error(int i)
{
    char keep[sizeof(g_err) + 32];
    snprintf(keep, sizeof(keep), "scenario %d: %s", i, g_err);
    memcpy(g_err, keep, sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
}

FcaHipBatch *fca_hip_batch_open(int n, const FcaGrid *const *g, const FcaRunCfg *const *cfg,
                                const FcaAcc *acc, const FcaHipOpts *opts)
{
    FcaHipBatch *b;
    FcaHipOpts bo;
    int i;
    g_err[0] = 0;
    if (n <= 0 || !g || !cfg || !acc) {
        fail(-1, "bad arguments");
        return NULL;
    }
    if (opts && opts->kernel != FCA_HIP_FUSED) {
        fail(-1, "the batched session runs the fused step only");
        return NULL;
    }
    for (i = 0; i < n; i++) {
        if (!g[i] || !cfg[i] || cfg[i]->phys.K != cfg[0]->phys.K) {
            fail(-1, "scenario %d: missing, or K differs from scenario 0", i);
            return NULL;
        }
        if (cfg[i]->n_force_days > 0) {
            fail(-1, "scenario %d: multi-day forcing is not supported in a batch", i);
            return NULL;
        }
    }
    b = batch_new(n, cfg[0]->phys.K, opts);
    if (!b) {
        return NULL;
    }
    opts = batch_opts(opts, &bo);
    for (i = 0; i < n; i++) {
        b->sims[i] = sim_create(g[i], cfg[i], &acc[i], opts, b->st);
        if (!b->sims[i]) {
            scenario_error(i);
            fca_hip_batch_close(b);
            return NULL;
        }
        batch_add(b, i, acc[i].step);
    }
    return batch_finish(b);
}

FcaHipBatch *fca_hip_batch_open_crops(const FcaHipBase *base, int n, const FcaHipCrop *crops,
                                      const FcaRunCfg *const *cfg, const FcaHipOpts *opts)
{
    FcaHipBatch *b;
    FcaHipOpts bo;
    int i;
    g_err[0] = 0;
    if (!base || n <= 0 || !crops || !cfg) {
        fail(-1, "bad arguments");
        return NULL;
    }
    if (opts && opts->kernel != FCA_HIP_FUSED) {
        fail(-1, "the batched session runs the fused step only");
        return NULL;
    }
    for (i = 0; i < n; i++) {
        if (!cfg[i] || cfg[i]->phys.K != cfg[0]->phys.K) {
            fail(-1, "scenario %d: missing, or K differs from scenario 0", i);
            return NULL;
        }
        if (cfg[i]->n_force_days > 0) {
            fail(-1, "scenario %d: multi-day forcing is not supported in a batch", i);
            return NULL;
        }
    }
    b = batch_new(n, cfg[0]->phys.K, opts);
    if (!b) {
        return NULL;
    }
    opts = batch_opts(opts, &bo);
    for (i = 0; i < n; i++) {
        b->sims[i] = sim_create_crop(base, &crops[i], cfg[i], opts, b->st);
        if (!b->sims[i]) {
            scenario_error(i);
            fca_hip_batch_close(b);
            return NULL;
        }
        batch_add(b, i, 0);
    }
    return batch_finish(b);
}

int fca_hip_batch_nscen(const FcaHipBatch *b)
{
    return b ? b->n : 0;
}

/* Runs every scenario to its end (T_sim, or cfg->steps): batches of launches
 * advance all active scenarios one step each per launch; between batches one
 * copy reads all controller records and finished scenarios leave the
 * workgroup list. */
int fca_hip_batch_run(FcaHipBatch *b)
{
    int i, j, nact;
    double t0;
    if (!b) {
        return fail(-1, "NULL batch");
    }
    t0 = now_s();
    for (;;) {
        int changed = 0;
        nact = 0;
        for (i = 0; i < b->n; i++) {
            nact += b->active[i];
        }
        if (nact == 0) {
            break;
        }
        if (b->batch % FCA_GRAPH_PERIOD == 0 && b->k % FCA_GRAPH_PERIOD == 0) {
            if (!b->gexec &&
                capture_steps(b->st, &b->gexec, b->batch, launch_batch_step, b) != 0) {
                return -2;
            }
            if (hipGraphLaunch(b->gexec, b->st) != hipSuccess) {
                return fail(-2, "graph launch failed");
            }
            b->k += b->batch;
        } else {
            for (j = 0; j < b->batch; j++) {
                if (launch_batch_step(b, b->k) != 0) {
                    return fail(-2, "batched step launch failed: %s",
                                hipGetErrorString(hipGetLastError()));
                }
                b->k++;
            }
        }
        HIP_OR(hipMemcpyAsync(b->h_ctrl, b->ctrl_pool, (size_t)b->n * 2 * sizeof(FcaDevCtrl),
                              hipMemcpyDeviceToHost, b->st), -2);
        HIP_OR(hipStreamSynchronize(b->st), -2);
        for (i = 0; i < b->n; i++) {
            const FcaRunCfg *cfg = b->sims[i]->cfg;
            const FcaDevCtrl *c = &b->h_ctrl[2 * i + (int)(b->k & 1)];
            int done;
            if (!b->active[i]) {
                continue;
            }
            if (c->state == FCA_CTL_PAUSE) {
                return fail(-2, "scenario %d paused for forcing", i);
            }
            /* a fixed-step scenario that made no progress in a whole batch
             * cannot advance (the reference would spin) */
            done = c->state == FCA_CTL_DONE ||
                   (cfg->steps > 0 && (c->step >= cfg->steps || c->step == b->last_step[i])) ||
                   (cfg->steps <= 0 && c->t >= (double)cfg->T_sim - 1e-15);
            b->last_step[i] = c->step;
            if (done) {
                b->active[i] = 0;
                b->k_done[i] = b->k;
                changed = 1;
            }
        }
        if (changed && batch_list(b) != 0) {
            return -2;
        }
    }
    /* hmax and tmax complete on the device before anything reads them (the
     * flooded counts, count rasters and the device fields handed out) */
    for (i = 0; i < b->n; i++) {
        b->sims[i]->k = b->k_done[i];
        if (sim_finalize(b->sims[i], &b->h_ctrl[2 * i + (int)(b->k_done[i] & 1)]) != 0) {
            return -2;
        }
    }
    HIP_OR(hipStreamSynchronize(b->st), -2);
    b->wall += now_s() - t0;
    return 0;
}

/* A scenario's session sees its own launch count: the whole batch's while
 * it is in the list, the count at which it left otherwise. */
static void batch_sync_k(FcaHipBatch *b, int i)
{
    b->sims[i]->k = b->active[i] ? b->k : b->k_done[i];
}

int fca_hip_batch_read(FcaHipBatch *b, int i, real *h, real *hmax, real *tmax, FcaAcc *acc)
{
    if (!b || i < 0 || i >= b->n) {
        return fail(-1, "bad arguments");
    }
    batch_sync_k(b, i);
    return fca_hip_sim_read(b->sims[i], h, hmax, tmax, acc);
}

int fca_hip_batch_steplog(FcaHipBatch *b, int i, FcaHipStepRec *out, int cap)
{
    if (!b || i < 0 || i >= b->n) {
        return fail(-1, "bad arguments");
    }
    batch_sync_k(b, i);
    return fca_hip_sim_steplog(b->sims[i], out, cap);
}

/* LAND cells of scenario i with hmax >= thr[j] and the largest LAND hmax
 * (ens_common.c:166-188), reduced on the device. */
int fca_hip_batch_flooded(FcaHipBatch *b, int i, int nthr, const double *thr, long *flooded,
                          double *max_depth)
{
    FcaHipSim *s;
    int nb, k, j;
    if (!b || i < 0 || i >= b->n || nthr < 0 || nthr > 4 || (nthr > 0 && (!thr || !flooded)) ||
        !max_depth) {
        return fail(-1, "bad arguments");
    }
    s = b->sims[i];
    nb = fca_k_ens_count(&s->d, NULL, nthr, thr, b->st);
    if (nb < 0) {
        return fail(-2, "count launch failed");
    }
    HIP_OR(hipMemcpyAsync(s->h_red, s->d.red, (size_t)nb * 5 * sizeof(double),
                          hipMemcpyDeviceToHost, b->st), -2);
    HIP_OR(hipStreamSynchronize(b->st), -2);
    for (j = 0; j < nthr; j++) {
        double f = 0.0;
        for (k = 0; k < nb; k++) {
            f += s->h_red[(size_t)k * 5 + j];
        }
        flooded[j] = (long)f;
    }
    *max_depth = 0.0;
    for (k = 0; k < nb; k++) {
        const double m = s->h_red[(size_t)k * 5 + 4];
        *max_depth = m > *max_depth ? m : *max_depth;
    }
    return 0;
}

int fca_hip_batch_add_counts_u32(FcaHipBatch *b, int nthr, const double *thr, uint32_t *host)
{
    uint32_t *dev = NULL, *tmp = NULL;
    size_t N, n, j;
    int i, rc = 0;
    if (!b || nthr < 1 || nthr > 4 || !thr || !host) {
        return fail(-1, "bad arguments");
    }
    N = b->sims[0]->N;
    for (i = 1; i < b->n; i++) {
        if (b->sims[i]->d.nr != b->sims[0]->d.nr || b->sims[i]->d.nc != b->sims[0]->d.nc) {
            return fail(-1, "count rasters need scenarios of one shape (scenario %d differs)", i);
        }
    }
    n = (size_t)nthr * N;
    HIP_OR(hipMalloc((void **)&dev, n * sizeof(uint32_t)), -2);
    tmp = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!tmp) {
        hipFree(dev);
        return fail(-2, "out of host memory");
    }
    if (fca_k_zero(dev, n * sizeof(uint32_t), b->st) != 0) {
        rc = fail(-2, "count clear failed");
    }
    /* counts[j*N + a] += (LAND && hmax >= thr[j]) over every scenario, on the
     * device; the order of the additions does not matter (integers) */
    for (i = 0; i < b->n && rc == 0; i++) {
        if (fca_k_ens_count(&b->sims[i]->d, dev, nthr, thr, b->st) < 0) {
            rc = fail(-2, "count launch failed");
        }
    }
    if (rc == 0 && (hipMemcpyAsync(tmp, dev, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   b->st) != hipSuccess ||
                    hipStreamSynchronize(b->st) != hipSuccess)) {
        rc = fail(-2, "count read-back failed");
    }
    for (j = 0; j < n && rc == 0; j++) {
        host[j] += tmp[j];
    }
    free(tmp);
    hipFree(dev);
    return rc;
}

int fca_hip_batch_status(FcaHipBatch *b, int i, int *step, double *t)
{
    FcaDevCtrl c;
    if (!b || i < 0 || i >= b->n || !step || !t) {
        return fail(-1, "bad arguments");
    }
    batch_sync_k(b, i);
    if (read_ctrl(b->sims[i], &c) != 0) {
        return -2;
    }
    *step = c.step;
    *t = c.t;
    return 0;
}

int fca_hip_batch_device_fields(FcaHipBatch *b, int i, void **hmax, void **mask, int *nr,
                                int *nc)
{
    if (!b || i < 0 || i >= b->n || !hmax || !mask || !nr || !nc) {
        return fail(-1, "bad arguments");
    }
    *hmax = b->sims[i]->d.hmax;
    *mask = b->sims[i]->d.mask;
    *nr = b->sims[i]->d.nr;
    *nc = b->sims[i]->d.nc;
    return 0;
}

double fca_hip_batch_wall(const FcaHipBatch *b)
{
    return b ? b->wall : 0.0;
}
