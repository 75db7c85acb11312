/*
 * main_ens_hip.c -- the reference's GPU ensemble (src/hybrid/main_ens_mpi_cuda.c,
 * "B6") as one process on one MI355X, with this library as its backend.
 *
 * The reference runs a scenario table through an MPI master-worker loop, one
 * rank per GPU, each rank keeping the static grid on its device
 * (fca_cuda_ens_*), and reduces the per-cell flood counts to rank 0. On one
 * GPU there is nothing to distribute, so this driver keeps the rest as it is:
 *
 *   input   the reference's CLI (fca_cli_load: --grid, --tsim, --neigh, --out,
 *           --csv, --exp, ...) plus the ensemble's own options --scenario-file
 *           CSV, --rain-dir DIR, --first K, --nscen N, --ens-csv FILE,
 *           --poll-every STEPS (main_ens_mpi_cuda.c:137-142); the scenario
 *           table scenario_id,base_date,lambda,D_h,phi,seed
 *           (prep/floodca_prep/scenarios.py:18, read as ens_common.c:27-70
 *           reads it);
 *   each    rain_<base_date>.bin scaled by lambda (fca_grid_load_rain), depth
 *           reset to the grid's initial h, hmax = h, tmax = 0, a triangular
 *           hyetograph of D_h and phi, the run to --tsim
 *           (main_ens_mpi_cuda.c:71-120);
 *   output  count_010/025/050.bin (uint16, LAND cells with hmax >= 0.10, 0.25,
 *           0.50 m, summed over scenarios) with their JSON headers under --out
 *           (ens_common.c:190-221), one row per scenario appended to --ens-csv
 *           in the reference's columns (ens_common.c:222-258), and the done line
 *           with scenarios per hour as main_ens_mpi_cuda.c:228-230 defines it
 *           (scenarios / wall of the whole ensemble, counts included).
 *
 * --batch 0 (default) runs the scenarios one after another through the
 * ensemble session fca_hip_ens_*, the calls the reference's rank makes;
 * --batch B runs them B at a time in one batched session (fca_hip_batch_*),
 * each scenario with its own time step, for throughput. Both give the same
 * depths, counts and step counts bit for bit (the batch's per-scenario
 * arithmetic is a session's); the FP64 balance sums, added per workgroup,
 * may differ in their last digits, as the batch walks strips of its own
 * height.
 *
 * Compiled against the reference's src/common (oracle/build_hip_main.sh), so
 * grid loading, rain files, accumulators and the bench CSV row are the
 * reference's own code.
 */
#include "cli.h"
#include "config.h"
#include "diag.h"
#include "grid.h"
#include "log.h"
#include "model.h"

#include "floodca_hip.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#define ENS_NTHR 3
#define ENS_MAX_SCEN 4096

static const double ENS_THR[ENS_NTHR] = {0.10, 0.25, 0.50};
static const char *const ENS_TAG[ENS_NTHR] = {"010", "025", "050"};

/* One row of the scenario table; the strings are kept as read (the output
 * CSV echoes them), in fields as wide as the reference's (ens_common.h:15-22). */
typedef struct Scen {
    int id;
    char date[16];
    char lambda[32];
    char D_h[16];
    char phi[32];
    char seed[24];
} Scen;

/* One scenario's results, the fields of the reference's per-scenario row. */
typedef struct Done {
    int steps;
    double t, wall, V_rain, V_sea, E_mass, max_depth;
    long flooded[ENS_NTHR];
} Done;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* The value after `key` on the command line, or dflt. */
static const char *opt(int argc, char **argv, const char *key, const char *dflt)
{
    int i;
    for (i = 1; i + 1 < argc; i++) {
        if (!strcmp(argv[i], key)) {
            return argv[i + 1];
        }
    }
    return dflt;
}

/* Splits line in place into its first six fields. A field ends at ',', '\r',
 * '\n' or the end of the line; the first five must end at a separator.
 * Returns 1 when six fields were found. */
static int six_fields(char *line, char *f[6])
{
    char *p = line;
    int k;
    for (k = 0; k < 6; k++) {
        size_t len = strcspn(p, ",\r\n");
        f[k] = p;
        if (p[len] == 0) {
            return k == 5;
        }
        p[len] = 0;
        p += len + 1;
    }
    return 1;
}

/* Reads up to nmax scenarios after the header row, skipping the first
 * `first` data rows; rows with fewer than six fields do not count. Returns
 * the number read, or -1 if the file cannot be read. */
static int read_scenarios(const char *path, Scen *s, int first, int nmax)
{
    char line[512];
    int n = 0;
    FILE *fp = fopen(path, "r");
    if (!fp) {
        perror(path);
        return -1;
    }
    if (!fgets(line, sizeof(line), fp)) { /* header */
        fclose(fp);
        return -1;
    }
    while (n < nmax && fgets(line, sizeof(line), fp)) {
        char *f[6];
        if (!six_fields(line, f)) {
            continue;
        }
        if (first > 0) {
            first--;
            continue;
        }
        s[n].id = atoi(f[0]);
        snprintf(s[n].date, sizeof(s[n].date), "%s", f[1]);
        snprintf(s[n].lambda, sizeof(s[n].lambda), "%s", f[2]);
        snprintf(s[n].D_h, sizeof(s[n].D_h), "%s", f[3]);
        snprintf(s[n].phi, sizeof(s[n].phi), "%s", f[4]);
        snprintf(s[n].seed, sizeof(s[n].seed), "%s", f[5]);
        n++;
    }
    fclose(fp);
    return n;
}

/* The scenario's configuration and forcing on g: rain file times lambda into
 * Rday, depth reset to h0, hmax = h0, tmax = 0. */
static int scenario_setup(FcaGrid *g, const real *h0, const FcaRunCfg *base, const Scen *sc,
                          const char *rain_dir, FcaRunCfg *cfg)
{
    char path[768];
    size_t i, N = fca_ncells(g);
    *cfg = *base;
    cfg->lambda = (real)atof(sc->lambda);
    cfg->D_h = (real)atof(sc->D_h);
    cfg->phi = (real)atof(sc->phi);
    cfg->hyeto = FCA_HYETO_TRIANGLE;
    snprintf(path, sizeof(path), "%s/rain_%s.bin", rain_dir, sc->date);
    if (fca_grid_load_rain(g, path, cfg->lambda) != 0) {
        fprintf(stderr, "floodca_ens_hip: failed to load %s\n", path);
        return -1;
    }
    for (i = 0; i < N; i++) {
        g->h[i] = h0[i];
        g->hmax[i] = h0[i];
        g->tmax[i] = 0;
    }
    return 0;
}

static void fill_done(Done *d, const FcaAcc *acc, double wall, const long *flooded,
                      double max_depth)
{
    int j;
    d->steps = acc->step;
    d->t = acc->t;
    d->wall = wall;
    d->V_rain = acc->V_rain;
    d->V_sea = acc->V_sea;
    d->E_mass = fca_mass_error(acc);
    d->max_depth = max_depth;
    for (j = 0; j < ENS_NTHR; j++) {
        d->flooded[j] = flooded[j];
    }
}

/* Scenarios one after another through the ensemble session. */
static int run_sessions(FcaGrid *g, const real *h0, const FcaRunCfg *base, const Scen *sc,
                        int nscen, const char *rain_dir, int poll_every, Done *done,
                        uint32_t *counts)
{
    FcaHipEns *e = fca_hip_ens_open(g, (FcaRunCfg *)base, ENS_NTHR, ENS_THR);
    int k;
    if (!e) {
        fprintf(stderr, "floodca_ens_hip: %s\n", fca_hip_last_error());
        return -1;
    }
    for (k = 0; k < nscen; k++) {
        FcaRunCfg cfg;
        FcaAcc acc;
        FcaTimers tm;
        long flooded[ENS_NTHR];
        double max_depth = 0.0;
        if (scenario_setup(g, h0, base, &sc[k], rain_dir, &cfg) != 0) {
            fca_hip_ens_close(e);
            return -1;
        }
        fca_timers_init(&tm);
        fca_acc_init(&acc, g, cfg.phys.A);
        if (fca_hip_ens_run(e, g, &cfg, &acc, &tm, flooded, &max_depth, NULL, NULL,
                            poll_every) != 0) {
            fprintf(stderr, "floodca_ens_hip: scenario %d: %s\n", sc[k].id, fca_hip_last_error());
            fca_hip_ens_close(e);
            return -1;
        }
        fill_done(&done[k], &acc, tm.wall_total, flooded, max_depth);
    }
    if (fca_hip_ens_counts_u32(e, counts) != 0) {
        fprintf(stderr, "floodca_ens_hip: %s\n", fca_hip_last_error());
        fca_hip_ens_close(e);
        return -1;
    }
    fca_hip_ens_close(e);
    return 0;
}

/* Scenarios `bsize` at a time, resident together in one batched session:
 * each scenario gets its own Rday, h, hmax and tmax, and shares z, n, C and
 * mask with the loaded grid. Every scenario of a batch reports the batch's
 * run time as its wall. */
static int run_batches(FcaGrid *g, const real *h0, const FcaRunCfg *base, const Scen *sc,
                       int nscen, const char *rain_dir, int bsize, Done *done, uint32_t *counts)
{
    const size_t N = fca_ncells(g);
    FcaGrid *gs = (FcaGrid *)calloc((size_t)bsize, sizeof(FcaGrid));
    const FcaGrid **gp = (const FcaGrid **)calloc((size_t)bsize, sizeof(FcaGrid *));
    FcaRunCfg *cfgs = (FcaRunCfg *)calloc((size_t)bsize, sizeof(FcaRunCfg));
    const FcaRunCfg **cp = (const FcaRunCfg **)calloc((size_t)bsize, sizeof(FcaRunCfg *));
    FcaAcc *accs = (FcaAcc *)calloc((size_t)bsize, sizeof(FcaAcc));
    real *fields = (real *)malloc((size_t)bsize * 4 * N * sizeof(real));
    real *rday0 = g->Rday, *h00 = g->h, *hm0 = g->hmax, *tm0 = g->tmax;
    int k0, i, rc = 0;
    if (!gs || !gp || !cfgs || !cp || !accs || !fields) {
        fprintf(stderr, "floodca_ens_hip: out of host memory for a batch of %d\n", bsize);
        rc = -1;
    }
    for (k0 = 0; k0 < nscen && rc == 0; k0 += bsize) {
        const int nb = nscen - k0 < bsize ? nscen - k0 : bsize;
        FcaHipBatch *b;
        for (i = 0; i < nb && rc == 0; i++) {
            real *f = fields + (size_t)i * 4 * N;
            gs[i] = *g;
            gs[i].Rday = f;
            gs[i].h = f + N;
            gs[i].hmax = f + 2 * N;
            gs[i].tmax = f + 3 * N;
            rc = scenario_setup(&gs[i], h0, base, &sc[k0 + i], rain_dir, &cfgs[i]);
            fca_acc_init(&accs[i], &gs[i], cfgs[i].phys.A);
            gp[i] = &gs[i];
            cp[i] = &cfgs[i];
        }
        if (rc != 0) {
            break;
        }
        b = fca_hip_batch_open(nb, gp, cp, accs, NULL);
        if (!b || fca_hip_batch_run(b) != 0) {
            fprintf(stderr, "floodca_ens_hip: batch at scenario %d: %s\n", sc[k0].id,
                    fca_hip_last_error());
            if (b) {
                fca_hip_batch_close(b);
            }
            rc = -1;
            break;
        }
        for (i = 0; i < nb && rc == 0; i++) {
            FcaAcc acc;
            long flooded[ENS_NTHR];
            double max_depth = 0.0;
            if (fca_hip_batch_read(b, i, NULL, NULL, NULL, &acc) != 0 ||
                fca_hip_batch_flooded(b, i, ENS_NTHR, ENS_THR, flooded, &max_depth) != 0) {
                rc = -1;
                break;
            }
            fill_done(&done[k0 + i], &acc, fca_hip_batch_wall(b), flooded, max_depth);
        }
        if (rc == 0 && fca_hip_batch_add_counts_u32(b, ENS_NTHR, ENS_THR, counts) != 0) {
            rc = -1;
        }
        if (rc != 0) {
            fprintf(stderr, "floodca_ens_hip: %s\n", fca_hip_last_error());
        }
        fca_hip_batch_close(b);
    }
    g->Rday = rday0, g->h = h00, g->hmax = hm0, g->tmax = tm0;
    free(gs);
    free(gp);
    free(cfgs);
    free(cp);
    free(accs);
    free(fields);
    return rc;
}

/* count_<tag>.bin (uint16, one raster per threshold) and its JSON header, in
 * the format the reference's readers expect (ens_common.c:190-221). */
static int write_counts(const char *dir, const FcaGrid *g, const uint16_t *c16, int nscen)
{
    const size_t N = fca_ncells(g);
    int j;
    for (j = 0; j < ENS_NTHR; j++) {
        char path[768];
        FILE *fp;
        size_t wrote;
        snprintf(path, sizeof(path), "%s/count_%s.bin", dir, ENS_TAG[j]);
        fp = fopen(path, "wb");
        if (!fp) {
            perror(path);
            return -1;
        }
        wrote = fwrite(c16 + (size_t)j * N, sizeof(uint16_t), N, fp);
        fclose(fp);
        if (wrote != N) {
            perror(path);
            return -1;
        }
        snprintf(path, sizeof(path), "%s/count_%s.json", dir, ENS_TAG[j]);
        fp = fopen(path, "w");
        if (!fp) {
            perror(path);
            return -1;
        }
        fprintf(fp,
                "{\"name\":\"count_%s\",\"nrows\":%d,\"ncols\":%d,\"dx\":%.9g,\"dtype\":\"uint16\","
                "\"crs\":\"EPSG:32643\",\"x0\":%.9g,\"y0\":%.9g,\"threshold_m\":%.2f,"
                "\"n_scenarios\":%d}\n",
                ENS_TAG[j], g->nrows, g->ncols, (double)g->dx, (double)g->x0, (double)g->y0,
                ENS_THR[j], nscen);
        fclose(fp);
    }
    return 0;
}

/* Appends one row per scenario (and the header to a new file) in the
 * reference's ensemble CSV columns (ens_common.c:222-258). */
static void write_csv(const char *path, const char *backend, int device, const Scen *sc,
                      const Done *done, int nscen)
{
    FILE *fp = fopen(path, "r");
    const int fresh = fp == NULL;
    int k;
    if (fp) {
        fclose(fp);
    }
    fp = fopen(path, "a");
    if (!fp) {
        perror(path);
        return;
    }
    if (fresh) {
        fputs("scenario_id,base_date,lambda,D_h,phi,seed,backend,worker_rank,device,"
              "steps,sim_time_s,wall_s,rain_volume_m3,sea_volume_m3,mass_err_rel,"
              "flooded_cells_010,flooded_cells_025,flooded_cells_050,max_depth_m\n",
              fp);
    }
    for (k = 0; k < nscen; k++) {
        const Done *d = &done[k];
        fprintf(fp, "%d,%s,%s,%s,%s,%s,%s,%d,gpu%d,%d,%.9g,%.6f,%.9e,%.9e,%.6e,%ld,%ld,%ld,%.6f\n",
                sc[k].id, sc[k].date, sc[k].lambda, sc[k].D_h, sc[k].phi, sc[k].seed, backend, 0,
                device, d->steps, d->t, d->wall, d->V_rain, d->V_sea, d->E_mass, d->flooded[0],
                d->flooded[1], d->flooded[2], d->max_depth);
    }
    fclose(fp);
}

int main(int argc, char **argv)
{
    FcaRunCfg cfg;
    FcaGrid g;
    Scen *scen;
    Done *done;
    uint32_t *counts;
    uint16_t *c16;
    real *h0;
    const char *rain_dir, *ens_csv;
    int rc, nscen, nmax, first, bsize, poll_every, device;
    double t_start, t_wall;
    size_t N, i;

    rc = fca_cli_load(&cfg, argc, argv, "ens_hip");
    if (rc != 0) {
        if (rc > 0) {
            printf("floodca_ens_hip --grid DIR --scenario-file CSV --rain-dir DIR [--first K]\n"
                   "    --nscen N --tsim S --neigh K [--batch B] [--poll-every STEPS]\n"
                   "    [--out DIR] [--ens-csv FILE] [--csv bench.csv --exp ID]\n");
        }
        return rc > 0 ? 0 : 1;
    }
    rain_dir = opt(argc, argv, "--rain-dir", "data/processed/udupi_coast/rain/90");
    ens_csv = opt(argc, argv, "--ens-csv", "");
    nmax = atoi(opt(argc, argv, "--nscen", "8"));
    first = atoi(opt(argc, argv, "--first", "0"));
    bsize = atoi(opt(argc, argv, "--batch", "0"));
    poll_every = atoi(opt(argc, argv, "--poll-every", "50"));
    if (nmax > ENS_MAX_SCEN) {
        nmax = ENS_MAX_SCEN;
    }
    device = fca_hip_select_device(0);
    if (device < 0) {
        fprintf(stderr, "floodca_ens_hip: no HIP device\n");
        return 1;
    }
    cfg.threads = 1;
    cfg.ranks = 1;
    cfg.gpus = 1;
    cfg.steps = 0;
    fca_hip_device_name(cfg.gpu_model, sizeof(cfg.gpu_model));

    scen = (Scen *)calloc(ENS_MAX_SCEN, sizeof(Scen));
    nscen = scen ? read_scenarios(cfg.scenario_file, scen, first, nmax) : -1;
    if (nscen <= 0) {
        fprintf(stderr, "floodca_ens_hip: no scenarios read from '%s'\n", cfg.scenario_file);
        free(scen);
        return 1;
    }
    memset(&g, 0, sizeof(g));
    if (fca_grid_load_dir(&g, cfg.grid_dir) != 0) {
        free(scen);
        return 1;
    }
    cfg.phys.dx = g.dx;
    cfg.phys.A = cfg.phys.dx * cfg.phys.dx;
    N = fca_ncells(&g);
    h0 = (real *)malloc(N * sizeof(real));
    done = (Done *)calloc((size_t)nscen, sizeof(Done));
    counts = (uint32_t *)calloc((size_t)ENS_NTHR * N, sizeof(uint32_t));
    c16 = (uint16_t *)malloc((size_t)ENS_NTHR * N * sizeof(uint16_t));
    if (!h0 || !done || !counts || !c16) {
        fprintf(stderr, "floodca_ens_hip: out of host memory\n");
        return 1;
    }
    memcpy(h0, g.h, N * sizeof(real));
    fca_cfg_print(&cfg);
    printf("ensemble ranks=1 gpus=1 scenarios=%d mode=%s device=%d (%s)\n", nscen,
           bsize > 0 ? "batch" : "session", device, cfg.gpu_model);
    fflush(stdout);

    t_start = now_s();
    rc = bsize > 0 ? run_batches(&g, h0, &cfg, scen, nscen, rain_dir, bsize, done, counts)
                   : run_sessions(&g, h0, &cfg, scen, nscen, rain_dir, poll_every, done, counts);
    /* the reference's uint16 rasters (the sum of its ranks' uint16 counts);
     * at most 4096 scenarios, so no count overflows */
    for (i = 0; i < (size_t)ENS_NTHR * N; i++) {
        c16[i] = (uint16_t)(counts[i] > 65535u ? 65535u : counts[i]);
    }
    t_wall = now_s() - t_start;
    if (rc == 0) {
        printf("done backend=ens_hip scenarios=%d ranks=1 wall=%.3f scen_per_hour=%.3f gpu=%s\n",
               nscen, t_wall, (double)nscen / t_wall * 3600.0, cfg.gpu_model);
        if (cfg.out_dir[0] && !cfg.no_output) {
            mkdir(cfg.out_dir, 0755);
            rc = write_counts(cfg.out_dir, &g, c16, nscen);
        }
        if (rc == 0 && ens_csv[0]) {
            write_csv(ens_csv, "ens_hip", device, scen, done, nscen);
        }
        if (rc == 0 && cfg.csv_path[0]) {
            FcaAcc acc;
            FcaTimers tm;
            int k;
            memset(&acc, 0, sizeof(acc));
            fca_timers_init(&tm);
            tm.wall_total = t_wall;
            for (k = 0; k < nscen; k++) {
                acc.step += done[k].steps;
            }
            fca_report(&cfg, &g, &acc, &tm,
                       getenv("FCA_NODE_TYPE") ? getenv("FCA_NODE_TYPE") : "", "", "");
        }
    }
    free(c16);
    free(counts);
    free(done);
    free(h0);
    free(scen);
    fca_grid_free(&g);
    return rc == 0 ? 0 : 1;
}
