/*
 * main_hip.c -- the reference's command-line driver with this library as its
 * backend: the executable a maintainer adds beside floodca_serial, floodca_omp
 * and floodca_cuda (src/cuda/main_cuda.c:13-80 is the CUDA one).
 *
 * It is compiled against the reference's own src/common (CLI, grid loading,
 * warm start, rain file, accumulators, CSV report), so it accepts the same
 * options and writes the same files and CSV columns as the other backends;
 * only the run itself goes to fca_hip_run. It includes the reference's
 * headers and declares the two entry points it calls with the reference's
 * types, which floodca_hip.h restates layout for layout (the ABI tests hold
 * the two layouts equal). The done line adds the algorithmic HBM rate of the
 * step (bytes per LAND cell and per other cell, as bench.py counts them).
 */
#include "cli.h"
#include "config.h"
#include "diag.h"
#include "grid.h"
#include "log.h"
#include "model.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* libfloodca_hip.so (include/floodca_hip.h) */
int fca_hip_run(FcaGrid *g, FcaRunCfg *cfg, FcaAcc *acc, FcaTimers *tm);
void fca_hip_device_name(char *buf, size_t n);
const char *fca_hip_last_error(void);

static long land_cells(const FcaGrid *g)
{
    size_t i, n = fca_ncells(g);
    long k = 0;
    for (i = 0; i < n; i++) {
        k += g->mask[i] == FCA_LAND;
    }
    return k;
}

int main(int argc, char **argv)
{
    FcaRunCfg cfg;
    FcaGrid g;
    FcaAcc acc;
    FcaTimers tm;
    int rc = fca_cli_load(&cfg, argc, argv, "hip");

    if (rc != 0) { /* 1: --help was printed */
        return rc > 0 ? 0 : 1;
    }
    if (!cfg.grid_dir[0]) {
        fprintf(stderr, "floodca_hip: --grid is required\n");
        return 1;
    }
    cfg.threads = 1;
    cfg.gpus = 1;
    fca_hip_device_name(cfg.gpu_model, sizeof(cfg.gpu_model));
    memset(&g, 0, sizeof(g));
    fca_timers_init(&tm);
    if (fca_grid_load_dir(&g, cfg.grid_dir) != 0) {
        return 1;
    }
    cfg.phys.dx = g.dx;
    cfg.phys.A = g.dx * g.dx;
    fca_cfg_print(&cfg);
    rc = 0;
    if (cfg.warmstart[0]) {
        rc = fca_load_warmstart_f32(cfg.warmstart, g.h, g.hmax, fca_ncells(&g));
    }
    if (rc == 0 && cfg.rain_file[0]) {
        rc = fca_grid_load_rain(&g, cfg.rain_file, cfg.lambda);
    }
    if (rc != 0) {
        fca_grid_free(&g);
        return 1;
    }
    tm.t_init = 0;
    tm.t_io = 0;
    fca_acc_init(&acc, &g, cfg.phys.A);
    if (cfg.steps > 0 || cfg.T_sim > 0) {
        const long nland = land_cells(&g), nall = (long)fca_ncells(&g);
        const double bytes = (double)nland * (6.0 * sizeof(real) + 1.0) + (double)(nall - nland);
        if (fca_hip_run(&g, &cfg, &acc, &tm) != 0) {
            fprintf(stderr, "floodca_hip: %s\n", fca_hip_last_error());
            fca_grid_free(&g);
            return 1;
        }
        printf("done backend=hip steps=%d t=%.6g E_mass=%.3e V=%.6g wall=%.3f h2d=%.3f "
               "d2h=%.3f mcups=%.6g hbm_alg_gbps=%.6g gpu=%s\n",
               acc.step, acc.t, fca_mass_error(&acc), acc.V, tm.wall_total, tm.t_h2d, tm.t_d2h,
               tm.wall_total > 0 ? (double)nland * acc.step / tm.wall_total / 1e6 : 0.0,
               tm.wall_total > 0 ? bytes * acc.step / tm.wall_total / 1e9 : 0.0, cfg.gpu_model);
    }
    if (cfg.out_dir[0] && !cfg.no_output) {
        fca_ensure_dir(cfg.out_dir);
        if (fca_grid_write_hmax(&g, cfg.out_dir) != 0 ||
            fca_grid_write_field(&g, cfg.out_dir, "h_final", g.h) != 0) {
            fprintf(stderr, "floodca_hip: cannot write hmax/h_final under %s\n", cfg.out_dir);
            fca_grid_free(&g);
            return 1;
        }
    }
    fca_report(&cfg, &g, &acc, &tm, getenv("FCA_NODE_TYPE") ? getenv("FCA_NODE_TYPE") : "", "",
               "");
    fca_grid_free(&g);
    return 0;
}
