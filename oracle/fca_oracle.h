/*
 * fca_oracle.h -- CPU restatement of the FloodCA WCA step, used ONLY as the
 * checker of the HIP backend (tests/, __graft_entry__.smoke, bench.py's CPU
 * baseline). Nothing in the product links or calls it.
 *
 * Every function restates one piece of the reference with the same operation
 * order, so that with contraction off (-ffp-contract=off) its results are
 * bitwise those of the reference serial build; tests/test_oracle_vs_ref.py
 * checks exactly that against oracle/_ref (the reference sources compiled
 * here). Parity status per function is listed in the header of fca_oracle.c.
 */
#ifndef FCA_ORACLE_H
#define FCA_ORACLE_H

#include "floodca_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Face flux a->b, positive = a to b (src/kernels/physics.h:51-125). */
real or_face_flux(real za, real ha, real na, uint8_t ma, real zb, real hb, real nb,
                  uint8_t mb, real d, real w, real dt, int K, real A, const FcaPhys *p,
                  int *cap_active);
/* (real)cbrt((double)hf*(double)hf) (physics.h:19-24). */
real or_cbrt_sq(real hf);
/* Limiter factor (physics.h:127-134). */
real or_limiter_s(real out_vol, real h, real A);
/* Per-cell stable step of a wet LAND cell, dt_max otherwise
 * (model.c:123-137 with physics.h:158-217). */
real or_cell_dt(const real *z, const real *h, const real *n, const uint8_t *mask, int r,
                int c, int nrows, int ncols, const FcaPhys *p);
/* Global held step with the dt_min/dt_max clamp (model.c:111-156). */
real or_compute_dt_base(const FcaGrid *g, const FcaPhys *p, FcaAcc *acc);
/* Hyetograph integral and next event (src/common/rain.c:30-93). */
double or_hyeto_integral(const FcaRunCfg *cfg, double t, double dt);
double or_next_event_time(const FcaRunCfg *cfg, double t);
int or_forcing_day_index(double t);
/* LAND volume and mass-balance error (grid.c:125-141, diag.c:5-14). */
double or_volume_land(const FcaGrid *g, real A);
void or_acc_init(FcaAcc *acc, const FcaGrid *g, real A);
double or_mass_error(const FcaAcc *acc);

/* One step of fca_step (model.c:321-605) on g's z, h, n, C, mask, Rday, hmax,
 * tmax. Scratch face arrays are allocated per call. Returns 1 if a step was
 * taken, 0 if dt <= 0 or the run ended on a missing forcing file, -1 on
 * allocation failure. rec (may be NULL) receives the controller's values. */
int or_step(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, FcaHipStepRec *rec);
/* fca_run (model.c:607-643) without the progress prints. log (may be NULL)
 * receives the first log_cap steps; *n_logged the number of steps taken. */
int or_run(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, FcaHipStepRec *log, int log_cap,
           int *n_logged);

/* Reference float-mode libm expressions over a bit range (see fca_oracle.c). */
void or_libm_range(int kind, uint32_t first, uint32_t count, float g, float *out);

/* Floats in [first, first+count) whose float-mode cbrt_sq is not the
 * correctly rounded x^(2/3) (see fca_oracle.c). */
long or_cbrt_sq_cr_mismatches(uint32_t first, uint32_t count);

/* Ensemble count of one scenario (src/hybrid/ens_common.c:166-188) with
 * uint32 counts: counts[j*N + a] += (LAND && hmax >= thr[j]). */
void or_ens_count(const FcaGrid *g, int nthr, const double *thr, uint32_t *counts,
                  long *flooded, double *max_depth);

#ifdef __cplusplus
}
#endif

#endif
