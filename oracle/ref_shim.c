/*
 * ref_shim.c -- compiled together with the reference's own src/common/*.c into
 * oracle/_ref/libfca_ref_{f32,f64}.so by oracle/build_oracle.sh, against the
 * reference's headers. It only exposes to ctypes what the reference keeps
 * static inline (src/kernels/physics.h) and logs the reference's own fca_step
 * loop; it restates no arithmetic. Test infrastructure, never shipped.
 */
#include "fca_types.h"
#include "grid.h"
#include "model.h"
#include "rain.h"
#include "physics.h"

real ref_face_flux(real za, real ha, real na, uint8_t ma, real zb, real hb, real nb,
                   uint8_t mb, real d, real w, real dt, int K, real A, const FcaPhys *p,
                   int *cap_active)
{
    return fca_face_flux(za, ha, na, ma, zb, hb, nb, mb, d, w, dt, K, A, p, cap_active);
}

real ref_cbrt_sq(real hf)
{
    return fca_cbrt_sq(hf);
}

real ref_limiter_s(real out_vol, real h, real A)
{
    return fca_limiter_s(out_vol, h, A);
}

/* The per-cell body of fca_compute_dt_base (model.c:123-137) for a whole grid. */
real ref_cell_dt(const real *z, const real *h, const real *n, const uint8_t *mask, int r,
                 int c, int nrows, int ncols, const FcaPhys *p)
{
    int a = r * ncols + c, open_edge;
    real smax;
    if (mask[a] != FCA_LAND || h[a] < p->tau_h) {
        return p->dt_max;
    }
    open_edge = p->boundary == FCA_BND_OPEN &&
                (c == 0 || c == ncols - 1 || r == 0 || r == nrows - 1);
    smax = fca_cell_smax(z, h, mask, r, c, nrows, ncols, p, open_edge);
    return fca_cell_dt_from_smax(h[a], n[a], smax, p);
}

/* fca_run's loop (model.c:607-643) with the three controller values the
 * reference keeps per step logged: t before and after the step, and the held
 * step dt_base. Stops where fca_run stops; with steps > 0 it also stops when
 * a call takes no step (where fca_run would spin). Returns steps taken. */
int ref_run_logged(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, double *log, int cap)
{
    int max_steps = cfg->steps > 0 ? cfg->steps : 2000000000, n = 0;
    while (acc->step < max_steps) {
        double t_before = acc->t;
        int step_before = acc->step;
        if (cfg->steps <= 0 && acc->t >= (double)cfg->T_sim - 1e-15) {
            break;
        }
        fca_step(g, cfg, acc, NULL, 0.0);
        if (acc->step != step_before) {
            if (log && n < cap) {
                log[3 * n + 0] = t_before;
                log[3 * n + 1] = acc->t;
                log[3 * n + 2] = (double)acc->dt_base;
            }
            n++;
        } else if (cfg->steps > 0) {
            break;
        }
        if (cfg->steps <= 0 && acc->t <= t_before) {
            break;
        }
    }
    acc->V = fca_volume_land(g, cfg->phys.A);
    return n;
}
