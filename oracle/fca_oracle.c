/*
 * fca_oracle.c -- CPU restatement of the FloodCA WCA step (test infrastructure).
 *
 * Built with -ffp-contract=off -fno-fast-math, once per precision
 * (liboracle_f32.so, and liboracle_f64.so with -DFCA_DOUBLE), by
 * oracle/build_oracle.sh. Only tests/, __graft_entry__.smoke and bench.py's CPU
 * baseline load it, and only as the checker.
 *
 * Parity status of each function (how it is pinned):
 *   or_face_flux, or_cbrt_sq, or_limiter_s, or_cell_dt, or_compute_dt_base,
 *   or_step, or_run, or_volume_land, or_acc_init
 *       -- bitwise against the reference compiled here (oracle/_ref, same
 *          flags) on random faces and on whole runs (K=4/8, open/closed,
 *          uniform/triangle rain, FP32/FP64): tests/test_oracle_vs_ref.py;
 *          plus closed forms (wall/sea/dry/lake-at-rest faces, supercritical
 *          cap, Qcap branch, single wet cell dt, closed-box conservation):
 *          tests/test_oracle_kat.py; plus the committed golden runs under
 *          tests/golden/ that carry the reference's own outputs to the GPU box.
 *   or_hyeto_integral, or_next_event_time, or_forcing_day_index
 *       -- against the reference's functions (ref shim) on random arguments,
 *          and closed forms (uniform F = dt/86400, triangle integrates to 1
 *          over [0, T], events at day boundaries and out_every).
 *   or_mass_error -- closed form of diag.c:5-14 (three-line identity).
 *   or_ens_count -- closed form on hand-made hmax rasters
 *       (tests/test_oracle_kat.py); the reference's ensemble needs MPI and
 *       does not build here, so its counts are pinned by that rule only.
 *   Forcing reload inside or_step (model.c:338-358) -- against the reference's
 *       fca_step on a two-day forcing CSV (tests/test_oracle_vs_ref.py).
 */
#include "fca_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- face and cell formulas (src/kernels/physics.h) --------------------- */

/* The float-mode libm detours. A build for the tolerance study
 * (oracle/tolerance_study.py) may replace them with approximations through
 * these macros; the oracle's own build keeps the reference's expressions. */
#ifndef ORACLE_CBRT_SQ
#define ORACLE_CBRT_SQ(hf, sq) ((real)cbrt(sq))
#endif
#ifndef ORACLE_SQRT_SLOPE
#define ORACLE_SQRT_SLOPE(x) ((real)sqrt((double)(x)))
#endif

/* physics.h:19-24: hf^(2/3) through a double cbrt of the exact double hf^2. */
real or_cbrt_sq(real hf)
{
    double sq = (double)hf * (double)hf;
    (void)hf;
    return ORACLE_CBRT_SQ(hf, sq);
}

/* physics.h:51-125. Positive result = flow from cell a to cell b. */
real or_face_flux(real za, real ha, real na, uint8_t ma, real zb, real hb, real nb,
                  uint8_t mb, real d, real w, real dt, int K, real A, const FcaPhys *p,
                  int *cap_active)
{
    real bed_a, bed_b, lev_a, lev_b, dL, hf, nf, slope, v, vcr, Qmag, Qcap;
    uint8_t up;

    if (cap_active) {
        *cap_active = 0;
    }
    /* walls close the face (physics.h:64-66) */
    if (ma == FCA_WALL || mb == FCA_WALL) {
        return (real)0;
    }
    /* a SEA side sits at the fixed sea level with no depth (physics.h:68-73) */
    bed_a = (ma == FCA_SEA) ? p->z_sea : za;
    bed_b = (mb == FCA_SEA) ? p->z_sea : zb;
    lev_a = bed_a + ((ma == FCA_SEA) ? (real)0 : ha);
    lev_b = bed_b + ((mb == FCA_SEA) ? (real)0 : hb);
    dL = lev_a - lev_b;
    /* level-equalisation dead band (physics.h:75-81); NaN falls through */
    if (dL < (real)0) {
        if (-dL < p->tau_l) {
            return (real)0;
        }
    } else if (dL < p->tau_l) {
        return (real)0;
    }
    /* flow depth over the higher bed (physics.h:83-86) */
    hf = ((lev_a > lev_b) ? lev_a : lev_b) - ((bed_a > bed_b) ? bed_a : bed_b);
    if (hf < p->tau_h) {
        return (real)0;
    }
    /* no flow out of the sea (physics.h:88-92) */
    up = (dL > (real)0) ? ma : mb;
    if (up == FCA_SEA) {
        return (real)0;
    }
    /* face roughness (physics.h:94-103) */
    if (ma == FCA_SEA) {
        nf = nb;
    } else if (mb == FCA_SEA) {
        nf = na;
    } else {
        nf = (real)0.5 * (na + nb);
    }
    if (nf < (real)1e-12) {
        nf = (real)1e-12;
    }
    /* Manning velocity, capped at the critical velocity (physics.h:105-112) */
    slope = ((dL < (real)0) ? -dL : dL) / d;
    v = ((real)1 / nf) * or_cbrt_sq(hf) * ORACLE_SQRT_SLOPE(slope);
    vcr = (real)sqrt((double)p->g * (double)hf);
    if (v > vcr) {
        v = vcr;
    }
    /* discharge, capped so one step cannot overshoot the level
     * equalisation shared among 2K faces (physics.h:113-122) */
    Qmag = v * hf * w;
    if (dt > (real)0 && K > 0) {
        Qcap = ((dL < (real)0) ? -dL : dL) * A / ((real)(2 * K) * dt);
        if (Qcap < Qmag) {
            if (cap_active) {
                *cap_active = 1;
            }
            Qmag = Qcap;
        }
    }
    return (dL > (real)0) ? Qmag : -Qmag;
}

/* physics.h:127-134 */
real or_limiter_s(real out_vol, real h, real A)
{
    real hold = h * A;
    if (out_vol > hold && out_vol > (real)0) {
        return hold / out_vol;
    }
    return (real)1;
}

/* Water level a neighbour presents to the slope test (physics.h:136-145). */
static real level_of(uint8_t m, real z, real h, const FcaPhys *p)
{
    if (m == FCA_SEA) {
        return p->z_sea;
    }
    if (m == FCA_WALL) {
        return z;
    }
    return z + h;
}

/* physics.h:147-154 */
static real abs_slope(real la, real lb, real d)
{
    real dL = la - lb;
    if (dL < (real)0) {
        dL = -dL;
    }
    return dL / d;
}

/* model.c:123-137 with fca_cell_smax (physics.h:158-191) and
 * fca_cell_dt_from_smax (physics.h:193-217), for the whole-grid case where the
 * owned window is the grid and every side is a domain edge. */
real or_cell_dt(const real *z, const real *h, const real *n, const uint8_t *mask, int r,
                int c, int nrows, int ncols, const FcaPhys *p)
{
    /* neighbour offsets in the reference's visiting order E, W, S, N, SE, SW,
     * NE, NW (physics.h:172-181); the max does not depend on it except for
     * NaN, which never wins a '>' test either way */
    static const int dr[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    static const int dc[8] = {1, -1, 0, 0, 1, -1, 1, -1};
    int a = r * ncols + c, k, nk = (p->K == 8) ? 8 : 4;
    real la, smax = (real)0, s, nn, hh, v, vcr;
    real ddiag = (real)sqrt(2.0) * p->dx;

    if (mask[a] != FCA_LAND || h[a] < p->tau_h) {
        return p->dt_max;
    }
    la = z[a] + h[a];
    for (k = 0; k < nk; k++) {
        int rb = r + dr[k], cb = c + dc[k], b;
        if (rb < 0 || rb >= nrows || cb < 0 || cb >= ncols) {
            continue;
        }
        b = rb * ncols + cb;
        if (mask[b] == FCA_WALL) { /* fca_face_exists with a LAND */
            continue;
        }
        s = abs_slope(la, level_of(mask[b], z[b], h[b], p), k < 4 ? p->dx : ddiag);
        if (s > smax) {
            smax = s;
        }
    }
    if (p->boundary == FCA_BND_OPEN &&
        (c == 0 || c == ncols - 1 || r == 0 || r == nrows - 1)) {
        real lv = z[a] - p->S_bnd * p->dx; /* virtual neighbour, h = 0 */
        s = abs_slope(la, lv, p->dx);
        if (s > smax) {
            smax = s;
        }
    }
    /* physics.h:196-216 */
    hh = h[a];
    if (hh < p->tau_h) {
        return p->dt_max;
    }
    nn = n[a];
    if (nn < (real)1e-12) {
        nn = (real)1e-12;
    }
    if (smax < (real)0) {
        smax = (real)0;
    }
    v = ((real)1 / nn) * or_cbrt_sq(hh) * ORACLE_SQRT_SLOPE(smax);
    vcr = (real)sqrt((double)p->g * (double)hh);
    if (v > vcr) {
        v = vcr;
    }
    if (v < (real)1e-6) {
        v = (real)1e-6;
    }
    return p->alpha * p->dx / v;
}

/* model.c:111-156 (no MPI hook). The "any wet cell" flag of the reference
 * only restores dt_max, which the running minimum already starts from. */
real or_compute_dt_base(const FcaGrid *g, const FcaPhys *p, FcaAcc *acc)
{
    int r, c;
    real dtmin = p->dt_max;
    for (r = 0; r < g->nrows; r++) {
        for (c = 0; c < g->ncols; c++) {
            real d = or_cell_dt(g->z, g->h, g->n, g->mask, r, c, g->nrows, g->ncols, p);
            if (d < dtmin) {
                dtmin = d;
            }
        }
    }
    if (dtmin < p->dt_min) {
        dtmin = p->dt_min;
        if (acc) {
            acc->dt_min_hits += 1;
        }
    }
    if (dtmin > p->dt_max) {
        dtmin = p->dt_max;
    }
    return dtmin;
}

/* ---- forcing time line (src/common/rain.c) ----------------------------- */

/* Unit triangular hyetograph of duration T peaking at tp (rain.c:13-28). */
static double tri_rate(double t, double T, double tp)
{
    if (t < 0.0 || t >= T || T <= 0.0) {
        return 0.0;
    }
    if (t < tp) {
        if (tp <= 0.0) {
            return 0.0;
        }
        return (2.0 / T) * (t / tp);
    }
    if (T - tp <= 0.0) {
        return 0.0;
    }
    return (2.0 / T) * ((T - t) / (T - tp));
}

/* rain.c:30-67 */
double or_hyeto_integral(const FcaRunCfg *cfg, double t, double dt)
{
    double t1, T, tp, sum, a, b;
    if (dt <= 0.0) {
        return 0.0;
    }
    t1 = t + dt;
    if (cfg->hyeto == FCA_HYETO_UNIFORM) {
        return dt / 86400.0;
    }
    T = (double)cfg->D_h * 3600.0;
    tp = (double)cfg->phi * T;
    if (t1 <= 0.0 || t >= T) {
        return 0.0;
    }
    if (t < 0.0) {
        t = 0.0;
    }
    if (t1 > T) {
        t1 = T;
    }
    /* exact trapezoids on the linear pieces, split at the peak */
    sum = 0.0;
    a = t;
    while (a < t1 - 1e-15) {
        b = (a < tp && t1 > tp) ? tp : t1;
        sum += 0.5 * (tri_rate(a, T, tp) + tri_rate(b, T, tp)) * (b - a);
        a = b;
    }
    return sum;
}

/* rain.c:69-93: earliest of T_sim, the next day boundary and the next
 * out_every multiple strictly after t (1e-9 s guard). */
double or_next_event_time(const FcaRunCfg *cfg, double t)
{
    double t_next = (double)cfg->T_sim, t_day, t_out, oe;
    const double day = FCA_DAY_S;
    t_day = (floor(t / day) + 1.0) * day;
    if (t_day - t < 1e-9) {
        t_day += day;
    }
    if (t_day < t_next) {
        t_next = t_day;
    }
    if (cfg->out_every > (real)0) {
        oe = (double)cfg->out_every;
        t_out = (floor(t / oe) + 1.0) * oe;
        if (t_out - t < 1e-9) {
            t_out += oe;
        }
        if (t_out < t_next) {
            t_next = t_out;
        }
    }
    return t_next;
}

/* rain.c:95-106 */
int or_forcing_day_index(double t)
{
    int d;
    if (t < 0.0) {
        return 0;
    }
    d = (int)floor(t / FCA_DAY_S);
    return d < 0 ? 0 : d;
}

/* ---- volumes (grid.c:125-141, model.c:102-109, diag.c:5-14) ------------- */

double or_volume_land(const FcaGrid *g, real A)
{
    size_t i, N = (size_t)g->nrows * (size_t)g->ncols;
    double V = 0.0;
    for (i = 0; i < N; i++) {
        if (g->mask[i] == FCA_LAND) {
            V += (double)g->h[i] * (double)A;
        }
    }
    return V;
}

void or_acc_init(FcaAcc *acc, const FcaGrid *g, real A)
{
    memset(acc, 0, sizeof(*acc));
    acc->V0 = or_volume_land(g, A);
    acc->V = acc->V0;
    acc->dt_base = 0;
    acc->force_day = -1;
}

double or_mass_error(const FcaAcc *acc)
{
    double denom = acc->V_rain + acc->V0, num;
    if (denom < 1e-30) {
        denom = 1e-30;
    }
    num = acc->V + acc->V_sea + acc->V_bnd - acc->V_rain - acc->V0 - acc->clamp_err;
    return fabs(num) / denom;
}

/* Float32 raster into real[] (grid.c:143-160, 239-254). */
static int load_f32(const char *path, real *dst, size_t n)
{
    FILE *fp = fopen(path, "rb");
    float *tmp;
    size_t i, got;
    if (!fp) {
        return -1;
    }
    tmp = (float *)malloc(n * sizeof(float));
    if (!tmp) {
        fclose(fp);
        return -1;
    }
    got = fread(tmp, sizeof(float), n, fp);
    fclose(fp);
    if (got != n) {
        free(tmp);
        return -1;
    }
    for (i = 0; i < n; i++) {
        dst[i] = (real)tmp[i];
    }
    free(tmp);
    return 0;
}

/* ---- the step (model.c:321-605) --------------------------------------- */

/* Face arrays of one step, laid out as in fca_types.h:121-125. */
typedef struct Faces {
    real *qx, *qy, *qd1, *qd2, *qo, *s;
} Faces;

/* One of the up to eight faces of a cell, in the reference's gather order
 * E, W, S, N, SE, NW, SW, NE (model.c:185-233, 247-311). `mine` is set when
 * the cell is the face's a side, i.e. positive Q leaves the cell. */
typedef struct FaceRef {
    real *q;
    int nb;
    int mine;
} FaceRef;

static int cell_faces(const Faces *f, int r, int c, int nr, int nc, int K, FaceRef out[8])
{
    int a = r * nc + c, m = 0;
    if (c + 1 < nc) {
        out[m].q = &f->qx[r * (nc - 1) + c], out[m].nb = a + 1, out[m].mine = 1, m++;
    }
    if (c > 0) {
        out[m].q = &f->qx[r * (nc - 1) + c - 1], out[m].nb = a - 1, out[m].mine = 0, m++;
    }
    if (r + 1 < nr) {
        out[m].q = &f->qy[r * nc + c], out[m].nb = a + nc, out[m].mine = 1, m++;
    }
    if (r > 0) {
        out[m].q = &f->qy[(r - 1) * nc + c], out[m].nb = a - nc, out[m].mine = 0, m++;
    }
    if (K == 8) {
        if (r + 1 < nr && c + 1 < nc) {
            out[m].q = &f->qd1[r * (nc - 1) + c], out[m].nb = a + nc + 1, out[m].mine = 1, m++;
        }
        if (r > 0 && c > 0) {
            out[m].q = &f->qd1[(r - 1) * (nc - 1) + c - 1], out[m].nb = a - nc - 1;
            out[m].mine = 0, m++;
        }
        if (r + 1 < nr && c > 0) {
            out[m].q = &f->qd2[r * (nc - 1) + c - 1], out[m].nb = a + nc - 1, out[m].mine = 1;
            m++;
        }
        if (r > 0 && c + 1 < nc) {
            out[m].q = &f->qd2[(r - 1) * (nc - 1) + c], out[m].nb = a - nc + 1, out[m].mine = 0;
            m++;
        }
    }
    return m;
}

int or_step(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, FcaHipStepRec *rec)
{
    const FcaPhys *p = &cfg->phys;
    const int nr = g->nrows, nc = g->ncols, K = p->K;
    const size_t N = (size_t)nr * (size_t)nc;
    const size_t nqx = (size_t)nr * (size_t)(nc > 0 ? nc - 1 : 0);
    const size_t nqy = (size_t)(nr > 0 ? nr - 1 : 0) * (size_t)nc;
    const size_t nqd = (size_t)(nr > 0 ? nr - 1 : 0) * (size_t)(nc > 0 ? nc - 1 : 0);
    const real dx = p->dx, A = p->A;
    const real w = (K == 8) ? ((real)sqrt(2.0) - (real)1) * dx : dx; /* physics.h:37-43 */
    const real ddiag = (real)sqrt(2.0) * dx;                          /* physics.h:45-48 */
    real dt;
    double t_next, remain, F;
    double v_rain = 0, clamp = 0, v_sea = 0, v_bnd = 0;
    long cap_t = 0, cap_a = 0;
    int hit = 0, r, c, cap;
    size_t i;
    Faces f;

    /* multi-day forcing: reload the day's Rday when the day changes
     * (model.c:338-358); a missing file ends the run at T_sim */
    if (cfg->n_force_days > 0) {
        int d = or_forcing_day_index(acc->t);
        if (d >= cfg->n_force_days) {
            d = cfg->n_force_days - 1;
        }
        if (d != acc->force_day) {
            if (load_f32(cfg->force_files[d], g->Rday, N) != 0) {
                acc->t = (double)cfg->T_sim;
                return 0;
            }
            if (cfg->lambda != (real)1) {
                for (i = 0; i < N; i++) {
                    g->Rday[i] *= cfg->lambda;
                }
            }
            acc->force_day = d;
        }
    }

    /* held step, event clamp and rain integral (model.c:363-382) */
    if (acc->step % (p->U > 0 ? p->U : 1) == 0) {
        acc->dt_base = or_compute_dt_base(g, p, acc);
    }
    dt = acc->dt_base;
    t_next = or_next_event_time(cfg, acc->t);
    remain = t_next - acc->t;
    if (remain <= (double)dt) {
        if (remain < 0) {
            remain = 0;
        }
        dt = (real)remain;
        hit = 1;
    }
    if (dt <= (real)0) {
        return 0;
    }
    F = or_hyeto_integral(cfg, acc->t, (double)dt);
    if (rec) {
        rec->t = acc->t;
        rec->F = F;
        rec->dt_base = acc->dt_base;
        rec->dt = dt;
        rec->step = acc->step;
        rec->hit_event = hit;
    }

    f.qx = (real *)calloc(nqx ? nqx : 1, sizeof(real));
    f.qy = (real *)calloc(nqy ? nqy : 1, sizeof(real));
    f.qd1 = (real *)calloc(nqd ? nqd : 1, sizeof(real));
    f.qd2 = (real *)calloc(nqd ? nqd : 1, sizeof(real));
    f.qo = (real *)calloc(N > 0 ? 4 * N : 1, sizeof(real));
    f.s = (real *)calloc(N ? N : 1, sizeof(real));
    if (!f.qx || !f.qy || !f.qd1 || !f.qd2 || !f.qo || !f.s) {
        free(f.qx), free(f.qy), free(f.qd1), free(f.qd2), free(f.qo), free(f.s);
        return -1;
    }

    /* raw fluxes on every face (model.c:393-482) */
#define FACE(A_, B_, D_) \
    or_face_flux(g->z[A_], g->h[A_], g->n[A_], g->mask[A_], g->z[B_], g->h[B_], g->n[B_], \
                 g->mask[B_], (D_), w, dt, K, A, p, &cap)
    for (r = 0; r < nr; r++) {
        for (c = 0; c < nc; c++) {
            int a = r * nc + c;
            if (c + 1 < nc) {
                f.qx[r * (nc - 1) + c] = FACE(a, a + 1, dx);
                cap_t++, cap_a += cap;
            }
            if (r + 1 < nr) {
                f.qy[r * nc + c] = FACE(a, a + nc, dx);
                cap_t++, cap_a += cap;
            }
            if (K == 8 && r + 1 < nr && c + 1 < nc) {
                f.qd1[r * (nc - 1) + c] = FACE(a, a + nc + 1, ddiag);   /* SE */
                cap_t++, cap_a += cap;
                f.qd2[r * (nc - 1) + c] = FACE(a + 1, a + nc, ddiag);   /* SW */
                cap_t++, cap_a += cap;
            }
            /* open-boundary virtual faces: z_b = z_a - S_bnd*dx, h_b = 0,
             * n_b = n_a, outflow only (model.c:166-177, 442-479) */
            if (p->boundary == FCA_BND_OPEN && g->mask[a] == FCA_LAND) {
                int slot_on[4] = {c == 0, c == nc - 1, r == 0, r == nr - 1}, k;
                for (k = 0; k < 4; k++) {
                    real zb, q;
                    if (!slot_on[k]) {
                        continue;
                    }
                    zb = g->z[a] - p->S_bnd * p->dx;
                    q = or_face_flux(g->z[a], g->h[a], g->n[a], g->mask[a], zb, (real)0,
                                     g->n[a], FCA_LAND, p->dx, w, dt, K, A, p, &cap);
                    if (q < (real)0) {
                        q = (real)0;
                    }
                    f.qo[4 * (size_t)a + k] = q;
                    cap_t++, cap_a += cap;
                }
            }
        }
    }
#undef FACE
    acc->cap_total += cap_t;
    acc->cap_active += cap_a;

    /* outgoing volume and limiter per cell (model.c:180-238, 488-502) */
    for (r = 0; r < nr; r++) {
        for (c = 0; c < nc; c++) {
            FaceRef fr[8];
            int a = r * nc + c, m = cell_faces(&f, r, c, nr, nc, K, fr), k;
            real out = 0;
            const real *qo = &f.qo[4 * (size_t)a];
            for (k = 0; k < m; k++) {
                real q = *fr[k].q;
                if (fr[k].mine ? (q > 0) : (q < 0)) {
                    out += fr[k].mine ? q : -q;
                }
            }
            out += qo[0] + qo[1] + qo[2] + qo[3];
            out = out * dt;
            f.s[a] = (g->mask[a] == FCA_LAND) ? or_limiter_s(out, g->h[a], A) : (real)1;
        }
    }

    /* scale every face by the limiter of its upwind cell (model.c:511-559) */
    for (r = 0; r < nr; r++) {
        for (c = 0; c < nc; c++) {
            int a = r * nc + c;
            real q;
            if (c + 1 < nc) {
                q = f.qx[r * (nc - 1) + c];
                f.qx[r * (nc - 1) + c] = q * ((q >= 0) ? f.s[a] : f.s[a + 1]);
            }
            if (r + 1 < nr) {
                q = f.qy[r * nc + c];
                f.qy[r * nc + c] = q * ((q >= 0) ? f.s[a] : f.s[a + nc]);
            }
            if (K == 8 && r + 1 < nr && c + 1 < nc) {
                q = f.qd1[r * (nc - 1) + c];
                f.qd1[r * (nc - 1) + c] = q * ((q >= 0) ? f.s[a] : f.s[a + nc + 1]);
                q = f.qd2[r * (nc - 1) + c];
                f.qd2[r * (nc - 1) + c] = q * ((q >= 0) ? f.s[a + 1] : f.s[a + nc]);
            }
        }
    }
    for (i = 0; i < N; i++) {
        real su = f.s[i];
        f.qo[4 * i + 0] *= su;
        f.qo[4 * i + 1] *= su;
        f.qo[4 * i + 2] *= su;
        f.qo[4 * i + 3] *= su;
    }

    /* net inflow, rain and depth update (model.c:241-319, 561-597) */
    for (r = 0; r < nr; r++) {
        for (c = 0; c < nc; c++) {
            FaceRef fr[8];
            int a = r * nc + c, m = cell_faces(&f, r, c, nr, nc, K, fr), k;
            int land = g->mask[a] == FCA_LAND;
            const real *qo = &f.qo[4 * (size_t)a];
            real net = 0, qos, rain;
            double sea = 0;
            for (k = 0; k < m; k++) {
                real q = *fr[k].q;
                if (fr[k].mine) {
                    net -= q;
                    if (land && g->mask[fr[k].nb] == FCA_SEA && q > 0) {
                        sea += (double)dt * (double)q;
                    }
                } else {
                    net += q;
                    if (land && g->mask[fr[k].nb] == FCA_SEA && q < 0) {
                        sea += (double)dt * (double)(-q);
                    }
                }
            }
            qos = qo[0] + qo[1] + qo[2] + qo[3];
            net -= qos;
            v_sea += sea;
            v_bnd += (double)dt * (double)qos;
            if (!land) {
                continue;
            }
            rain = g->C[a] * g->Rday[a] * (real)F;
            v_rain += (double)A * (double)rain;
            g->h[a] += dt * (net / A) + rain;
            if (g->h[a] < (real)0) {
                clamp += (double)(-g->h[a]) * (double)A;
                g->h[a] = (real)0;
            }
            if (g->h[a] > g->hmax[a]) {
                g->hmax[a] = g->h[a];
                g->tmax[a] = (real)(acc->t + (double)dt);
            }
        }
    }
    acc->V_rain += v_rain;
    acc->clamp_err += clamp;
    acc->V_sea += v_sea;
    acc->V_bnd += v_bnd;

    /* land exactly on the event (model.c:599-601) */
    acc->t = hit ? t_next : acc->t + (double)dt;
    acc->step += 1;

    free(f.qx), free(f.qy), free(f.qd1), free(f.qd2), free(f.qo), free(f.s);
    return 1;
}

/* model.c:607-643. One deviation: with cfg->steps > 0 the reference loops
 * forever once no step can be taken (dt <= 0 at T_sim, model.c:379-381 with
 * the exit at 618-620 disabled); here the loop ends instead. */
int or_run(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, FcaHipStepRec *log, int log_cap,
           int *n_logged)
{
    int max_steps = cfg->steps > 0 ? cfg->steps : 2000000000, n = 0;
    while (acc->step < max_steps) {
        double t_before = acc->t;
        FcaHipStepRec rec;
        int rc;
        if (cfg->steps <= 0 && acc->t >= (double)cfg->T_sim - 1e-15) {
            break;
        }
        rc = or_step(g, cfg, acc, &rec);
        if (rc < 0) {
            return -1;
        }
        if (rc == 1) {
            if (log && n < log_cap) {
                log[n] = rec;
            }
            n++;
        }
        if (cfg->steps <= 0 && acc->t <= t_before) {
            break;
        }
        if (cfg->steps > 0 && rc == 0) {
            break;
        }
    }
    acc->V = or_volume_land(g, cfg->phys.A);
    if (n_logged) {
        *n_logged = n;
    }
    return 0;
}

/* ---- libm reference values over float bit ranges ----------------------- */

/* out[i] = the reference's float-mode expression for the float with bit
 * pattern first + i: kind 0 (float)cbrt((double)x*(double)x) (physics.h:19-24),
 * kind 1 (float)sqrt((double)x) (physics.h:106), kind 2 the critical speed
 * (float)sqrt((double)g*(double)x) (physics.h:108). glibc libm; these feed the
 * exhaustive checks of the device implementations. */
void or_libm_range(int kind, uint32_t first, uint32_t count, float g, float *out)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (long)count; i++) {
        uint32_t b = first + (uint32_t)i;
        float x;
        memcpy(&x, &b, 4);
        if (kind == 0) {
            out[i] = (float)cbrt((double)x * (double)x);
        } else if (kind == 1) {
            out[i] = (float)sqrt((double)x);
        } else {
            out[i] = (float)sqrt((double)g * (double)x);
        }
    }
}

/* Sign of m^3 - x, exact for m with <= 26 significant bits (m^2 exact, m^3
 * as a double-double via fma) and x a double. */
static int cube_sign(double m, double x)
{
    double m2 = m * m, p = m2 * m, e = fma(m2, m, -p), d = (p - x) + e;
    return (d > 0) - (d < 0);
}

static float f32_bits_add(float f, int k)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    b += (uint32_t)k;
    memcpy(&f, &b, 4);
    return f;
}

/* Counts the floats in [first, first+count) for which the reference's
 * (float)cbrt((double)x*x) differs from the correctly rounded x^(2/3):
 * glibc's result is within one float ulp, and exact cube comparisons at the
 * two neighbouring midpoints decide the correctly rounded value. */
long or_cbrt_sq_cr_mismatches(uint32_t first, uint32_t count)
{
    long i, bad = 0;
#pragma omp parallel for schedule(dynamic, 65536) reduction(+ : bad)
    for (i = 0; i < (long)count; i++) {
        uint32_t b = first + (uint32_t)i;
        float x, g, cr, up, dn;
        double sq;
        memcpy(&x, &b, 4);
        sq = (double)x * (double)x;
        g = (float)cbrt(sq);
        up = f32_bits_add(g, 1);
        dn = f32_bits_add(g, -1);
        cr = g;
        if (cube_sign(((double)g + (double)up) * 0.5, sq) < 0) {
            cr = up;
        } else if (cube_sign(((double)dn + (double)g) * 0.5, sq) > 0) {
            cr = dn;
        }
        bad += cr != g;
    }
    return bad;
}

/* ---- ensemble counts (src/hybrid/ens_common.c:166-188) ----------------- */

void or_ens_count(const FcaGrid *g, int nthr, const double *thr, uint32_t *counts,
                  long *flooded, double *max_depth)
{
    size_t i, N = (size_t)g->nrows * (size_t)g->ncols;
    int j;
    *max_depth = 0;
    for (j = 0; j < nthr; j++) {
        flooded[j] = 0;
    }
    for (i = 0; i < N; i++) {
        double hm;
        if (g->mask[i] != FCA_LAND) {
            continue;
        }
        hm = (double)g->hmax[i];
        if (hm > *max_depth) {
            *max_depth = hm;
        }
        for (j = 0; j < nthr; j++) {
            if (hm >= thr[j]) {
                counts[(size_t)j * N + i]++;
                flooded[j]++;
            }
        }
    }
}
