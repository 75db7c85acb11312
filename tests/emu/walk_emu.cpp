/*
 * walk_emu.cpp -- host emulation of the fused step, for the CPU test suite
 * only (never shipped, never a fallback: the product path is the HIP library).
 *
 * It runs floodca_amd/csrc/fca_walk.h's walk_strip with 64 or 128 lanes in lockstep
 * (pull_left/pull_right are array shifts with 0 at the wavefront edge, as the
 * DPP shifts read 0), tile by tile in the order of the kernel's workgroups,
 * with the kernel's controller (ctrl_eval) and candidate rules. The CPU
 * suite checks it bit for bit against the oracle, which pins the walk's
 * dataflow (halos, ring, edges, K=8 diagonals) before any GPU run.
 */
#include "fca_walk.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef real R;

template <int W>
struct CpuLanes {
    static constexpr int WIDTH = W;
    WalkLane<R> s[W];
    template <class F>
    void each(F f)
    {
        for (int l = 0; l < W; l++) {
            f(s[l], l);
        }
    }
    template <class T>
    void pull_right(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        T tmp[W];
        for (int l = 0; l < W; l++) {
            tmp[l] = (l + 1 < W) ? s[l + 1].*src : (T)0;
        }
        for (int l = 0; l < W; l++) {
            s[l].*dst = tmp[l];
        }
    }
    template <class T>
    void pull_left(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        T tmp[W];
        for (int l = 0; l < W; l++) {
            tmp[l] = (l > 0) ? s[l - 1].*src : (T)0;
        }
        for (int l = 0; l < W; l++) {
            s[l].*dst = tmp[l];
        }
    }
    void strip_begun() {}
    /* the capped faces of written columns where c holds, per lane */
    void count_caps(WalkLane<R> &t, bool c)
    {
        t.cap += (c && t.outcol) ? 1 : 0;
    }
    /* the device's range test of the LAND form's rows (float) */
    uint64_t out_of_range_lanes()
    {
        uint64_t b = 0;
        if (sizeof(R) == 4) {
            for (int l = 0; l < W; l++) {
                b |= (uint64_t)(!(fabs(s[l].lv0) <= 0x1p20) || !(fabs(s[l].h0) <= 0x1p14)) << (l & 63);
            }
        }
        return b;
    }
    long form_rows[3];
    void count_form(int form)
    {
        form_rows[form]++;
    }
    void loads_settled() {}
    /* nonzero exactly when some lane qualifies (lanes folded onto 64 bits) */
    uint64_t land_lanes(int WalkLane<R>::*m)
    {
        uint64_t b = 0;
        for (int l = 0; l < W; l++) {
            b |= (uint64_t)(s[l].*m == FCA_LAND) << (l & 63);
        }
        return b;
    }
    uint64_t non_land_lanes(int WalkLane<R>::*m)
    {
        uint64_t b = 0;
        for (int l = 0; l < W; l++) {
            b |= (uint64_t)(s[l].*m != FCA_LAND) << (l & 63);
        }
        return b;
    }
    bool any(int WalkLane<R>::*f)
    {
        for (int l = 0; l < W; l++) {
            if (s[l].*f) {
                return true;
            }
        }
        return false;
    }
    void wave_max_to(R WalkLane<R>::*dst, R WalkLane<R>::*src)
    {
        R v = s[0].*src;
        for (int l = 1; l < W; l++) {
            v = (s[l].*src > v) ? s[l].*src : v;
        }
        for (int l = 0; l < W; l++) {
            s[l].*dst = v;
        }
    }
};

/* -1: the LAND form where the data allow it (as on the device); 0: never
 * (tests compare both forms). */
static int emu_land_form = -1;
extern "C" void emu_set_land_form(int v)
{
    emu_land_form = v;
}

/* 1: skip quiet tiles as the kernel does (the default); 0: walk every tile. */
static int emu_skip = 1;
static long emu_skipped;
extern "C" void emu_set_skip(int v)
{
    emu_skip = v;
}
/* Lanes of the emulated wavefront: 64 (one column per work-item) or 128
 * (two columns per work-item); tiles write width - 6 columns. */
static int emu_width = 64;
extern "C" void emu_set_width(int w)
{
    emu_width = (w == 128) ? 128 : 64;
}

/* Tiles skipped by the last emu_run. */
extern "C" long emu_skipped_tiles(void)
{
    return emu_skipped;
}

/* The kernel's rain-free tile map (k_tile_norain). */
static void tile_norain(const R *cr, const uint8_t *mask, int nr, int nc, int strip_rows,
                        int ngroups, int ntiles, int outc, uint8_t *out)
{
    for (int t = 0; t < ntiles; t++) {
        const int strip = t / ngroups, grp = t % ngroups;
        const int r0 = strip * strip_rows, r1 = (r0 + strip_rows < nr) ? r0 + strip_rows : nr;
        int rain = 0, land = 0;
        for (int r = r0; r < r1; r++) {
            for (int l = 0; l < outc; l++) {
                const int c = grp * outc + l;
                if (c < nc) {
                    const size_t i = (size_t)r * nc + c;
                    land |= mask[i] == FCA_LAND;
                    rain |= mask[i] == FCA_LAND && !(cr[i] == (R)0);
                }
            }
        }
        out[t] = (uint8_t)((rain ? 0 : 1) | (land ? 2 : 0));
    }
}

/* Rows walked in the general, the LAND and the no-LAND form by the last
 * emu_run. */
static long emu_rows[3];
extern "C" void emu_form_rows(long *general, long *land, long *noland)
{
    *general = emu_rows[0];
    *land = emu_rows[1];
    *noland = emu_rows[2];
}

static int load_day(const char *path, R *dst, size_t n, R lambda)
{
    FILE *fp = fopen(path, "rb");
    float *tmp;
    size_t i, got;
    if (!fp) {
        return -1;
    }
    tmp = (float *)malloc(n * sizeof(float));
    got = tmp ? fread(tmp, sizeof(float), n, fp) : 0;
    fclose(fp);
    if (got != n) {
        free(tmp);
        return -1;
    }
    for (i = 0; i < n; i++) {
        dst[i] = (R)tmp[i];
        if (lambda != (R)1) {
            dst[i] *= lambda;
        }
    }
    free(tmp);
    return 0;
}

/* Runs the fused step like fca_hip_run, on the host, with wavefronts of W
 * lanes. g is updated in place (h, hmax, tmax, Rday on reloads); acc as
 * fca_hip_run sets it. */
template <int W>
static int emu_run_w(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, int strip_rows,
                     FcaHipStepRec *rec, int rec_cap, int *nrec)
{
    const int outc = FCA_TILE_OUT(W);
    const int nr = g->nrows, nc = g->ncols, K8 = cfg->phys.K == 8;
    const size_t N = (size_t)nr * (size_t)nc;
    PhysK<R> p = make_phys<R>(cfg->phys);
    FcaCtlParams ctl;
    FcaDevCtrl c, nx;
    R *h[2], *cr, cand = (R)INFINITY;
    long faces = 0;
    int ngroups, nstrips, r, cc, hb = 0, nrc = 0;
    size_t i;
    const FcaAcc acc0 = *acc;

    if (strip_rows <= 0) {
        strip_rows = 256;
    }
    ctl.T_sim = (double)cfg->T_sim, ctl.out_every = (double)cfg->out_every;
    ctl.D_h = (double)cfg->D_h, ctl.phi = (double)cfg->phi;
    ctl.hyeto = cfg->hyeto, ctl.steps = cfg->steps, ctl.U = cfg->phys.U;
    ctl.n_force_days = cfg->n_force_days;
    h[0] = (R *)malloc(N * sizeof(R));
    h[1] = (R *)malloc(N * sizeof(R));
    cr = (R *)malloc(N * sizeof(R));
    if (!h[0] || !h[1] || !cr) {
        free(h[0]), free(h[1]), free(cr);
        return -1;
    }
    memcpy(h[0], g->h, N * sizeof(R));
    memcpy(h[1], g->h, N * sizeof(R));
    for (i = 0; i < N; i++) {
        cr[i] = g->C[i] * g->Rday[i];
    }
    for (r = 0; r < nr; r++) {
        for (cc = 0; cc < nc; cc++) {
            R d = cell_dt_at<R>(g->z, h[0], g->n, g->mask, r, cc, nr, nc, p);
            size_t a = (size_t)r * nc + cc;
            if (g->mask[a] == FCA_LAND && h[0][a] >= p.tau_h && d == d && d < cand) {
                cand = d;
            }
            if (g->mask[a] == FCA_LAND && p.open) {
                faces += (cc == 0) + (cc == nc - 1) + (r == 0) + (r == nr - 1);
            }
        }
    }
    faces += (long)nr * (nc - 1) + (long)(nr - 1) * nc + (K8 ? 2L * (nr - 1) * (nc - 1) : 0);
    memset(&c, 0, sizeof(c));
    c.t = acc->t, c.step = acc->step, c.force_day = acc->force_day;
    c.state = FCA_CTL_RUN, c.dt_base = (double)acc->dt_base;
    /* the walk's LAND form depends on the data only; host arithmetic needs
     * none of the device's division flags */
    p.fast = emu_land_form == 0 ? 0 : fca_land_form_ok(&cfg->phys, g->n, g->mask, N);
    ngroups = (nc + outc - 1) / outc;
    nstrips = (nr + strip_rows - 1) / strip_rows;
    uint8_t *tflags = (uint8_t *)calloc((size_t)2 * ngroups * nstrips, 1);
    uint8_t *tnorain = (uint8_t *)calloc((size_t)ngroups * nstrips, 1);
    if (!tflags || !tnorain) {
        free(tflags), free(tnorain), free(h[0]), free(h[1]), free(cr);
        return -1;
    }
    tile_norain(cr, g->mask, nr, nc, strip_rows, ngroups, ngroups * nstrips, outc, tnorain);
    emu_skipped = 0;

    for (;;) {
        StepVals<R> sv;
        R next = (R)INFINITY;
        if (c.state == FCA_CTL_DONE) {
            break;
        }
        if (c.state == FCA_CTL_PAUSE) {
            if (load_day(cfg->force_files[c.want_day], g->Rday, N, cfg->lambda) != 0) {
                c.t = (double)cfg->T_sim;
                c.state = FCA_CTL_DONE;
                break;
            }
            for (i = 0; i < N; i++) {
                cr[i] = g->C[i] * g->Rday[i];
            }
            tile_norain(cr, g->mask, nr, nc, strip_rows, ngroups, ngroups * nstrips, outc, tnorain);
            c.force_day = c.want_day;
            c.state = FCA_CTL_RUN;
            continue;
        }
        if (cfg->steps > 0 && c.step >= cfg->steps) {
            break;
        }
        ctrl_eval<R>(c, cand, ctl, p, sv, nx);
        if (!sv.active) {
            const int stuck = nx.state == FCA_CTL_RUN;
            c = nx;
            if (stuck) {
                break; /* no step possible: the reference would spin */
            }
            continue;
        }
        if (rec && nrc < rec_cap) {
            FcaHipStepRec q;
            q.t = c.t, q.F = sv.F, q.dt_base = sv.dt_base, q.dt = sv.dt;
            q.step = c.step, q.hit_event = sv.hit;
            rec[nrc] = q;
        }
        nrc++;
        {
            WalkArgs<R> wa;
            wa.z = g->z, wa.n = g->n, wa.cr = cr, wa.mask = g->mask;
            wa.h_in = h[hb], wa.h_out = h[1 - hb], wa.hmax = g->hmax, wa.tmax = g->tmax;
            wa.nr = nr, wa.nc = nc, wa.strip_rows = strip_rows, wa.p = p, wa.sv = sv;
            wa.lgg = speed_log_half_g<R>(p.g);
            for (int t = 0; t < ngroups * nstrips; t++) {
                static CpuLanes<W> lp;
                if (t == 0 && c.step == acc0.step) {
                    lp.form_rows[0] = lp.form_rows[1] = lp.form_rows[2] = 0;
                }
                const int strip = t / ngroups, grp = t % ngroups;
                const int r0 = strip * strip_rows;
                const int r1 = (r0 + strip_rows < nr) ? r0 + strip_rows : nr;
                const int ntl = ngroups * nstrips;
                const uint8_t *prev = tflags + (size_t)((c.step + 1) & 1) * ntl;
                uint8_t *cur = tflags + (size_t)(c.step & 1) * ntl;
                walk_acc_init<R>(lp);
                if (emu_skip && tile_skippable(prev, t, ngroups, nstrips,
                                               sv.rain_f == (R)0 || (tnorain[t] & 1))) {
                    cur[t] = FCA_TILE_DRY | FCA_TILE_SAME;
                    emu_skipped++;
                    continue;
                }
                if (K8 && emu_skip) {
                    walk_strip<R, 1, 1>(lp, wa, r0, r1, grp * outc - FCA_HALO);
                } else if (K8) {
                    walk_strip<R, 1, 0>(lp, wa, r0, r1, grp * outc - FCA_HALO);
                } else if (emu_skip) {
                    walk_strip<R, 0, 1>(lp, wa, r0, r1, grp * outc - FCA_HALO);
                } else {
                    walk_strip<R, 0, 0>(lp, wa, r0, r1, grp * outc - FCA_HALO);
                }
                cur[t] = (uint8_t)((lp.any(&WalkLane<R>::qwet) ? 0 : FCA_TILE_DRY) |
                                   (lp.any(&WalkLane<R>::qchg) ? 0 : FCA_TILE_SAME));
                walk_acc_finish<R>(lp, p);
                emu_rows[0] = lp.form_rows[0], emu_rows[1] = lp.form_rows[1];
                emu_rows[2] = lp.form_rows[2];
                for (int l = 0; l < W; l++) {
                    acc->V_rain += lp.s[l].v_rain;
                    acc->clamp_err += lp.s[l].clamp;
                    acc->V_sea += lp.s[l].v_sea;
                    acc->V_bnd += lp.s[l].v_bnd;
                    acc->cap_active += (long)lp.s[l].cap;
                    if (lp.s[l].dtc < next) {
                        next = lp.s[l].dtc;
                    }
                }
            }
        }
        nx.hbuf = 1 - c.hbuf;
        hb = nx.hbuf;
        cand = sv.need_cand ? next : (R)INFINITY;
        c = nx;
    }
    memcpy(g->h, h[hb], N * sizeof(R));
    /* the walk stores hmax and tmax only when a cell stops rising: the cells
     * still rising take their depth and the last step's tmax (as the host
     * library's finalise step) */
    if (c.armed) {
        for (i = 0; i < N; i++) {
            if (g->mask[i] == FCA_LAND && g->h[i] > g->hmax[i]) {
                g->hmax[i] = g->h[i];
                g->tmax[i] = (R)c.tmax_last;
            }
        }
    }
    acc->cap_total = acc0.cap_total + faces * (long)(c.step - acc0.step);
    acc->dt_min_hits = acc0.dt_min_hits + (long)c.dt_min_hits;
    acc->t = c.t;
    acc->step = c.step;
    acc->dt_base = (R)c.dt_base;
    acc->force_day = c.force_day;
    acc->V = 0.0;
    for (i = 0; i < N; i++) {
        if (g->mask[i] == FCA_LAND) {
            acc->V += (double)g->h[i] * (double)p.A;
        }
    }
    if (nrec) {
        *nrec = nrc;
    }
    free(h[0]), free(h[1]), free(cr), free(tflags), free(tnorain);
    return 0;
}

extern "C" int emu_run(FcaGrid *g, const FcaRunCfg *cfg, FcaAcc *acc, int strip_rows,
                       FcaHipStepRec *rec, int rec_cap, int *nrec)
{
    return emu_width == 128 ? emu_run_w<128>(g, cfg, acc, strip_rows, rec, rec_cap, nrec)
                            : emu_run_w<64>(g, cfg, acc, strip_rows, rec, rec_cap, nrec);
}
