/*
 * fca_walk.h -- the fused WCA step as a column walk.
 *
 * One wavefront owns W adjacent columns (W = 64, one column per work-item,
 * or 128, two adjacent columns per work-item) and walks down a strip of rows.
 * Lane l (0..W-1) holds column c = c_first + l; the three lanes at each side
 * are halo, so a wavefront writes W-6 columns. At row r the walk
 *   A. reads row r (z, h, n, mask) and forms the faces that touch it: the
 *      east face of row r, the face between rows r-1 and r and the two K=8
 *      diagonals between them (physics.h:51-125, model.c:393-482);
 *   B. sums the outflow of row r-1, with its open-boundary faces, and its
 *      limiter s (model.c:158-238, 488-502);
 *   C. updates row r-2: upwind-scaled net inflow, rain, clamp, hmax/tmax and
 *      the FP64 balance terms (model.c:241-319, 511-597);
 *      hmax and tmax are stored only when a cell stops rising (see below);
 *   D. bounds the next step's stable time step on row r-3 from the new
 *      depths (model.c:111-156, physics.h:158-217).
 * Rows r-1..r-4 stay in a register ring, so every field crosses HBM once per
 * step; horizontal neighbours come from the adjacent lanes.
 *
 * Two forms of a row. Where rows r-4..r are LAND on all W lanes, away from
 * the grid's edges, a row takes the LAND form: no wall, sea or open-boundary
 * terms, levels formed once per cell. Any other row takes the general form,
 * which keeps every rule. Both evaluate the same expressions in the same
 * order, so the choice never changes a result.
 *
 * Stage D keeps, per wavefront, the largest exact cell speed found so far
 * (vthr). A cell whose speed is bounded above by vthr cannot lower the
 * step's minimum time step (alpha*dx / v is non-increasing in v), so its
 * exact speed is only evaluated when a cheap upper bound exceeds vthr.
 *
 * The walk is a template over a lane policy LP of LP::WIDTH lanes: on the
 * GPU a lane is one work-item (or one of a work-item's two columns) and
 * pull_left/pull_right are DPP wavefront shifts; the CPU emulation runs the
 * lanes in lockstep with array shifts, so the CPU test suite checks the same
 * dataflow bit for bit against the oracle.
 *
 * hmax and tmax. The reference sets hmax = h and tmax = t + dt whenever a
 * step leaves h > hmax (model.c:585-590), which under steady rain is every
 * LAND cell on every step. The walk stores them only when a cell stops
 * rising, keeping the invariant that after any step the true hmax is
 * max(stored hmax, h) and that h > stored hmax exactly when the cell rose
 * in that step (its true tmax then being that step's t + dt). So a cell with
 * h_old > stored hmax rose in the previous step; if the new depth does not
 * exceed h_old it has stopped, and h_old and the previous step's tmax value
 * are stored. The first step after an upload makes no such store (the
 * caller's h may exceed its hmax before any step); the host finalises the
 * cells still rising (hmax = h, tmax = the last step's value) before any
 * read of hmax or tmax.
 *
 * Lane validity: faces are valid on lanes 0..W-2, limiters on 1..W-2, new
 * depths on 2..W-3 and time-step candidates on 3..W-4, which are the lanes a
 * wavefront writes. Values a lane pulls from outside the wavefront only
 * reach lanes that are not valid.
 */
#ifndef FCA_WALK_H
#define FCA_WALK_H

#include "fca_device.h"

#define FCA_WAVE 64
#define FCA_HALO 3
/* columns a wavefront of `width` lanes writes */
#define FCA_TILE_OUT(width) ((width) - 2 * FCA_HALO)
#define FCA_OUTC FCA_TILE_OUT(FCA_WAVE)

/* Logarithms of a cell's depth and roughness in the speed test of stage D
 * (speed_logs, below). */
template <class R>
struct SpeedLogs {
    R lh, lcr, lnn;
};

/* Inputs of one launch of the fused step. */
template <class R>
struct WalkArgs {
    const R *z, *n, *cr; /* cr = C*Rday, rounded as (C*Rday) in rain = C*Rday*F */
    const R *h_in;
    R *h_out, *hmax, *tmax;
    const uint8_t *mask;
    int nr, nc;
    int strip_rows;
    PhysK<R> p;
    StepVals<R> sv;
    R lgg; /* (1/2) lg g for speed_may_exceed_fast (speed_log_half_g) */
};

/* Per-lane state of the walk (registers on the GPU). Suffix k = row r-k. */
template <class R>
struct WalkLane {
    int c, colok, outcol;
    uint32_t boff; /* byte offset of column c in a row of R */
    R z0, z1, z2, z3, z4;
    int m0, m1, m2, m3, m4;
    R h0, h1, h2;
    R n0, n1, n2, n3;
    R lv0, lv1; /* z+h of rows r, r-1 (meaningful on LAND cells) */
    R hn2, hn3, hn4;
    R qx0, qx1, qx2;
    R qy1, qy2, qy3;       /* faces (r-1 -> r), (r-2 -> r-1), (r-3 -> r-2) */
    R qd1_1, qd1_2, qd1_3; /* SE diagonal faces (q,c)->(q+1,c+1), q = r-1.. */
    R qd2_1, qd2_2, qd2_3; /* SW diagonal faces (q,c+1)->(q+1,c), q = r-1.. */
    R s1, s2, s3;
    /* values pulled from neighbour lanes */
    R xz, xh, xn, xl, xq, xr, xs, xl1, xt, xu, xv, xw, xy, xa;
    int xm, xm2, xm3, xm4;
    R dz1, dz2, dz3, dz4, dh1, dh2, dh3, dh4;
    int dm1, dm2, dm3, dm4;
    /* prefetched: mask of rows r+1 and r+2, z/h/n of row r+1, cr/hmax of row
     * r-2 (this row's update) and of row r-1 (the next one) */
    int pm1, pm2;
    R pz, ph, pn, pcr, phm, pcr1, phm1;
    /* stores of row r-3 deferred to the next row's memory block: bit 0 h,
     * bit 1 hmax (= hst, the cell's depth before the step) and tmax */
    int stq;
    R hst;
    /* this tile's output cells: some depth >= tau_h after the step (qwet),
     * some h, hmax or tmax changed by the step (qchg) */
    int qwet, qchg;
    /* accumulators of this lane's output cells; v_rain and clamp hold sums
     * of depths (rain, and -h of clamped cells), scaled by A once the walk
     * ends (walk_acc_finish): each product (double)A*(double)x of two floats
     * is exact, so the scaling changes only the sums' roundings. cap counts
     * capped faces per lane, capw per wavefront where the lane policy counts
     * them with ballots (count_caps) */
    double v_rain, clamp, v_sea, v_bnd;
    long long cap, capw;
    R dtc, vmax, vthr;
    R lthr; /* log2 form of vthr for speed_may_exceed_fast (vthr and lthr are
             * the same on every lane of a wavefront) */
    /* stage D: largest |level difference| per face length, and whether the
     * exact speed is needed */
    R dax, dad;
    /* stage D of the LAND form: levels z + h (new depth) of rows r-2, r-3,
     * r-4 on this lane, and row r-3's level pulled from the east and west */
    R dl2, dl3, dl4, dle, dlw;
    R cnet; /* stage C of the LAND form (K = 4): net after the E and W terms */
    SpeedLogs<R> dlg; /* its speed test's logarithms (speed_logs) */

    int dneed;
};

/* Loads the element at byte offset boff of a row: a uniform row address plus
 * a 32-bit lane offset. Every caller passes a wavefront-uniform row. On the
 * device there are two forms. With SB false, a plain global access: the
 * compiler keeps each field's 64-bit lane address (base + boff) in two
 * vector registers and adds the row's offset with one vector add per access.
 * With SB true, the row's address is pinned to two scalar registers and boff
 * passed as the access's 32-bit vector offset: no vector register holds an
 * address, at the cost of scalar work per row. (The empty asm statements pin
 * the row to scalar registers and hide where boff comes from, so that the
 * compiler cannot hoist a 64-bit lane address out of the row loop; the cast
 * to the global address space restores what the asm hides, without which the
 * access would be a flat one.) The plain form is the faster wherever the walk
 * fits its occupancy anyway; the scalar base frees the registers the K = 8
 * walks need to run one more wavefront per SIMD. */
template <bool SB, class T>
FCA_HDI T ld_row(const T *row, uint32_t boff)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (SB) {
        typedef const __attribute__((address_space(1))) char gchar;
        typedef const __attribute__((address_space(1))) T gT;
        __asm__("" : "+s"(row));
        __asm__("" : "+v"(boff));
        return *(gT *)((gchar *)row + boff);
    }
#endif
    return *(const T *)((const char *)row + boff);
}
template <bool SB, class T>
FCA_HDI void st_row(T *row, uint32_t boff, T v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (SB) {
        typedef __attribute__((address_space(1))) char gchar;
        typedef __attribute__((address_space(1))) T gT;
        __asm__("" : "+s"(row));
        __asm__("" : "+v"(boff));
        *(gT *)((gchar *)row + boff) = v;
        return;
    }
#endif
    *(T *)((char *)row + boff) = v;
}

/* What a row of the walk needs besides the lane state. */
struct WalkRow {
    int r, r_out0, r_out1, c_first;
    int edge_cols; /* the wavefront holds column 0 or nc-1 */
};

/* Quiet tiles. A tile whose output cells and their neighbours all hold less
 * than tau_h of water, and on which no rain falls this step, is left exactly
 * as it is by the step: every face touching its cells is dead (hf < tau_h for
 * two such cells, for a sea or open face too), so net = 0 and
 * h + (dt*(0/A) + 0) = h, no hmax rises, no balance term or time step comes
 * from it. Such a tile need not be walked, provided the depth buffer the
 * step would write already holds its cells' values, i.e. the tile did not
 * change in the previous step either. Per tile and step, bit 0: all output
 * depths < tau_h after the step; bit 1: the step changed none of them. */
#define FCA_TILE_DRY 1
#define FCA_TILE_SAME 2

/* Whether tile t may skip this step given the flags of the previous step
 * (tiles of ngroups columns, nstrips rows) and whether no rain falls on it. */
FCA_HDI int tile_skippable(const uint8_t *prev, int t, int ngroups, int nstrips, int no_rain)
{
    const int st = t / ngroups, gr = t - st * ngroups;
    if (!no_rain || (prev[t] & (FCA_TILE_DRY | FCA_TILE_SAME)) != (FCA_TILE_DRY | FCA_TILE_SAME)) {
        return 0;
    }
    for (int ds = -1; ds <= 1; ds++) {
        for (int dg = -1; dg <= 1; dg++) {
            const int s2 = st + ds, g2 = gr + dg;
            if (s2 >= 0 && s2 < nstrips && g2 >= 0 && g2 < ngroups &&
                !(prev[s2 * ngroups + g2] & FCA_TILE_DRY)) {
                return 0;
            }
        }
    }
    return 1;
}

/* Upper bound test for stage D: whether a wet LAND cell's exact speed
 * (cell_speed) might exceed thr > 0, from its depth h, roughness n and an
 * upper bound s_hi of its largest slope (flat: the slope is exactly 0).
 * v = RN(RN(RN(1/n')*h^(2/3))*sqrt(s)), n' = max(n, 1e-12), is within
 * (1+2^-24)^5 of (1/n') h^(2/3) s^(1/2); comparing sixth powers,
 * h^4 s^3 < (thr*n')^6, avoids every root. With h >= 2^-20 and s_hi >= 2^-30
 * every partial product is a normal number, so each side is within 2^-20 of
 * its exact value and the margin 2^-12 covers both. The critical speed bounds
 * v too, so a cell whose vcr is below thr is skipped as well. A flat cell's speed is
 * the floor 1e-6, which never exceeds a thr >= 1e-6 (a thr found by the walk
 * is an exact speed, hence at least 1e-6). */
template <class R>
FCA_HDI int speed_may_exceed(R h, R n, R s_hi, int flat, R thr, R g)
{
    if (!(thr > (R)0)) {
        return 1;
    }
    if (flat) {
        return !(thr >= (R)1e-6);
    }
    /* the critical speed caps v: vcr = RN(sqrt(g*h)) <= thr when
     * fl(g*h)*(1+2^-20) <= fl(thr*thr) (each side within 2^-23) */
    if (g * h * ((R)1 + (R)0x1p-20) <= thr * thr && thr >= (R)1e-6 && thr <= (R)0x1p60) {
        return 0;
    }
    {
        const R nn = (n < (R)1e-12) ? (R)1e-12 : n;
        const R h2 = h * h, s2 = s_hi * s_hi;
        const R lhs = ((h2 * h2) * (s2 * s_hi)) * ((R)1 + (R)0x1p-12);
        const R tn = thr * nn, tn2 = tn * tn;
        const R rhs = (tn2 * tn2) * tn2;
        return !(h >= (R)0x1p-20 && s_hi >= (R)0x1p-30 && lhs >= (R)0x1p-100 && lhs < rhs);
    }
}

/* The same test in base-2 logarithms on the device (float): with lg the
 * hardware log2, whose absolute error over every positive normal float is at
 * most 2^-14 (checked on the GPU over all of them, tests/test_gpu_math.py),
 *   L   = (2/3) lg h + (1/2) lg s_hi - lg n'   bounds log2 of the Manning speed,
 *   Lcr = (1/2) lg h + (1/2) lg g              bounds log2 of the critical speed,
 * each within 3.2 * 2^-14 + 2^-15 of the exact value (the lg errors, the
 * roundings of the fmas at |L| < 512 and of RN(2/3)), and the computed speed
 * is within 2^-21 of the exact product in log2. So min(L, Lcr) < lthr =
 * lg(thr) - 2^-10 implies v < thr. Both logs need normal arguments; lthr is
 * -inf until the wavefront has a thr >= 1e-6 (the speed floor), so nothing
 * is skipped before then. lgg = (1/2) lg g. */
FCA_HDI float speed_log_thr(float thr)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* thr is the same on every lane: so is the result, which readfirstlane
     * lets the compiler keep in a scalar register */
    const float l = (thr >= 1e-6f && thr <= 0x1p100f) ? __builtin_amdgcn_logf(thr) - 0x1p-10f
                                                       : -INFINITY;
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(l)));
#else
    (void)thr;
    return -INFINITY;
#endif
}
FCA_HDI double speed_log_thr(double) { return -INFINITY; }

template <class R>
FCA_HDI int speed_may_exceed_fast(R h, R n, R s_hi, int flat, R thr, R lthr, R lgg, R g)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        if (flat) {
            return !(thr >= (R)1e-6);
        }
        {
            const float nn = (n < 1e-12f) ? 1e-12f : (float)n;
            const float lh = __builtin_amdgcn_logf((float)h);
            const float lcr = __builtin_fmaf(lh, 0.5f, (float)lgg);
            const float l = __builtin_fmaf(
                lh, 2.0f / 3.0f,
                __builtin_fmaf(__builtin_amdgcn_logf((float)s_hi), 0.5f, -__builtin_amdgcn_logf(nn)));
            const float lm = (l < lcr) ? l : lcr;
            return !(lm < (float)lthr && __builtin_fminf((float)h, (float)s_hi) >= 0x1p-126f);
        }
    }
#endif
    (void)lthr, (void)lgg;
    return speed_may_exceed<R>(h, n, s_hi, flat, thr, g);
}

/* speed_may_exceed_fast in two parts: the logarithms that depend on the
 * cell's depth and roughness alone (speed_logs), which a walk may form
 * before the cell's slope is known, and the test given its slope bound
 * (speed_may_exceed_logs); the same operations, so the same result. */
template <class R>
FCA_HDI SpeedLogs<R> speed_logs(R h, R n, R lgg)
{
    SpeedLogs<R> o;
    o.lh = o.lcr = o.lnn = (R)0;
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        const float nn = (n < 1e-12f) ? 1e-12f : (float)n;
        o.lh = (R)__builtin_amdgcn_logf((float)h);
        o.lcr = (R)__builtin_fmaf((float)o.lh, 0.5f, (float)lgg);
        o.lnn = (R)__builtin_amdgcn_logf(nn);
    }
#endif
    (void)h, (void)n, (void)lgg;
    return o;
}
template <class R>
FCA_HDI int speed_may_exceed_logs(R h, R n, R s_hi, int flat, R thr, R lthr,
                                  const SpeedLogs<R> &lg, R g)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        if (flat) {
            return !(thr >= (R)1e-6);
        }
        {
            const float l = __builtin_fmaf(
                (float)lg.lh, 2.0f / 3.0f,
                __builtin_fmaf(__builtin_amdgcn_logf((float)s_hi), 0.5f, -(float)lg.lnn));
            const float lm = (l < (float)lg.lcr) ? l : (float)lg.lcr;
            return !(lm < (float)lthr && __builtin_fminf((float)h, (float)s_hi) >= 0x1p-126f);
        }
    }
#endif
    (void)lthr, (void)lg;
    return speed_may_exceed<R>(h, n, s_hi, flat, thr, g);
}

/* (1/2) lg g for speed_may_exceed_fast (device, float). */
template <class R>
FCA_HDI R speed_log_half_g(R g)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        /* +inf (no skip by the critical speed) unless g is a normal float */
        const float l = (g >= (R)0x1p-126 && g <= (R)0x1p127)
                            ? 0.5f * __builtin_amdgcn_logf((float)g) : INFINITY;
        return (R)__int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(l)));
    }
#endif
    (void)g;
    return (R)0;
}

/* Open-boundary faces of a LAND cell of row q (model.c:166-177). */
template <class R>
struct OpenFaces {
    R w, e, n, s;
    int cap;
};
template <class R>
FCA_HDI OpenFaces<R> open_faces(R z, R h, R n, int m, int c, int q, int nr, int nc, R dt, R den,
                               R rden, const PhysK<R> &p)
{
    OpenFaces<R> o;
    int cw = 0, ce = 0, cn = 0, cs = 0;
    o.w = o.e = o.n = o.s = (R)0;
    if (m == FCA_LAND) {
        if (c == 0) {
            o.w = open_flux<R>(z, h, n, dt, den, rden, p, cw);
        }
        if (c == nc - 1) {
            o.e = open_flux<R>(z, h, n, dt, den, rden, p, ce);
        }
        if (q == 0) {
            o.n = open_flux<R>(z, h, n, dt, den, rden, p, cn);
        }
        if (q == nr - 1) {
            o.s = open_flux<R>(z, h, n, dt, den, rden, p, cs);
        }
    }
    o.cap = cw + ce + cn + cs;
    return o;
}

/* One row of the walk after row r's inputs were taken into the ring.
 * LANDF: rows r-4..r are LAND on every lane, 1 <= r-4, r <= nr-2 and no lane
 * holds column 0 or nc-1. */
template <class R, int K8, int LANDF, int QT, class LP>
FCA_HDI void walk_row(LP &lp, const WalkArgs<R> &a, const WalkRow &w)
{
    typedef WalkLane<R> L;
    const int nr = a.nr, nc = a.nc, r = w.r;
    const PhysK<R> &p = a.p;
    const StepVals<R> &sv = a.sv;
    const R dt = sv.dt, den = sv.den, rden = sv.rden, A = p.A;
    const int dfast = sv.den_fast;
    const int rowA = r < nr;                            /* row r exists */
    const int outA = r >= w.r_out0 && r < w.r_out1;    /* row r is written here */
    const int qB = r - 1, qC = r - 2, qD = r - 3;
    const int outB = qB >= w.r_out0 && qB < w.r_out1;
    const int dD = sv.need_cand && qD >= w.r_out0 && qD < w.r_out1 && qD < nr;
    /* the LAND form for K = 4 in float forms a row's independent work first
     * (below); in double the longer live ranges cost more than they save */
    constexpr bool HOIST = LANDF && !K8 && sizeof(R) == 4;

    /* ---- D, first part (LAND form, K = 4): the level differences of row
     * r-3 that do not wait for this row's update -- east, west and north,
     * from depths of earlier rows -- are formed before the faces, so their
     * work can share the faces' basic blocks; the south one follows stage C.
     * (A maximum of magnitudes does not depend on its order.) */
    if (HOIST && dD) {
        lp.each([&](L &s, int) { s.dl3 = s.z3 + s.hn3, s.dl4 = s.z4 + s.hn4; });
        lp.pull_right(&L::dle, &L::dl3);
        lp.pull_left(&L::dlw, &L::dl3);
        lp.each([&](L &s, int) {
            s.dax = fmax_r(fmax_r(fmax_r((R)0, fabs_r(s.dl3 - s.dle)), fabs_r(s.dl3 - s.dlw)),
                           fabs_r(s.dl3 - s.dl4));
            s.dlg = speed_logs<R>(s.hn3, s.n3, a.lgg);
        });
    }
    /* ---- C, first part (LAND form, K = 4): the east and west terms of row
     * r-2's net inflow, whose limiters are known from the previous row, in
     * the reference's order (net = -E + W, then -S + N after stage B) */
    if (HOIST) {
        lp.pull_right(&L::xw, &L::s2);
        lp.pull_left(&L::xy, &L::s2);
        lp.pull_left(&L::xa, &L::qx2);
        lp.each([&](L &s, int) {
            R net = (R)0;
            net -= s.qx2 * ((s.qx2 >= (R)0) ? s.s2 : s.xw);
            net += s.xa * ((s.xa >= (R)0) ? s.xy : s.s2);
            s.cnet = net;
        });
    }

    /* ---- A: the faces that touch row r ---------------------------------- */
    if (LANDF) {
        lp.pull_right(&L::xz, &L::z0);
        lp.pull_right(&L::xl, &L::lv0);
        lp.pull_right(&L::xn, &L::n0);
        if (K8) {
            lp.pull_right(&L::xq, &L::z1);
            lp.pull_right(&L::xl1, &L::lv1);
            lp.pull_right(&L::xs, &L::n1);
        }
        lp.each([&](L &s, int) {
            /* the east face of row r and the face from row r-1 to row r
             * together (and the two diagonals), so their work interleaves */
            R q[2];
            bool cb[2];
            {
                const R za[2] = {s.z0, s.z1}, la[2] = {s.lv0, s.lv1}, na[2] = {s.n0, s.n1};
                const R zb[2] = {s.xz, s.z0}, lb[2] = {s.xl, s.lv0}, nb[2] = {s.xn, s.n0};
                face_land_pair<R>(za, la, na, zb, lb, nb, p.dx, p.rdx, dt, den, rden, dfast, p, q,
                                  cb);
            }
            s.qx0 = q[0], s.qy1 = q[1];
            lp.count_caps(s, outA && cb[0]);
            lp.count_caps(s, outB && cb[1]);
            if (K8) {
                const R za[2] = {s.z1, s.xq}, la[2] = {s.lv1, s.xl1}, na[2] = {s.n1, s.xs};
                const R zb[2] = {s.xz, s.z0}, lb[2] = {s.xl, s.lv0}, nb[2] = {s.xn, s.n0};
                face_land_pair<R>(za, la, na, zb, lb, nb, p.ddiag, p.rddiag, dt, den, rden, dfast,
                                  p, q, cb);
                s.qd1_1 = q[0], s.qd2_1 = q[1];
                lp.count_caps(s, outB && cb[0]);
                lp.count_caps(s, outB && cb[1]);
            }
        });
    } else {
        lp.pull_right(&L::xz, &L::z0);
        lp.pull_right(&L::xh, &L::h0);
        lp.pull_right(&L::xn, &L::n0);
        lp.pull_right(&L::xm, &L::m0);
        if (K8) {
            lp.pull_right(&L::xq, &L::z1);
            lp.pull_right(&L::xr, &L::h1);
            lp.pull_right(&L::xs, &L::n1);
            lp.pull_right(&L::xm2, &L::m1);
        }
        lp.each([&](L &s, int) {
            int cap;
            const int east = s.colok && s.c + 1 < nc;
            s.qx0 = (R)0;
            s.qy1 = (R)0;
            s.qd1_1 = s.qd2_1 = (R)0;
            if (!rowA) {
                return;
            }
            if (east) {
                s.qx0 = face_flux<R>(s.z0, s.h0, s.n0, s.m0, s.xz, s.xh, s.xn, s.xm, p.dx, p.rdx,
                                     p.fdx, dt, den, rden, p, cap);
                if (outA && s.outcol) {
                    s.cap += cap;
                }
            }
            if (r >= 1 && s.colok) {
                s.qy1 = face_flux<R>(s.z1, s.h1, s.n1, s.m1, s.z0, s.h0, s.n0, s.m0, p.dx, p.rdx,
                                     p.fdx, dt, den, rden, p, cap);
                if (outB && s.outcol) {
                    s.cap += cap;
                }
                if (K8 && east) {
                    s.qd1_1 = face_flux<R>(s.z1, s.h1, s.n1, s.m1, s.xz, s.xh, s.xn, s.xm,
                                           p.ddiag, p.rddiag, p.fddiag, dt, den, rden, p, cap);
                    if (outB && s.outcol) {
                        s.cap += cap;
                    }
                    s.qd2_1 = face_flux<R>(s.xq, s.xr, s.xs, s.xm2, s.z0, s.h0, s.n0, s.m0,
                                           p.ddiag, p.rddiag, p.fddiag, dt, den, rden, p, cap);
                    if (outB && s.outcol) {
                        s.cap += cap;
                    }
                }
            }
        });
    }

    /* ---- B: outflow volume and limiter of row r-1 ---------------------- */
    lp.pull_left(&L::xt, &L::qx1);
    if (K8) {
        lp.pull_left(&L::xu, &L::qd1_2);
        lp.pull_left(&L::xv, &L::qd2_1);
    }
    if (LANDF) {
        /* out += q for q > 0, in the order E, W, S, N[, SE, NW, SW, NE]; adding
         * a zero instead of skipping leaves the (non-negative) sum unchanged */
        lp.each([&](L &s, int) {
            R out = pos_part(s.qx1) + pos_part(-s.xt);
            out = out + pos_part(s.qy1);
            out = out + pos_part(-s.qy2);
            if (K8) {
                out = out + pos_part(s.qd1_1);
                out = out + pos_part(-s.xu);
                out = out + pos_part(s.xv);
                out = out + pos_part(-s.qd2_2);
            }
            out = out * dt;
            s.s1 = limiter_s_wave<R>(out, s.h1, A);
        });
    } else if (qB >= 0 && qB < nr) {
        const int open_row = p.open && (w.edge_cols || qB == 0 || qB == nr - 1);
        lp.each([&](L &s, int) {
            R out = (R)0;
            if (!s.colok) {
                s.s1 = (R)1;
                return;
            }
            if (s.c + 1 < nc && s.qx1 > (R)0) {
                out += s.qx1;
            }
            if (s.c > 0 && s.xt < (R)0) {
                out += -s.xt;
            }
            if (qB + 1 < nr && s.qy1 > (R)0) {
                out += s.qy1;
            }
            if (qB > 0 && s.qy2 < (R)0) {
                out += -s.qy2;
            }
            if (K8) {
                if (qB + 1 < nr && s.c + 1 < nc && s.qd1_1 > (R)0) {
                    out += s.qd1_1;
                }
                if (qB > 0 && s.c > 0 && s.xu < (R)0) {
                    out += -s.xu;
                }
                if (qB + 1 < nr && s.c > 0 && s.xv > (R)0) {
                    out += s.xv;
                }
                if (qB > 0 && s.c + 1 < nc && s.qd2_2 < (R)0) {
                    out += -s.qd2_2;
                }
            }
            if (open_row) {
                const OpenFaces<R> o =
                    open_faces<R>(s.z1, s.h1, s.n1, s.m1, s.c, qB, nr, nc, dt, den, rden, p);
                out += o.w + o.e + o.n + o.s;
                if (outB && s.outcol) {
                    s.cap += o.cap;
                }
            }
            out = out * dt;
            s.s1 = (s.m1 == FCA_LAND) ? limiter_s<R>(out, s.h1, A) : (R)1;
        });
    }

    /* ---- C: update row r-2 --------------------------------------------- */
    if (!HOIST) {
        lp.pull_right(&L::xw, &L::s2);
        lp.pull_left(&L::xy, &L::s2);
        lp.pull_left(&L::xa, &L::qx2);
    }
    if (K8) {
        lp.pull_right(&L::dz1, &L::s1); /* s of (q+1, c+1) */
        lp.pull_left(&L::dz2, &L::s1);  /* s of (q+1, c-1) */
        lp.pull_right(&L::dz3, &L::s3); /* s of (q-1, c+1) */
        lp.pull_left(&L::dz4, &L::s3);  /* s of (q-1, c-1) */
        lp.pull_left(&L::dh1, &L::qd1_3);
        lp.pull_left(&L::dh2, &L::qd2_2);
    }
    if (!LANDF) {
        lp.pull_right(&L::xm3, &L::m2);
        lp.pull_left(&L::xm4, &L::m2);
        if (K8) {
            lp.pull_right(&L::dm1, &L::m1);
            lp.pull_left(&L::dm2, &L::m3);
            lp.pull_left(&L::dm3, &L::m1);
            lp.pull_right(&L::dm4, &L::m3);
        }
    }
    if (qC >= 0 && qC < nr) {
        const int outC = qC >= w.r_out0 && qC < w.r_out1;
        const int open_row = !LANDF && p.open && (w.edge_cols || qC == 0 || qC == nr - 1);
        lp.each([&](L &s, int) {
            R net = (R)0, q, rain, hnew;
            double sea = 0.0;
            const int land = LANDF || s.m2 == FCA_LAND;
            if (!LANDF && !s.colok) {
                return;
            }
            /* E, W, S, N[, SE, NW, SW, NE] in the reference's order */
            if (HOIST) {
                net = s.cnet;
            } else if (LANDF || s.c + 1 < nc) {
                q = s.qx2 * ((s.qx2 >= (R)0) ? s.s2 : s.xw);
                net -= q;
                if (!LANDF && land && s.xm3 == FCA_SEA && q > (R)0) {
                    sea += (double)dt * (double)q;
                }
            }
            if (!HOIST && (LANDF || s.c > 0)) {
                q = s.xa * ((s.xa >= (R)0) ? s.xy : s.s2);
                net += q;
                if (!LANDF && land && s.xm4 == FCA_SEA && q < (R)0) {
                    sea += (double)dt * (double)(-q);
                }
            }
            if (LANDF || qC + 1 < nr) {
                q = s.qy2 * ((s.qy2 >= (R)0) ? s.s2 : s.s1);
                net -= q;
                if (!LANDF && land && s.m1 == FCA_SEA && q > (R)0) {
                    sea += (double)dt * (double)q;
                }
            }
            if (LANDF || qC > 0) {
                q = s.qy3 * ((s.qy3 >= (R)0) ? s.s3 : s.s2);
                net += q;
                if (!LANDF && land && s.m3 == FCA_SEA && q < (R)0) {
                    sea += (double)dt * (double)(-q);
                }
            }
            if (K8) {
                if (LANDF || (qC + 1 < nr && s.c + 1 < nc)) {
                    q = s.qd1_2 * ((s.qd1_2 >= (R)0) ? s.s2 : s.dz1);
                    net -= q;
                    if (!LANDF && land && s.dm1 == FCA_SEA && q > (R)0) {
                        sea += (double)dt * (double)q;
                    }
                }
                if (LANDF || (qC > 0 && s.c > 0)) {
                    q = s.dh1 * ((s.dh1 >= (R)0) ? s.dz4 : s.s2);
                    net += q;
                    if (!LANDF && land && s.dm2 == FCA_SEA && q < (R)0) {
                        sea += (double)dt * (double)(-q);
                    }
                }
                if (LANDF || (qC + 1 < nr && s.c > 0)) {
                    q = s.dh2 * ((s.dh2 >= (R)0) ? s.s2 : s.dz2);
                    net -= q;
                    if (!LANDF && land && s.dm3 == FCA_SEA && q > (R)0) {
                        sea += (double)dt * (double)q;
                    }
                }
                if (LANDF || (qC > 0 && s.c + 1 < nc)) {
                    q = s.qd2_3 * ((s.qd2_3 >= (R)0) ? s.dz3 : s.s2);
                    net += q;
                    if (!LANDF && land && s.dm4 == FCA_SEA && q < (R)0) {
                        sea += (double)dt * (double)(-q);
                    }
                }
            }
            if (open_row) {
                /* the open faces of row r-2 again (counted in B one row ago) */
                const OpenFaces<R> o =
                    open_faces<R>(s.z2, s.h2, s.n2, s.m2, s.c, qC, nr, nc, dt, den, rden, p);
                const R qos = o.w * s.s2 + o.e * s.s2 + o.n * s.s2 + o.s * s.s2;
                net -= qos;
                if (outC && s.outcol) {
                    s.v_bnd += (double)dt * (double)qos;
                }
            }
            if (!LANDF && outC && s.outcol) {
                s.v_sea += sea;
            }
            s.hst = s.h2; /* the depth a stop-rising store would record */
            if (!land) {
                s.hn2 = s.h2;
                return;
            }
            {
                rain = s.pcr * sv.rain_f;
                const R hraw =
                    s.h2 + (dt * (LANDF ? div_c_wave(net, A, p.rA) : div_c(net, A, p.rA, p.fA)) +
                            rain);
                const bool neg = hraw < (R)0; /* clamped to +0 */
                hnew = neg ? (R)0 : hraw;
                s.hn2 = hnew;
                if (outC && s.outcol) {
                    /* rose in the previous step (h_old > stored hmax) and
                     * not in this one: store hmax = h_old and its tmax */
                    const int stop = sv.armed && s.h2 > s.phm && !(hnew > s.h2);
                    s.v_rain += (double)rain;
                    /* the clamped depth is added only where some lane of
                     * the wavefront clamps (adding +0 elsewhere would leave
                     * the non-negative sum as it is): a branch, not a select */
                    if (wave_any(hraw < (R)0)) {
                        FCA_COLD_PATH();
                        if (neg) {
                            s.clamp += (double)(-hraw);
                        }
                    }
                    s.stq = stop ? 3 : 1;
                    if (QT) {
                        /* changed: h, a store, or the true hmax (h_new >
                         * stored hmax also where h does not move, as on the
                         * first step from a caller's h above its hmax) */
                        s.qwet |= !(hnew < p.tau_h);
                        s.qchg |= !same_bits(hnew, s.h2) || stop || hnew > s.phm;
                    }
                }
            }
        });
    }

    /* ---- D: next step's stable time step on row r-3 -------------------- */
    if (dD) {
      if (HOIST) {
        lp.each([&](L &s, int) { s.dl2 = s.z2 + s.hn2; });
      } else if (LANDF) {
        /* LAND form: each lane forms its own levels once and the neighbours'
         * are pulled as levels (the same sums, formed by their own lanes) */
        lp.each([&](L &s, int) {
            s.dl2 = s.z2 + s.hn2, s.dl3 = s.z3 + s.hn3, s.dl4 = s.z4 + s.hn4;
        });
        lp.pull_right(&L::dle, &L::dl3);
        lp.pull_left(&L::dlw, &L::dl3);
        if (K8) {
            lp.pull_right(&L::dz1, &L::dl2);
            lp.pull_left(&L::dz2, &L::dl2);
            lp.pull_right(&L::dz3, &L::dl4);
            lp.pull_left(&L::dz4, &L::dl4);
        }
      } else {
        lp.pull_right(&L::xz, &L::z3);
        lp.pull_right(&L::xh, &L::hn3);
        lp.pull_left(&L::xq, &L::z3);
        lp.pull_left(&L::xr, &L::hn3);
        if (!LANDF) {
            lp.pull_right(&L::xm, &L::m3);
            lp.pull_left(&L::xm2, &L::m3);
        }
        if (K8) {
            lp.pull_right(&L::dz1, &L::z2);
            lp.pull_right(&L::dh1, &L::hn2);
            lp.pull_left(&L::dz2, &L::z2);
            lp.pull_left(&L::dh2, &L::hn2);
            lp.pull_right(&L::dz3, &L::z4);
            lp.pull_right(&L::dh3, &L::hn4);
            lp.pull_left(&L::dz4, &L::z4);
            lp.pull_left(&L::dh4, &L::hn4);
            if (!LANDF) {
                lp.pull_right(&L::dm1, &L::m2);
                lp.pull_left(&L::dm2, &L::m2);
                lp.pull_right(&L::dm3, &L::m4);
                lp.pull_left(&L::dm4, &L::m4);
            }
        }
      }
        lp.each([&](L &s, int) {
            R la, ax = (R)0, ad = (R)0, s_hi;
            s.dneed = 0;
            if (!s.outcol || (!LANDF && s.m3 != FCA_LAND) || s.hn3 < p.tau_h) {
                return;
            }
            la = LANDF ? s.dl3 : s.z3 + s.hn3;
            /* largest |level difference| per face length: max_k RN(a_k/d)
             * = RN(max_k a_k / d) as rounding is monotonic, so one
             * division per length replaces one per neighbour. The LAND form
             * takes fmax of magnitudes: it differs from the general form
             * only in the sign of a zero maximum, and a zero slope gives the
             * floor speed either way. */
#define FCA_LEVEL_DIFF(cond, m_, z_, h_, acc_)                                \
    if (!LANDF && (cond) && (m_) != FCA_WALL) {                                 \
        R dl_ = la - level_of<R>((m_), (z_), (h_), p);                          \
        if (dl_ < (R)0) {                                                       \
            dl_ = -dl_;                                                         \
        }                                                                       \
        if (dl_ > acc_) {                                                       \
            acc_ = dl_;                                                         \
        }                                                                       \
    }
            if (HOIST) {
                ax = fmax_r(s.dax, fabs_r(la - s.dl2));
            } else if (LANDF) {
                ax = fmax_r(ax, fabs_r(la - s.dle));
                ax = fmax_r(ax, fabs_r(la - s.dlw));
                ax = fmax_r(ax, fabs_r(la - s.dl2));
                ax = fmax_r(ax, fabs_r(la - s.dl4));
            }
            FCA_LEVEL_DIFF(s.c + 1 < nc, s.xm, s.xz, s.xh, ax)
            FCA_LEVEL_DIFF(s.c > 0, s.xm2, s.xq, s.xr, ax)
            FCA_LEVEL_DIFF(qD + 1 < nr, s.m2, s.z2, s.hn2, ax)
            FCA_LEVEL_DIFF(qD > 0, s.m4, s.z4, s.hn4, ax)
            if (!LANDF && p.open && (s.c == 0 || s.c == nc - 1 || qD == 0 || qD == nr - 1)) {
                R dl_ = la - (s.z3 - p.S_bnd * p.dx);
                if (dl_ < (R)0) {
                    dl_ = -dl_;
                }
                if (dl_ > ax) {
                    ax = dl_;
                }
            }
            if (K8 && LANDF) {
                ad = fmax_r(ad, fabs_r(la - s.dz1));
                ad = fmax_r(ad, fabs_r(la - s.dz2));
                ad = fmax_r(ad, fabs_r(la - s.dz3));
                ad = fmax_r(ad, fabs_r(la - s.dz4));
            }
            if (K8) {
                FCA_LEVEL_DIFF(qD + 1 < nr && s.c + 1 < nc, s.dm1, s.dz1, s.dh1, ad)
                FCA_LEVEL_DIFF(qD + 1 < nr && s.c > 0, s.dm2, s.dz2, s.dh2, ad)
                FCA_LEVEL_DIFF(qD > 0 && s.c + 1 < nc, s.dm3, s.dz3, s.dh3, ad)
                FCA_LEVEL_DIFF(qD > 0 && s.c > 0, s.dm4, s.dz4, s.dh4, ad)
            }
#undef FCA_LEVEL_DIFF
            s.dax = ax, s.dad = ad;
            /* RN(ax*RN(1/dx)) is within 2^-23 of ax/dx, so after the factor
             * (1+2^-20) it bounds RN(ax/dx) from above */
            s_hi = ax * p.rdx;
            if (K8) {
                const R sd = ad * p.rddiag;
                s_hi = (sd > s_hi) ? sd : s_hi;
            }
            s_hi = s_hi * ((R)1 + (R)0x1p-20);
            s.dneed = HOIST
                          ? speed_may_exceed_logs<R>(s.hn3, s.n3, s_hi, ax == (R)0 && ad == (R)0,
                                                     s.vthr, s.lthr, s.dlg, p.g)
                          : speed_may_exceed_fast<R>(s.hn3, s.n3, s_hi,
                                                     ax == (R)0 && ad == (R)0, s.vthr, s.lthr,
                                                     a.lgg, p.g);
        });
        if (lp.any(&L::dneed)) {
            lp.each([&](L &s, int) {
                if (s.dneed) {
                    R smax = div_c(s.dax, p.dx, p.rdx, p.fdx), v;
                    if (K8) {
                        const R sd = div_c(s.dad, p.ddiag, p.rddiag, p.fddiag);
                        if (sd > smax) {
                            smax = sd;
                        }
                    }
                    v = cell_speed<R>(s.hn3, s.n3, smax, p);
                    if (v > s.vmax) {
                        s.vmax = v;
                    }
                }
            });
            lp.wave_max_to(&L::vthr, &L::vmax);
            lp.each([&](L &s, int) { s.lthr = speed_log_thr(s.vthr); });
        }
    }
}

/* Walks one strip for one wavefront. strip covers output rows
 * [r_out0, r_out1) and columns [c_first + FCA_HALO, c_first + LP::WIDTH - FCA_HALO).
 * QT: also record whether the strip's cells end wet or changed (qwet, qchg)
 * for quiet-tile skipping. */
template <class R, int K8, int QT, class LP>
FCA_HD void walk_strip(LP &lp, const WalkArgs<R> &a, int r_out0, int r_out1, int c_first)
{
    typedef WalkLane<R> L;
    /* row accesses with a scalar base for the K = 8 walk of one column per
     * work-item without the quiet-tile bookkeeping: the vector registers the
     * plain form spends on addresses are what keeps it from one more
     * wavefront per SIMD (four in float, three in double). With the
     * bookkeeping the extra wavefront loses on wet grids, as for K = 4, so
     * that walk keeps the plain form. */
    constexpr bool SB = K8 && !QT && LP::WIDTH == 64;
    const int nr = a.nr, nc = a.nc;
    const int rb = (r_out0 - 3 > 0) ? r_out0 - 3 : 0;
    const int re = (r_out1 + 2 < nr + 2) ? r_out1 + 2 : nr + 2; /* inclusive */
    /* every lane inside the grid and none on its first or last column */
    const int cols_inner = c_first >= 1 && c_first + LP::WIDTH - 1 <= nc - 2;
    WalkRow w;
    unsigned land_rows; /* bit k: row r-k is LAND on every lane */
    unsigned void_rows; /* bit k: row r-k has no LAND lane */
    int r;

    w.r_out0 = r_out0, w.r_out1 = r_out1, w.c_first = c_first;
    w.edge_cols = c_first <= 0 || c_first + LP::WIDTH - 1 >= nc - 1;
    lp.each([&](L &s, int lane) {
        s.c = c_first + lane;
        s.colok = s.c >= 0 && s.c < nc;
        s.outcol = s.colok && lane >= FCA_HALO && lane < LP::WIDTH - FCA_HALO;
        s.boff = (uint32_t)s.c * (uint32_t)sizeof(R);
        s.z0 = s.z1 = s.z2 = s.z3 = s.z4 = (R)0;
        s.m0 = s.m1 = s.m2 = s.m3 = s.m4 = FCA_NONE;
        s.h0 = s.h1 = s.h2 = (R)0;
        s.n0 = s.n1 = s.n2 = s.n3 = (R)0;
        s.lv0 = s.lv1 = (R)0;
        s.hn2 = s.hn3 = s.hn4 = (R)0;
        s.qx0 = s.qx1 = s.qx2 = s.qy1 = s.qy2 = s.qy3 = (R)0;
        s.qd1_1 = s.qd1_2 = s.qd1_3 = s.qd2_1 = s.qd2_2 = s.qd2_3 = (R)0;
        s.s1 = s.s2 = s.s3 = (R)1;
        s.pcr = s.phm = s.pcr1 = s.phm1 = (R)0;
        s.stq = 0;
        s.hst = (R)0;
        s.qwet = s.qchg = 0;
        s.pz = s.ph = s.pn = (R)0;
        s.pm1 = s.pm2 = FCA_NONE;
        if (s.colok && rb < nr) {
            const size_t ro = (size_t)rb * (size_t)nc;
            s.pm1 = a.mask[ro + (uint32_t)s.c];
            if (s.pm1 == FCA_LAND) {
                s.pz = ld_row<SB>(a.z + ro, s.boff);
                s.ph = ld_row<SB>(a.h_in + ro, s.boff);
                s.pn = ld_row<SB>(a.n + ro, s.boff);
            }
        }
        if (s.colok && rb + 1 < nr) {
            s.pm2 = a.mask[(size_t)(rb + 1) * (size_t)nc + (uint32_t)s.c];
        }
    });
    lp.strip_begun();
    /* rows before rb are not loaded; the ring holds them as outside the grid */
    land_rows = 0u;
    void_rows = ~0u;

    /* take row r (prefetched), prefetch z/h/n of row r+1 and the mask of row
     * r+2, and form row r's form bits */
    auto take_row = [&](int r) {
        /* Every memory access of a row is issued here, at the top: the
         * loads are consumed one row later, so each has a whole row of work
         * to arrive in, and the stores of row r-3 (computed one row ago)
         * follow the wait for the previous row's loads, so no wait ever
         * covers a store that was just issued. */
        {
            const int ok1 = r + 1 < nr, ok2 = r + 2 < nr;
            const int q = r - 1, okq = q >= 0 && q < nr;
            const int outq = q >= r_out0 && q < r_out1;
            const size_t ro1 = (size_t)(r + 1) * (size_t)nc, ro2 = (size_t)(r + 2) * (size_t)nc;
            const size_t roq = (size_t)(okq ? q : 0) * (size_t)nc;
            const R *zr = a.z + ro1, *hr = a.h_in + ro1, *nrw = a.n + ro1;
            const R *crr = a.cr + roq, *hmr = a.hmax + roq;
            const uint8_t *mr = a.mask + ro2;
            const int qs = r - 3, oks = qs >= 0 && qs < nr;
            const size_t ros = (size_t)(oks ? qs : 0) * (size_t)nc;
            R *hos = a.h_out + ros, *hms = a.hmax + ros, *tms = a.tmax + ros;
            const R tval = a.sv.tmax_prev;
            lp.each([&](L &s, int) {
                s.m0 = s.pm1;
                if (s.m0 == FCA_LAND) {
                    s.z0 = s.pz, s.h0 = s.ph, s.n0 = s.pn;
                }
                s.lv0 = s.z0 + s.h0;
            });
            /* the wait for the previous row's loads goes here, before this
             * row's stores are issued */
            lp.loads_settled();
            lp.each([&](L &s, int) {
                if (s.stq) {
                    st_row<SB>(hos, s.boff, s.hn3);
                    if (s.stq & 2) {
                        st_row<SB>(hms, s.boff, s.hst);
                        st_row<SB>(tms, s.boff, tval);
                    }
                    s.stq = 0;
                }
                s.pm1 = s.pm2;
                if (s.colok && ok1 && s.pm1 == FCA_LAND) {
                    s.pz = ld_row<SB>(zr, s.boff), s.ph = ld_row<SB>(hr, s.boff);
                    s.pn = ld_row<SB>(nrw, s.boff);
                }
                s.pm2 = (s.colok && ok2) ? (int)ld_row<SB>(mr, (uint32_t)s.c) : FCA_NONE;
                /* cr and hmax of row r-1 (m1), updated one row from now;
                 * hmax is read before that update writes it */
                if (s.colok && okq && s.m1 == FCA_LAND) {
                    s.pcr1 = ld_row<SB>(crr, s.boff);
                    if (s.outcol && outq) {
                        s.phm1 = ld_row<SB>(hmr, s.boff);
                    }
                }
            });
        }
        /* the row forms from row r's mask, fetched two rows ago, so that no
         * load of this row is waited for here */
        {
            /* a lane whose level or depth leaves the LAND form's range
             * (|z + h| <= 2^20, |h| <= 2^14) counts as not LAND here */
            const uint64_t nonland = lp.non_land_lanes(&L::m0) | lp.out_of_range_lanes();
            const uint64_t land = lp.land_lanes(&L::m0);
            land_rows = (land_rows << 1) | (nonland == 0 ? 1u : 0u);
            void_rows = (void_rows << 1) | (land == 0 ? 1u : 0u);
        }
        w.r = r;
    };
    /* row r takes the LAND form: rows r-4..r LAND on every lane, away from
     * the grid's edges */
    auto land_form = [&](int r) {
        return a.p.fast && cols_inner && (land_rows & 0x1Fu) == 0x1Fu && r - 4 >= 1 &&
               r <= nr - 2;
    };
    auto rotate = [&]() {
        lp.each([&](L &s, int) {
            s.z4 = s.z3, s.z3 = s.z2, s.z2 = s.z1, s.z1 = s.z0;
            s.m4 = s.m3, s.m3 = s.m2, s.m2 = s.m1, s.m1 = s.m0;
            s.h2 = s.h1, s.h1 = s.h0;
            s.n3 = s.n2, s.n2 = s.n1, s.n1 = s.n0;
            s.lv1 = s.lv0;
            s.hn4 = s.hn3, s.hn3 = s.hn2;
            s.qx2 = s.qx1, s.qx1 = s.qx0;
            s.qy3 = s.qy2, s.qy2 = s.qy1;
            s.qd1_3 = s.qd1_2, s.qd1_2 = s.qd1_1;
            s.qd2_3 = s.qd2_2, s.qd2_2 = s.qd2_1;
            s.s3 = s.s2, s.s2 = s.s1;
            s.pcr = s.pcr1, s.phm = s.phm1;
        });
    };
    /* one LAND row of a run, then take the next row; false when the run
     * ends (the walk's end, or the next row is not LAND form) */
    auto land_step = [&](int &r) {
        lp.count_form(1);
        walk_row<R, K8, 1, QT>(lp, a, w);
        rotate();
        if (++r > re) {
            return false;
        }
        take_row(r);
        return land_form(r);
    };

    r = rb;
    if (r <= re) {
        take_row(r);
    }
    while (r <= re) {
        /* row r is taken */
        if (land_form(r)) {
            /* A run of LAND rows walks two rows per trip: the ring's shifts
             * are then renames within the trip, and only the values a trip
             * leaves in the ring cross the loop's back edge, rather than a
             * move per ring slot and row. */
            while (land_step(r) && land_step(r)) {
            }
            if (r > re) {
                break;
            }
        }
        if ((void_rows & 0xFu) == 0xFu) {
            /* rows r-3..r hold no LAND cell: every face touching rows r-1
             * and r is dead, the limiter of row r-1 is 1, row r-2 keeps its
             * depth and row r-3 has no time step (the general form's values
             * for non-LAND cells) */
            lp.count_form(2);
            lp.each([&](L &s, int) {
                s.qx0 = s.qy1 = s.qd1_1 = s.qd2_1 = (R)0;
                s.s1 = (R)1;
                s.hn2 = s.h2;
            });
        } else {
            lp.count_form(0);
            walk_row<R, K8, 0, QT>(lp, a, w);
        }
        rotate();
        if (++r <= re) {
            take_row(r);
        }
    }
    /* the last row's update is never an output row (re = r_out1 + 2), so no
     * deferred store is left; flush one all the same if the bounds change */
    if (lp.any(&L::stq)) {
        const size_t ros = (size_t)(re - 2) * (size_t)nc;
        R *hos = a.h_out + ros, *hms = a.hmax + ros, *tms = a.tmax + ros;
        lp.each([&](L &s, int) {
            if (s.stq) {
                st_row<SB>(hos, s.boff, s.hn3);
                if (s.stq & 2) {
                    st_row<SB>(hms, s.boff, s.hst);
                    st_row<SB>(tms, s.boff, a.sv.tmax_prev);
                }
                s.stq = 0;
            }
        });
    }
}

/* Zeroes a lane's accumulators; a wavefront may walk several strips. */
template <class R, class LP>
FCA_HD void walk_acc_init(LP &lp)
{
    lp.each([&](WalkLane<R> &s, int) {
        s.v_rain = s.clamp = s.v_sea = s.v_bnd = 0.0;
        s.cap = s.capw = 0;
        s.dtc = (R)INFINITY;
        s.vmax = (R)0;
        s.vthr = (R)0;
        s.lthr = (R)-INFINITY;
    });
}

/* The lane's time-step candidate from its largest speed: dt_a = alpha*dx/v is
 * non-increasing in v (alpha*dx > 0 is checked by the host). */
template <class R, class LP>
FCA_HD void walk_acc_finish(LP &lp, const PhysK<R> &p)
{
    lp.each([&](WalkLane<R> &s, int) {
        s.dtc = (s.vmax > (R)0) ? (p.alpha * p.dx) / s.vmax : (R)INFINITY;
        s.v_rain *= (double)p.A;
        s.clamp *= (double)p.A;
    });
}

#endif
