/*
 * fca_walk.h -- the fused WCA step as a column walk.
 *
 * One 64-lane wavefront owns 64 adjacent columns and walks down a strip of
 * rows. Lane l holds column c = c_first + l; the three lanes at each side are
 * halo, so a wavefront writes 58 columns. At row r the walk
 *   A. reads row r (z, h, n, mask) and forms the faces that touch it: the
 *      east face of row r, the face between rows r-1 and r, the two K=8
 *      diagonals between them, and the open-boundary faces of row r;
 *   B. sums the outflow of row r-1 and its limiter s (model.c:180-238,488-502);
 *   C. updates row r-2: upwind-scaled net inflow, rain, clamp, hmax/tmax and
 *      the FP64 balance terms (model.c:241-319, 511-597);
 *   D. evaluates the next step's stable time step on row r-3 from the new
 *      depths (model.c:111-156, physics.h:158-217).
 * Rows r-1..r-4 stay in a register ring, so every field crosses HBM once per
 * step; horizontal neighbours come from the adjacent lanes.
 *
 * The walk is a template over a lane policy LP: on the GPU a lane is one
 * work-item and pull_left/pull_right are DPP wavefront shifts; the CPU
 * emulation runs 64 lanes in lockstep with array shifts, so the CPU test
 * suite checks the same dataflow bit for bit against the oracle.
 *
 * Lane validity: a value pulled from outside the wavefront is 0. Faces are
 * valid on lanes 0..62, limiters on 1..62, new depths on 2..61 and time-step
 * candidates on 3..60, which are the lanes a wavefront writes.
 */
#ifndef FCA_WALK_H
#define FCA_WALK_H

#include "fca_device.h"

#define FCA_WAVE 64
#define FCA_HALO 3
#define FCA_OUTC (FCA_WAVE - 2 * FCA_HALO)

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
};

/* Per-lane state of the walk (registers on the GPU). Suffix k = row r-k. */
template <class R>
struct WalkLane {
    int c, colok, outcol;
    R z0, z1, z2, z3, z4;
    int m0, m1, m2, m3, m4;
    R h0, h1, h2;
    R n0, n1, n2, n3;
    R hn2, hn3, hn4;
    R qx0, qx1, qx2;
    R qy1, qy2, qy3;       /* faces (r-1 -> r), (r-2 -> r-1), (r-3 -> r-2) */
    R qd1_1, qd1_2, qd1_3; /* SE diagonal faces (q,c)->(q+1,c+1), q = r-1.. */
    R qd2_1, qd2_2, qd2_3; /* SW diagonal faces (q,c+1)->(q+1,c), q = r-1.. */
    R oW0, oE0, oN0, oS0, oW1, oE1, oN1, oS1, oW2, oE2, oN2, oS2;
    R s1, s2, s3;
    /* values pulled from neighbour lanes */
    R xz, xh, xn, xq, xr, xs, xt, xu, xv, xw, xy, xa;
    int xm, xm2, xm3, xm4, xm5, xm6;
    R dz1, dz2, dz3, dz4, dh1, dh2, dh3, dh4;
    int dm1, dm2, dm3, dm4;
    /* prefetched: mask of rows r+1 and r+2, z/h/n of row r+1, and cr/hmax of
     * row r-1 (the next update row) */
    int pm1, pm2;
    R pz, ph, pn, pcr, phm;
    /* accumulators of this lane's output cells */
    double v_rain, clamp, v_sea, v_bnd;
    long long cap;
    R dtc, vmax;
};

/* Walks one strip for one wavefront. strip covers output rows
 * [r_out0, r_out1) and columns [c_first + FCA_HALO, c_first + FCA_HALO + 58). */
template <class R, int K8, class LP>
FCA_HD void walk_strip(LP &lp, const WalkArgs<R> &a, int r_out0, int r_out1, int c_first)
{
    typedef WalkLane<R> L;
    const int nr = a.nr, nc = a.nc;
    const PhysK<R> &p = a.p;
    const StepVals<R> &sv = a.sv;
    const R dt = sv.dt, den = sv.den, rden = sv.rden, A = p.A;
    const int rb = (r_out0 - 3 > 0) ? r_out0 - 3 : 0;
    const int re = (r_out1 + 2 < nr + 2) ? r_out1 + 2 : nr + 2; /* inclusive */
    int r;

    lp.each([&](L &s, int lane) {
        s.c = c_first + lane;
        s.colok = s.c >= 0 && s.c < nc;
        s.outcol = s.colok && lane >= FCA_HALO && lane < FCA_WAVE - FCA_HALO;
        s.z0 = s.z1 = s.z2 = s.z3 = s.z4 = (R)0;
        s.m0 = s.m1 = s.m2 = s.m3 = s.m4 = FCA_NONE;
        s.h0 = s.h1 = s.h2 = (R)0;
        s.n0 = s.n1 = s.n2 = s.n3 = (R)0;
        s.hn2 = s.hn3 = s.hn4 = (R)0;
        s.qx0 = s.qx1 = s.qx2 = s.qy1 = s.qy2 = s.qy3 = (R)0;
        s.qd1_1 = s.qd1_2 = s.qd1_3 = s.qd2_1 = s.qd2_2 = s.qd2_3 = (R)0;
        s.oW0 = s.oE0 = s.oN0 = s.oS0 = s.oW1 = s.oE1 = s.oN1 = s.oS1 = (R)0;
        s.oW2 = s.oE2 = s.oN2 = s.oS2 = (R)0;
        s.s1 = s.s2 = s.s3 = (R)1;
        s.pcr = s.phm = (R)0;
        s.pz = s.ph = s.pn = (R)0;
        s.pm1 = s.pm2 = FCA_NONE;
        if (s.colok && rb < nr) {
            const size_t i = (size_t)rb * (size_t)nc + (size_t)s.c;
            s.pm1 = a.mask[i];
            if (s.pm1 == FCA_LAND) {
                s.pz = a.z[i], s.ph = a.h_in[i], s.pn = a.n[i];
            }
        }
        if (s.colok && rb + 1 < nr) {
            s.pm2 = a.mask[(size_t)(rb + 1) * (size_t)nc + (size_t)s.c];
        }
    });

    for (r = rb; r <= re; r++) {
        const int rowA = r < nr;                       /* row r exists */
        const int outA = r >= r_out0 && r < r_out1;    /* row r is written here */
        const int outB = r - 1 >= r_out0 && r - 1 < r_out1;
        const int qB = r - 1, qC = r - 2, qD = r - 3;

        /* ---- A: take row r (prefetched), prefetch ahead, form the faces ---- */
        lp.each([&](L &s, int) {
            s.m0 = s.pm1;
            if (s.m0 == FCA_LAND) {
                s.z0 = s.pz, s.h0 = s.ph, s.n0 = s.pn;
            }
            /* row r+1: z/h/n of LAND cells (mask fetched last iteration) */
            s.pm1 = s.pm2;
            if (s.colok && r + 1 < nr && s.pm1 == FCA_LAND) {
                const size_t i = (size_t)(r + 1) * (size_t)nc + (size_t)s.c;
                s.pz = a.z[i], s.ph = a.h_in[i], s.pn = a.n[i];
            }
            /* row r+2: mask */
            s.pm2 = (s.colok && r + 2 < nr) ? (int)a.mask[(size_t)(r + 2) * (size_t)nc +
                                                           (size_t)s.c]
                                            : FCA_NONE;
        });
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
            s.oW0 = s.oE0 = s.oN0 = s.oS0 = (R)0;
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
            if (p.open && s.colok && s.m0 == FCA_LAND &&
                (s.c == 0 || s.c == nc - 1 || r == 0 || r == nr - 1)) {
                int cw = 0, ce = 0, cn = 0, cs = 0;
                if (s.c == 0) {
                    s.oW0 = open_flux<R>(s.z0, s.h0, s.n0, dt, den, rden, p, cw);
                }
                if (s.c == nc - 1) {
                    s.oE0 = open_flux<R>(s.z0, s.h0, s.n0, dt, den, rden, p, ce);
                }
                if (r == 0) {
                    s.oN0 = open_flux<R>(s.z0, s.h0, s.n0, dt, den, rden, p, cn);
                }
                if (r == nr - 1) {
                    s.oS0 = open_flux<R>(s.z0, s.h0, s.n0, dt, den, rden, p, cs);
                }
                if (outA && s.outcol) {
                    s.cap += cw + ce + cn + cs;
                }
            }
        });

        /* ---- B: outflow volume and limiter of row r-1 -------------------- */
        lp.pull_left(&L::xt, &L::qx1);
        if (K8) {
            lp.pull_left(&L::xu, &L::qd1_2);
            lp.pull_left(&L::xv, &L::qd2_1);
        }
        if (qB >= 0 && qB < nr) {
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
                out += s.oW1 + s.oE1 + s.oN1 + s.oS1;
                out = out * dt;
                s.s1 = (s.m1 == FCA_LAND) ? limiter_s<R>(out, s.h1, A) : (R)1;
            });
        }

        /* ---- C: update row r-2 ------------------------------------------- */
        lp.pull_right(&L::xw, &L::s2);
        lp.pull_left(&L::xy, &L::s2);
        lp.pull_left(&L::xa, &L::qx2);
        lp.pull_right(&L::xm3, &L::m2);
        lp.pull_left(&L::xm4, &L::m2);
        if (K8) {
            lp.pull_right(&L::dz1, &L::s1); /* s of (q+1, c+1) */
            lp.pull_left(&L::dz2, &L::s1);  /* s of (q+1, c-1) */
            lp.pull_right(&L::dz3, &L::s3); /* s of (q-1, c+1) */
            lp.pull_left(&L::dz4, &L::s3);  /* s of (q-1, c-1) */
            lp.pull_left(&L::dh1, &L::qd1_3);
            lp.pull_left(&L::dh2, &L::qd2_2);
            lp.pull_right(&L::dm1, &L::m1);
            lp.pull_left(&L::dm2, &L::m3);
            lp.pull_left(&L::dm3, &L::m1);
            lp.pull_right(&L::dm4, &L::m3);
        }
        if (qC >= 0 && qC < nr) {
            const int outC = qC >= r_out0 && qC < r_out1;
            lp.each([&](L &s, int) {
                R net = (R)0, q, qos, rain, hnew;
                double sea = 0.0;
                const int land = s.m2 == FCA_LAND;
                if (!s.colok) {
                    return;
                }
                /* E, W, S, N[, SE, NW, SW, NE] in the reference's order */
                if (s.c + 1 < nc) {
                    q = s.qx2 * ((s.qx2 >= (R)0) ? s.s2 : s.xw);
                    net -= q;
                    if (land && s.xm3 == FCA_SEA && q > (R)0) {
                        sea += (double)dt * (double)q;
                    }
                }
                if (s.c > 0) {
                    q = s.xa * ((s.xa >= (R)0) ? s.xy : s.s2);
                    net += q;
                    if (land && s.xm4 == FCA_SEA && q < (R)0) {
                        sea += (double)dt * (double)(-q);
                    }
                }
                if (qC + 1 < nr) {
                    q = s.qy2 * ((s.qy2 >= (R)0) ? s.s2 : s.s1);
                    net -= q;
                    if (land && s.m1 == FCA_SEA && q > (R)0) {
                        sea += (double)dt * (double)q;
                    }
                }
                if (qC > 0) {
                    q = s.qy3 * ((s.qy3 >= (R)0) ? s.s3 : s.s2);
                    net += q;
                    if (land && s.m3 == FCA_SEA && q < (R)0) {
                        sea += (double)dt * (double)(-q);
                    }
                }
                if (K8) {
                    if (qC + 1 < nr && s.c + 1 < nc) {
                        q = s.qd1_2 * ((s.qd1_2 >= (R)0) ? s.s2 : s.dz1);
                        net -= q;
                        if (land && s.dm1 == FCA_SEA && q > (R)0) {
                            sea += (double)dt * (double)q;
                        }
                    }
                    if (qC > 0 && s.c > 0) {
                        q = s.dh1 * ((s.dh1 >= (R)0) ? s.dz4 : s.s2);
                        net += q;
                        if (land && s.dm2 == FCA_SEA && q < (R)0) {
                            sea += (double)dt * (double)(-q);
                        }
                    }
                    if (qC + 1 < nr && s.c > 0) {
                        q = s.dh2 * ((s.dh2 >= (R)0) ? s.s2 : s.dz2);
                        net -= q;
                        if (land && s.dm3 == FCA_SEA && q > (R)0) {
                            sea += (double)dt * (double)q;
                        }
                    }
                    if (qC > 0 && s.c + 1 < nc) {
                        q = s.qd2_3 * ((s.qd2_3 >= (R)0) ? s.dz3 : s.s2);
                        net += q;
                        if (land && s.dm4 == FCA_SEA && q < (R)0) {
                            sea += (double)dt * (double)(-q);
                        }
                    }
                }
                qos = s.oW2 * s.s2 + s.oE2 * s.s2 + s.oN2 * s.s2 + s.oS2 * s.s2;
                net -= qos;
                if (outC && s.outcol) {
                    s.v_sea += sea;
                    s.v_bnd += (double)dt * (double)qos;
                }
                if (!land) {
                    s.hn2 = s.h2;
                    return;
                }
                {
                    size_t i = (size_t)qC * (size_t)nc + (size_t)s.c;
                    double clamp_here = 0.0;
                    rain = s.pcr * sv.rain_f;
                    hnew = s.h2 + (dt * div_c(net, A, p.rA, p.fA) + rain);
                    if (hnew < (R)0) {
                        clamp_here = (double)(-hnew) * (double)A;
                        hnew = (R)0;
                    }
                    s.hn2 = hnew;
                    if (outC && s.outcol) {
                        s.v_rain += (double)A * (double)rain;
                        s.clamp += clamp_here;
                        a.h_out[i] = hnew;
                        if (hnew > s.phm) {
                            a.hmax[i] = hnew;
                            a.tmax[i] = sv.tmax_val;
                        }
                    }
                }
            });
        }

        /* prefetch cr and hmax of row r-1, which the next iteration updates */
        lp.each([&](L &s, int) {
            const int q = r - 1;
            if (s.colok && q >= 0 && q < nr && s.m1 == FCA_LAND) {
                const size_t i = (size_t)q * (size_t)nc + (size_t)s.c;
                s.pcr = a.cr[i];
                if (s.outcol && q >= r_out0 && q < r_out1) {
                    s.phm = a.hmax[i];
                }
            }
        });

        /* ---- D: next step's stable time step on row r-3 ------------------ */
        if (sv.need_cand && qD >= r_out0 && qD < r_out1 && qD < nr) {
            lp.pull_right(&L::xz, &L::z3);
            lp.pull_right(&L::xh, &L::hn3);
            lp.pull_right(&L::xm, &L::m3);
            lp.pull_left(&L::xq, &L::z3);
            lp.pull_left(&L::xr, &L::hn3);
            lp.pull_left(&L::xm2, &L::m3);
            if (K8) {
                lp.pull_right(&L::dz1, &L::z2);
                lp.pull_right(&L::dh1, &L::hn2);
                lp.pull_right(&L::dm1, &L::m2);
                lp.pull_left(&L::dz2, &L::z2);
                lp.pull_left(&L::dh2, &L::hn2);
                lp.pull_left(&L::dm2, &L::m2);
                lp.pull_right(&L::dz3, &L::z4);
                lp.pull_right(&L::dh3, &L::hn4);
                lp.pull_right(&L::dm3, &L::m4);
                lp.pull_left(&L::dz4, &L::z4);
                lp.pull_left(&L::dh4, &L::hn4);
                lp.pull_left(&L::dm4, &L::m4);
            }
            lp.each([&](L &s, int) {
                R la, ax = (R)0, smax, v;
                if (!s.outcol || s.m3 != FCA_LAND || s.hn3 < p.tau_h) {
                    return;
                }
                la = s.z3 + s.hn3;
                /* largest |level difference| per face length: max_k RN(a_k/d)
                 * = RN(max_k a_k / d) as rounding is monotonic, so one
                 * division per length replaces one per neighbour */
#define FCA_LEVEL_DIFF(cond, m_, z_, h_, acc_)                                \
    if ((cond) && (m_) != FCA_WALL) {                                           \
        R dl_ = la - level_of<R>((m_), (z_), (h_), p);                          \
        if (dl_ < (R)0) {                                                       \
            dl_ = -dl_;                                                         \
        }                                                                       \
        if (dl_ > acc_) {                                                       \
            acc_ = dl_;                                                         \
        }                                                                       \
    }
                FCA_LEVEL_DIFF(s.c + 1 < nc, s.xm, s.xz, s.xh, ax)
                FCA_LEVEL_DIFF(s.c > 0, s.xm2, s.xq, s.xr, ax)
                FCA_LEVEL_DIFF(qD + 1 < nr, s.m2, s.z2, s.hn2, ax)
                FCA_LEVEL_DIFF(qD > 0, s.m4, s.z4, s.hn4, ax)
                if (p.open && (s.c == 0 || s.c == nc - 1 || qD == 0 || qD == nr - 1)) {
                    R dl_ = la - (s.z3 - p.S_bnd * p.dx);
                    if (dl_ < (R)0) {
                        dl_ = -dl_;
                    }
                    if (dl_ > ax) {
                        ax = dl_;
                    }
                }
                smax = div_c(ax, p.dx, p.rdx, p.fdx);
                if (K8) {
                    R ad = (R)0, sd;
                    FCA_LEVEL_DIFF(qD + 1 < nr && s.c + 1 < nc, s.dm1, s.dz1, s.dh1, ad)
                    FCA_LEVEL_DIFF(qD + 1 < nr && s.c > 0, s.dm2, s.dz2, s.dh2, ad)
                    FCA_LEVEL_DIFF(qD > 0 && s.c + 1 < nc, s.dm3, s.dz3, s.dh3, ad)
                    FCA_LEVEL_DIFF(qD > 0 && s.c > 0, s.dm4, s.dz4, s.dh4, ad)
                    sd = div_c(ad, p.ddiag, p.rddiag, p.fddiag);
                    if (sd > smax) {
                        smax = sd;
                    }
                }
#undef FCA_LEVEL_DIFF
                /* dt_a = alpha*dx / v is non-increasing in v, so the lane keeps
                 * the largest v and divides once at the end of the strip */
                v = cell_speed<R>(s.hn3, s.n3, smax, p);
                if (v > s.vmax) {
                    s.vmax = v;
                }
            });
        }

        /* ---- rotate the ring ----------------------------------------------- */
        lp.each([&](L &s, int) {
            s.z4 = s.z3, s.z3 = s.z2, s.z2 = s.z1, s.z1 = s.z0;
            s.m4 = s.m3, s.m3 = s.m2, s.m2 = s.m1, s.m1 = s.m0;
            s.h2 = s.h1, s.h1 = s.h0;
            s.n3 = s.n2, s.n2 = s.n1, s.n1 = s.n0;
            s.hn4 = s.hn3, s.hn3 = s.hn2;
            s.qx2 = s.qx1, s.qx1 = s.qx0;
            s.qy3 = s.qy2, s.qy2 = s.qy1;
            s.qd1_3 = s.qd1_2, s.qd1_2 = s.qd1_1;
            s.qd2_3 = s.qd2_2, s.qd2_2 = s.qd2_1;
            s.oW2 = s.oW1, s.oE2 = s.oE1, s.oN2 = s.oN1, s.oS2 = s.oS1;
            s.oW1 = s.oW0, s.oE1 = s.oE0, s.oN1 = s.oN0, s.oS1 = s.oS0;
            s.s3 = s.s2, s.s2 = s.s1;
        });
    }
}

/* Zeroes a lane's accumulators; a wavefront may walk several strips. */
template <class R, class LP>
FCA_HD void walk_acc_init(LP &lp)
{
    lp.each([&](WalkLane<R> &s, int) {
        s.v_rain = s.clamp = s.v_sea = s.v_bnd = 0.0;
        s.cap = 0;
        s.dtc = (R)INFINITY;
        s.vmax = (R)0;
    });
}

/* The lane's time-step candidate from its largest speed: dt_a = alpha*dx/v is
 * non-increasing in v (alpha*dx > 0 is checked by the host). */
template <class R, class LP>
FCA_HD void walk_acc_finish(LP &lp, const PhysK<R> &p)
{
    lp.each([&](WalkLane<R> &s, int) {
        s.dtc = (s.vmax > (R)0) ? (p.alpha * p.dx) / s.vmax : (R)INFINITY;
    });
}

#endif
