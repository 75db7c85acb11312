/*
 * fca_device.h -- scalar arithmetic of the WCA step shared by the HIP kernels
 * (compiled for gfx950) and by the host-side emulation of the fused walk used
 * in the CPU test suite. Every expression keeps the reference's operation
 * order (src/kernels/physics.h, src/common/model.c) and is compiled with
 * contraction off, so a float build reproduces the reference float build bit
 * for bit wherever the libm calls agree.
 */
#ifndef FCA_DEVICE_H
#define FCA_DEVICE_H

#include "floodca_hip.h"

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FCA_HD __host__ __device__ inline
#else
#define FCA_HD inline
#endif

/* Mask value used for lanes whose column lies outside the grid. */
#define FCA_NONE 3

/* Run-constant inputs of the step controller (C-visible). */
typedef struct FcaCtlParams {
    double T_sim, out_every, D_h, phi;
    int hyeto, steps, U, n_force_days;
} FcaCtlParams;

/* Controller state carried from step to step on the device (C-visible). */
typedef struct FcaDevCtrl {
    double t;              /* time at the start of the next step */
    long long dt_min_hits; /* dt_base clamped up to dt_min */
    int step;              /* steps taken so far */
    int force_day;         /* forcing day whose Rday is on the device, -1 none */
    int state;             /* FCA_CTL_RUN, _DONE or _PAUSE */
    int want_day;          /* with _PAUSE: the day to load */
    int nrec;              /* step records written */
    int hbuf;              /* which of the two depth buffers holds h */
    int armed;             /* a step was taken since the state was uploaded */
    int pad_;
    double dt_base;        /* held step, in the run's precision */
    double tmax_last;      /* tmax value of the last step taken, (real)(t + dt) */
} FcaDevCtrl;

enum { FCA_CTL_RUN = 0, FCA_CTL_DONE = 1, FCA_CTL_PAUSE = 2 };

/* Per-workgroup balance terms, accumulated over the steps of a run. */
typedef struct FcaAccSlot {
    double v_rain, clamp, v_sea, v_bnd;
    long long cap;
    long long pad;
} FcaAccSlot;

/* Whether a run admits the LAND form of the fused walk. The thresholds keep
 * the arguments of the face's divisions, square root and hf^(2/3) inside the
 * ranges where the device shortcuts need no range test (|dL| >= tau_l >=
 * 2^-80, RN(tau_l/d) >= 2^-96 for the longest face, hf >= tau_h >= 2^-15,
 * g*hf >= 2^-100 with g in [2^-80, 2^80]), and every LAND roughness in
 * [1e-12, 2^100] lets the face roughness go without its clamp and its
 * reciprocal without a range test. */
static inline int fca_land_phys_ok(const FcaPhys *ph)
{
    const real dx = ph->dx, dmax = (ph->K == 8) ? (real)sqrt(2.0) * dx : dx;
    const real A = ph->A;
    /* the row test of the LAND form (|z + h| <= 2^20, |h| <= 2^14 on every
     * lane, fca_walk.h) bounds a live face's |dL| by 2^21.1 and hf by
     * 2^14 + 1; with A <= 2^58 and dx >= 2^-50 this keeps RN(|dL|*A) <= 2^80
     * and |dL|/d <= 2^126, and RN(tau_l*A) >= 2^-80 keeps RN(|dL|*A) >= 2^-80 */
    return ph->tau_l >= (real)0x1p-80 && ph->tau_l / dmax >= (real)0x1p-96 &&
           ph->tau_h >= (real)0x1p-15 && ph->g >= (real)0x1p-80 && ph->g <= (real)0x1p80 &&
           A <= (real)0x1p58 && dx >= (real)0x1p-50 && (real)(ph->tau_l * A) >= (real)0x1p-80;
}
static inline int fca_land_n_ok(const real *n, const uint8_t *mask, size_t N)
{
    size_t i;
    for (i = 0; i < N; i++) {
        if (mask[i] == FCA_LAND && !(n[i] >= (real)1e-12 && n[i] <= (real)0x1p100)) {
            return 0;
        }
    }
    return 1;
}
static inline int fca_land_form_ok(const FcaPhys *ph, const real *n, const uint8_t *mask,
                                   size_t N)
{
    return fca_land_phys_ok(ph) && fca_land_n_ok(n, mask, N);
}

#ifdef __cplusplus

#if defined(__HIPCC__)
#define FCA_HDI __host__ __device__ inline __attribute__((always_inline))
#else
#define FCA_HDI inline
#endif

/* Marks a block the compiler must branch around rather than evaluate
 * unconditionally (its work is rare and costly). */
#if defined(__HIP_DEVICE_COMPILE__)
#define FCA_COLD_PATH() __asm__ volatile("")
#else
#define FCA_COLD_PATH() ((void)0)
#endif

/* Physical constants of one run in the kernel's precision, plus the derived
 * widths of physics.h:37-48. */
template <class R>
struct PhysK {
    R dx, A, g, alpha, tau_h, tau_l, z_sea, dt_max, dt_min, S_bnd;
    R w;     /* face width: dx (K=4) or (sqrt2-1)*dx on every face (K=8) */
    R ddiag; /* sqrt2*dx */
    int U, K, open;
    /* Division by the run constants dx, sqrt2*dx and A: reciprocals and, per
     * divisor, whether q0 = a*r, q = fma(fma(-q0, d, a), r, q0) was verified
     * on the host to equal a/d for every float mantissa (div_const_exact). */
    R rdx, rddiag, rA;
    int fdx, fddiag, fA;
    /* the LAND form may be used: fca_land_form_ok and, in float, fdx, fA
     * and (K=8) fddiag */
    int fast;
    /* RN(g*(1-2^-16)): the LAND form's test that a face speed is clearly
     * below the critical speed (face_land) */
    R gk;
};

/* a/d for a run-constant d with reciprocal rd = RN(1/d). With ok set (the
 * two-fma form was verified exact for every mantissa of a), the device uses
 * it inside the range where the scaling argument holds; otherwise, and on the
 * host, a plain correctly rounded division. */
FCA_HD float div_c(float a, float d, float rd, int ok)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float aa = __builtin_fabsf(a);
    if (ok && ((aa >= 0x1p-80f && aa <= 0x1p80f) || aa == 0.0f)) {
        const float q0 = a * rd; /* a signed zero stays itself */
        const float q = __builtin_fmaf(__builtin_fmaf(-q0, d, a), rd, q0);
        return (aa == 0.0f) ? q0 : q;
    }
#else
    (void)rd, (void)ok;
#endif
    return a / d;
}
FCA_HD double div_c(double a, double d, double rd, int ok)
{
    (void)rd, (void)ok;
    return a / d;
}

/* div_c for a divisor whose flag is set and a dividend known to be at least
 * 2^-80 in magnitude (the LAND form's |dL| >= tau_l). */
FCA_HD float div_c_big(float a, float d, float rd)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_fabsf(a) <= 0x1p80f) {
        const float q0 = a * rd;
        return __builtin_fmaf(__builtin_fmaf(-q0, d, a), rd, q0);
    }
#else
    (void)rd;
#endif
    return a / d;
}
FCA_HD double div_c_big(double a, double d, double rd)
{
    (void)rd;
    return a / d;
}

/* Correctly rounded 1/b: on the device a hardware reciprocal with one Newton
 * correction, exact for every float in [2^-100, 2^100] (checked exhaustively
 * over all mantissas on the GPU, tests/test_gpu_math.py). */
FCA_HD float rcp_cr(float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float ab = __builtin_fabsf(b);
    if (ab >= 0x1p-100f && ab <= 0x1p100f) {
        const float y = __builtin_amdgcn_rcpf(b);
        return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    }
#endif
    return 1.0f / b;
}
FCA_HD double rcp_cr(double b) { return 1.0 / b; }

/* rcp_cr for b known to lie in [2^-100, 2^100]. */
FCA_HD float rcp_in_range(float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
#else
    return 1.0f / b;
#endif
}
FCA_HD double rcp_in_range(double b) { return 1.0 / b; }

FCA_HDI int same_bits(float a, float b)
{
    uint32_t x, y;
    __builtin_memcpy(&x, &a, 4);
    __builtin_memcpy(&y, &b, 4);
    return x == y;
}
FCA_HDI int same_bits(double a, double b)
{
    uint64_t x, y;
    __builtin_memcpy(&x, &a, 8);
    __builtin_memcpy(&y, &b, 8);
    return x == y;
}
FCA_HDI float fabs_r(float x) { return __builtin_fabsf(x); }
FCA_HDI double fabs_r(double x) { return __builtin_fabs(x); }
FCA_HDI float fmax_r(float a, float b) { return __builtin_fmaxf(a, b); }
FCA_HDI double fmax_r(double a, double b) { return __builtin_fmax(a, b); }

#if defined(__HIPCC__)
/* A DPP move with control CTRL on rows ROW_MASK; lanes of other rows, and
 * lanes whose source is outside their row, read +0. */
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_or_zero(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, true));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_or_zero(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, ROW_MASK,
                                               0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float lane63(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

/* uniform: the largest v over the wavefront, for v >= +0 and not NaN on
 * every lane, with every lane active. A max-scan within each row of 16 lanes
 * (shifts of 1, 2, 4, 8), then lane 15 of rows 0 and 2 into rows 1 and 3 and
 * lane 31 into rows 2 and 3 leave the maximum in lane 63. Lanes with no
 * source read +0, the identity of this max. */
template <class R>
__device__ inline R wave_max_nonneg(R v)
{
    v = fmax_r(v, dpp_or_zero<0x111, 0xf>(v)); /* row_shr:1 */
    v = fmax_r(v, dpp_or_zero<0x112, 0xf>(v)); /* row_shr:2 */
    v = fmax_r(v, dpp_or_zero<0x114, 0xf>(v)); /* row_shr:4 */
    v = fmax_r(v, dpp_or_zero<0x118, 0xf>(v)); /* row_shr:8 */
    v = fmax_r(v, dpp_or_zero<0x142, 0xa>(v)); /* row_bcast:15 */
    v = fmax_r(v, dpp_or_zero<0x143, 0xc>(v)); /* row_bcast:31 */
    return lane63(v);
}
/* The float form in integers: floats >= +0 are ordered as their bit patterns
 * are as signed integers, so an integer maximum gives the float maximum with
 * no NaN quieting of its operands (which a float maximum of shifted or
 * loop-carried values needs), and +0 is the integer 0. */
__device__ inline float wave_max_nonneg(float v)
{
    int b = __float_as_int(v);
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x111, 0xf, 0xf, true)); /* row_shr:1 */
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x112, 0xf, 0xf, true)); /* row_shr:2 */
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x114, 0xf, 0xf, true)); /* row_shr:4 */
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x118, 0xf, 0xf, true)); /* row_shr:8 */
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x142, 0xa, 0xf, true)); /* row_bcast:15 */
    b = max(b, __builtin_amdgcn_update_dpp(0, b, 0x143, 0xc, 0xf, true)); /* row_bcast:31 */
    return __int_as_float(__builtin_amdgcn_readlane(b, 63));
}
#endif

/* A wavefront-wide test inside a lane's code. On the device: uniform, whether
 * c holds on some active lane. On the host the emulation runs the lanes one
 * at a time, so it is c itself, and a branch on wave_any followed by a test
 * of c takes the same lanes either way. */
FCA_HDI bool wave_any(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
}

/* x if x > 0, else +0: the term an outflow sum adds for x. */
template <class R>
FCA_HDI R pos_part(R x)
{
    return (x > (R)0) ? x : (R)0;
}
/* float: the same value by an integer maximum of the bit pattern (a positive
 * float is a positive integer; -0 and negative floats are negative integers,
 * whose maximum with 0 is +0); no NaN quieting of x is needed. */
FCA_HDI float pos_part(float x)
{
    int32_t b;
    __builtin_memcpy(&b, &x, 4);
    b = (b > 0) ? b : 0;
    __builtin_memcpy(&x, &b, 4);
    return x;
}

/* Rough sqrt (relative error < 2^-21) used only to skip exact work. */
FCA_HD float sqrt_rough(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
FCA_HD double sqrt_rough(double x) { return sqrt(x); }

#define FCA_FILTER_MARGIN ((R)0x1p-18)

/* ---- float-mode libm detours --------------------------------------------
 *
 * In float mode the reference evaluates hf^(2/3) as (float)cbrt((double)hf^2)
 * and the Manning slope term as (float)sqrt((double)S) (physics.h:19-24, 106,
 * 205). Both equal the correctly rounded float result for every float input:
 * sqrt by the double-rounding theorem (53 >= 2*24+2), cbrt by an exhaustive
 * check of all 2^31-2^23 positive finite floats against glibc 2.35
 * (tests/test_exhaustive_libm.py). The device therefore computes correctly
 * rounded float results from cheap hardware approximations:
 *   cbrt_sq: v_log/v_exp seed, one Newton step in double, and a certificate
 *     that the double result is far enough from a float rounding boundary;
 *     otherwise an exact midpoint test (m^3 vs hf^2 in double-double);
 *   sqrt: the compiler's correctly rounded sqrtf.
 * The host side (the CPU emulation) keeps the reference's double formulas;
 * the GPU test suite checks the device functions against numpy's libm over
 * every positive float. */

/* sign of m^3 - x, exact for m with <= 26 significant bits and x a double */
FCA_HD int cube_cmp(double m, double x)
{
    double m2 = m * m;
    double p = m2 * m;
    double e = __builtin_fma(m2, m, -p);
    double d = (p - x) + e;
    return (d > 0.0) - (d < 0.0);
}

FCA_HD float f32_step(float f, int k)
{
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    b += (uint32_t)k;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

/* Correctly rounded hf^(2/3) from a float estimate y within one ulp. */
FCA_HD float cbrt_sq_fix(float y, double x)
{
    float yn = f32_step(y, 1), yp = f32_step(y, -1);
    if (cube_cmp(((double)y + (double)yn) * 0.5, x) < 0) {
        return yn;
    }
    if (cube_cmp(((double)yp + (double)y) * 0.5, x) > 0) {
        return yp;
    }
    return y;
}

#if defined(__HIP_DEVICE_COMPILE__)
/* Fast-path domain of cbrt_sq: 2^-15 <= hf <= 2^15 keeps |log2 hf| < 16, so
 * the v_log/v_exp seed is within ~2^-20 and one Newton step lands within
 * ~2^-41 of hf^(2/3). */
#define FCA_CBRT_LO 0x1p-15f
#define FCA_CBRT_HI 0x1p15f
/* Reject when the double result is within 2^14 double ulps (2^-38 relative)
 * of a float rounding midpoint. */
#define FCA_CBRT_GUARD 16384

/* One Newton step from the hardware seed: a^(2/3) to ~2^-41 relative. The
 * step y0 + r/(3 y0^2) is one fma (one rounding fewer than a product and a
 * sum, so within the same bound), and r = a^2 - y0^3 another (y0^2 is exact
 * in double). */
__device__ inline double cbrt_sq_newton(float a, double sq)
{
    const float y0 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(a) * (2.0f / 3.0f));
    const double yd = (double)y0;
    const double r = __builtin_fma(-(yd * yd), yd, sq); /* a^2 - y0^3, one rounding */
    return __builtin_fma(r, (double)__builtin_amdgcn_rcpf(3.0f * y0 * y0), yd);
}
#endif

FCA_HD float cbrt_sq(float hf)
{
    double sq = (double)hf * (double)hf;
#if defined(__HIP_DEVICE_COMPILE__)
    const float a = __builtin_fabsf(hf);
    if (a >= FCA_CBRT_LO && a <= FCA_CBRT_HI) {
        const double y1 = cbrt_sq_newton(a, sq);
        uint64_t bits;
        __builtin_memcpy(&bits, &y1, 8);
        /* the 29 bits below float precision, relative to the midpoint */
        const int lo = (int)((uint32_t)bits & 0x1FFFFFFFu) - 0x10000000;
        if (lo > FCA_CBRT_GUARD || lo < -FCA_CBRT_GUARD) {
            return (float)y1;
        }
        return cbrt_sq_fix((float)y1, sq);
    }
    return cbrt_sq_fix((float)cbrt(sq), sq);
#else
    return (float)cbrt(sq);
#endif
}
FCA_HD double cbrt_sq(double hf)
{
    double sq = hf * hf;
    return cbrt(sq);
}

/* cbrt_sq for hf >= 2^-15 (the LAND form's hf >= tau_h): the fast path's
 * lower range test is dropped. */
FCA_HD float cbrt_sq_big(float hf)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double sq = (double)hf * (double)hf;
    if (hf <= FCA_CBRT_HI) {
        const double y1 = cbrt_sq_newton(hf, sq);
        uint64_t bits;
        __builtin_memcpy(&bits, &y1, 8);
        const int lo = (int)((uint32_t)bits & 0x1FFFFFFFu) - 0x10000000;
        if (lo > FCA_CBRT_GUARD || lo < -FCA_CBRT_GUARD) {
            return (float)y1;
        }
        return cbrt_sq_fix((float)y1, sq);
    }
    return cbrt_sq_fix((float)cbrt(sq), sq);
#else
    return cbrt_sq(hf);
#endif
}
FCA_HD double cbrt_sq_big(double hf) { return cbrt_sq(hf); }

#if defined(__HIP_DEVICE_COMPILE__)
/* Lane mask of a condition. Given a single compare it is that compare's own
 * result (no vector instruction); a condition built from several compares
 * costs a select and a compare to turn into a mask, so callers ballot each
 * compare and combine the masks with scalar & and |. */
__device__ inline uint64_t lane_mask(bool c)
{
    return __builtin_amdgcn_ballot_w64(c);
}

/* Whether the low 29 bits of a double estimate y (those below float
 * precision) lie within FCA_CBRT_GUARD double ulps of the float rounding
 * midpoint, i.e. (float)y might not be the correctly rounded value. */
__device__ inline bool near_float_midpoint(double y)
{
    uint64_t bits;
    __builtin_memcpy(&bits, &y, 8);
    const int lo = (int)((uint32_t)bits & 0x1FFFFFFFu) - 0x10000000;
    return !(lo > FCA_CBRT_GUARD || lo < -FCA_CBRT_GUARD);
}

/* cbrt_sq_big without branches: the correctly rounded hf^(2/3) for
 * 2^-15 <= hf <= 2^15 unless *fix is set (near a midpoint, or hf out of that
 * range), where the caller must take cbrt_sq_big. */
__device__ inline float cbrt_sq_cert(float hf, bool &fix)
{
    const double y1 = cbrt_sq_newton(hf, (double)hf * (double)hf);
    fix = fix || near_float_midpoint(y1) || !(hf <= FCA_CBRT_HI);
    return (float)y1;
}

/* (float)sqrt((double)g*(double)h), the critical speed, without branches:
 * x = g*h is exact in double; s = the hardware float sqrt of (float)x is
 * within 2^-22 of sqrt(x), one Newton step y = s + (x - s^2)/(2s) (the
 * residual exact in one fma, the reciprocal the hardware's) brings it within
 * 2^-44, and y is the correctly rounded float unless it lies within 2^-38 of
 * a float midpoint, where *fix is set and the caller must take crit_speed.
 * x within [2^-100, 2^100] keeps every step normal (the LAND form has
 * h >= 2^-15 and g in [2^-80, 2^80]). Checked against the double expression
 * for every h in [2^-15, 2^15] (tests/test_gpu_math.py). */
__device__ inline float crit_speed_cert(float g, float h, bool &fix)
{
    const double x = (double)g * (double)h;
    const float s = __builtin_amdgcn_sqrtf((float)x);
    const double sd = (double)s;
    const double r = __builtin_fma(-sd, sd, x);
    const double y = __builtin_fma(r, (double)__builtin_amdgcn_rcpf(2.0f * s), sd);
    fix = fix || near_float_midpoint(y);
    return (float)y;
}
#endif

FCA_HD float sqrt_via_double(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrtf(x);
#else
    return (float)sqrt((double)x);
#endif
}
FCA_HD double sqrt_via_double(double x) { return sqrt(x); }

/* Correctly rounded sqrt of x in [2^-96, 2^126]: the hardware square root is
 * within one ulp, and the residuals of its two neighbours pick the correctly
 * rounded one; no scaling or special-value handling is needed in this range
 * (checked for every float of the range on the GPU, tests/test_gpu_math.py). */
FCA_HD float sqrt_in_range(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = f32_step(s, -1), sp = f32_step(s, 1);
    const float em = __builtin_fmaf(-sm, s, x);
    const float ep = __builtin_fmaf(-sp, s, x);
    const float t = (em <= 0.0f) ? sm : s;
    return (ep > 0.0f) ? sp : t;
#else
    return (float)sqrt((double)x);
#endif
}
FCA_HD double sqrt_in_range(double x) { return sqrt(x); }

/* (real)sqrt((double)g*(double)h), the critical velocity (physics.h:108,207). */
FCA_HD float crit_speed(float g, float h) { return (float)sqrt((double)g * (double)h); }
FCA_HD double crit_speed(double g, double h) { return sqrt(g * h); }

/* RN(num/den) for the step constant den > 0 with rden = RN(1/den). On the
 * device: the estimate num*rden refined by one remainder step, then
 * certified by its exact remainder r = num - q*den (exact because q is within
 * an ulp of the quotient): q is the correctly rounded quotient when
 * |r| < (1-2^-10) * den * u/2, u the ulp of q towards zero (no quotient of two
 * floats lies exactly halfway). Anything else takes a plain division. */
FCA_HDI float div_step_cr(float num, float den, float rden)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float q = 0.0f;
    int ok = 0;
    if (__builtin_fabsf(num) >= 0x1p-80f && __builtin_fabsf(num) <= 0x1p80f) {
        const float q0 = num * rden;
        const float q1 = __builtin_fmaf(__builtin_fmaf(-q0, den, num), rden, q0);
        const float r1 = __builtin_fmaf(-q1, den, num);
        const float u = __builtin_fabsf(q1) - __builtin_fabsf(f32_step(q1, -1));
        ok = __builtin_fabsf(r1) < u * (den * (0.5f - 0x1p-11f));
        q = q1;
    }
    if (!ok) {
        FCA_COLD_PATH();
        q = num / den;
    }
    return q;
#else
    (void)rden;
    return num / den;
#endif
}
FCA_HDI double div_step_cr(double num, double den, double rden)
{
    (void)rden;
    return num / den;
}

/* Whether num/den for this step's den may be taken as the two-fma refinement
 * of num*RN(1/den) without a certificate: den in [2^-40, 2^40]. With
 * rden = RN(1/den) and q0 = RN(num*rden), q = fma(fma(-q0, den, num), rden,
 * q0) is the correctly rounded quotient for every divisor and dividend
 * significand (Markstein's refinement; checked on the GPU over all
 * 2^23 x 2^23 significand pairs, tests/test_gpu_math.py), and dividends in
 * [2^-80, 2^80] keep every intermediate a normal number. */
FCA_HD int step_den_fast(float den)
{
    return den >= 0x1p-40f && den <= 0x1p40f;
}
FCA_HD int step_den_fast(double) { return 0; }

/* div_step_cr with the step's den_fast flag: the certificate only where the
 * flag is off or num is out of range. */
FCA_HDI float div_step_fast(float num, float den, float rden, int den_fast)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float an = __builtin_fabsf(num);
    if (den_fast && an >= 0x1p-80f && an <= 0x1p80f) {
        const float q0 = num * rden;
        return __builtin_fmaf(__builtin_fmaf(-q0, den, num), rden, q0);
    }
    return div_step_cr(num, den, rden);
#else
    (void)rden, (void)den_fast;
    return num / den;
#endif
}
FCA_HDI double div_step_fast(double num, double den, double rden, int)
{
    return div_step_cr(num, den, rden);
}

/* The level-equalisation cap Qcap = |dL|*A / ((2K)*dt) (physics.h:113-121). */
template <class R>
FCA_HDI R level_cap(R qmag, R adL, R dt, R den, R rden, const PhysK<R> &p, int &cap,
                    int den_fast = 0)
{
    if (dt > (R)0 && p.K > 0) {
        const R qcap = div_step_fast(adL * p.A, den, rden, den_fast);
        if (qcap < qmag) {
            cap = 1;
            qmag = qcap;
        }
    }
    return qmag;
}

/* min(v, vcr) with vcr = (real)sqrt((double)g*(double)hf) (physics.h:
 * 108-111). The exact critical speed is only needed when v is not clearly
 * below it: fl(v*v) < fl(fl(g*hf)*(1-2^-16)) implies v < sqrt(g*hf)*(1-2^-18)
 * < vcr, as each rounding is within 2^-24 (2^-53 in double); g*hf >= 2^-100
 * keeps the comparison clear of underflow. GH_OK: the caller knows
 * g*hf >= 2^-100 (the LAND form: hf >= tau_h >= 2^-15 and g >= 2^-80). */
template <class R, int GH_OK = 0>
FCA_HDI R crit_cap(R v, R hf, const PhysK<R> &p)
{
    const R gh = p.g * hf;
    if (!(v * v < gh * ((R)1 - (R)0x1p-16) && (GH_OK || gh >= (R)0x1p-100))) {
        FCA_COLD_PATH();
        const R vcr = crit_speed(p.g, hf);
        if (v > vcr) {
            v = vcr;
        }
    }
    return v;
}

/* Face flux a->b, positive = a to b (physics.h:51-125). den = (real)(2K)*dt
 * and rden = RN(1/den) are evaluated once per step by the caller; d is dx or
 * sqrt2*dx, rd its reciprocal and dok its div_c flag. Two filters skip exact
 * work whose outcome is already decided: the critical-velocity cap is only
 * evaluated exactly when v is within 2^-18 of a rough sqrt(g*hf), and the
 * level-equalisation cap only when a rough (|dL|*A)*rden is not clearly above
 * Qmag; the rough values are within 2^-21 of the exact ones, so a skipped
 * test could not have changed the result. */
template <class R>
FCA_HD R face_flux(R za, R ha, R na, int ma, R zb, R hb, R nb, int mb, R d, R rd, int dok,
                   R dt, R den, R rden, const PhysK<R> &p, int &cap)
{
    /* Cheap tests first, all evaluated; the reference's early returns
     * (walls, dead band, dry face, upwind sea) become one `dead` predicate
     * and the expensive part runs under a single branch. Each expression is
     * the reference's own, so a live face's value is unchanged. */
    const int sea_a = ma == FCA_SEA, sea_b = mb == FCA_SEA;
    const R bed_a = sea_a ? p.z_sea : za;
    const R bed_b = sea_b ? p.z_sea : zb;
    const R lev_a = bed_a + (sea_a ? (R)0 : ha);
    const R lev_b = bed_b + (sea_b ? (R)0 : hb);
    const R dL = lev_a - lev_b;
    const R adL = (dL < (R)0) ? -dL : dL;
    const R hf = ((lev_a > lev_b) ? lev_a : lev_b) - ((bed_a > bed_b) ? bed_a : bed_b);
    const int up_sea = (dL > (R)0) ? sea_a : sea_b;
    const int dead = (ma == FCA_WALL) | (mb == FCA_WALL) | (adL < p.tau_l) | (hf < p.tau_h) |
                     up_sea;
    R qmag = (R)0;
    cap = 0;
    if (!dead) {
        R nf = sea_a ? nb : (sea_b ? na : (R)0.5 * (na + nb));
        R v;
        if (nf < (R)1e-12) {
            nf = (R)1e-12;
        }
        v = rcp_cr(nf) * cbrt_sq(hf) * sqrt_via_double(div_c(adL, d, rd, dok));
        v = crit_cap<R>(v, hf, p);
        qmag = v * hf * p.w;
        qmag = level_cap<R>(qmag, adL, dt, den, rden, p, cap);
        qmag = (dL > (R)0) ? qmag : -qmag;
    }
    return qmag;
}

/* Face flux between two LAND cells given their beds z and levels l = z + h,
 * for runs where fca_land_form_ok holds: the expressions of face_flux with
 * the wall, sea and clamp terms (all inactive) left out. */
template <class R>
FCA_HDI R face_land_exact(R za, R la, R na, R zb, R lb, R nb, R d, R rd, R dt, R den, R rden,
                    int den_fast, const PhysK<R> &p, int &cap)
{
    const R dL = la - lb;
    /* |dL| as a magnitude: where face_flux keeps a -0 this gives +0, and
     * either zero is below tau_l, so the face is dead in both */
    const R adL = fabs_r(dL);
    const R hf = ((la > lb) ? la : lb) - ((za > zb) ? za : zb);
    R qmag = (R)0;
    cap = 0;
    if (!((adL < p.tau_l) | (hf < p.tau_h))) {
        const R nf = (R)0.5 * (na + nb);
        R v = rcp_in_range(nf) * cbrt_sq_big(hf) * sqrt_in_range(div_c_big(adL, d, rd));
        v = crit_cap<R, 1>(v, hf, p);
        qmag = v * hf * p.w;
        qmag = level_cap<R>(qmag, adL, dt, den, rden, p, cap, den_fast);
        qmag = (dL > (R)0) ? qmag : -qmag;
    }
    return qmag;
}

/* face_land as the LAND form evaluates it. On the device (float) every lane
 * takes the same straight-line path -- dead faces computed on their own
 * inputs and masked at the end, the cube root and the critical speed in
 * their certified branch-free forms (cbrt_sq_cert, crit_speed_cert), Qcap
 * by Markstein's refinement -- and each lane raises a flag where a
 * certificate fails or an input leaves the ranges those forms assume; one
 * ballot then sends the flagged lanes, rarely any, through face_land_exact.
 * A wavefront with no live face skips the arithmetic. The values are
 * face_land_exact's, bit for bit. */
template <class R>
FCA_HDI R face_land(R za, R la, R na, R zb, R lb, R nb, R d, R rd, R dt, R den, R rden,
                    int den_fast, const PhysK<R> &p, bool &cap)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        const float dL = la - lb;
        /* la > lb exactly when la - lb > 0 (finite floats, gradual
         * underflow), so one compare serves the upper level and the sign */
        const bool up = dL > 0.0f;
        const float adL = __builtin_fabsf(dL);
        const float hf = (up ? la : lb) - ((za > zb) ? za : zb);
        /* flags are bools combined with && and ||, so each stays a lane mask
         * and its ballot costs no vector instruction */
        const bool live = !(adL < (float)p.tau_l) && !(hf < (float)p.tau_h);
        const uint64_t mlive = lane_mask(!(adL < (float)p.tau_l)) & lane_mask(!(hf < (float)p.tau_h));
        cap = 0;
        if (mlive == 0) {
            return (R)0;
        }
        {
            /* dead lanes compute on their own (possibly out-of-range) inputs:
             * every value they produce, their flags included, is masked by
             * the live lanes below, and floating-point exceptions do not trap.
             * mfix collects, one compare at a time, the lanes whose certificate
             * fails. The ranges the fast forms assume hold on every live lane
             * by the LAND form's row test (|z + h| <= 2^20 and |h| <= 2^14,
             * so |dL| < 2^21.1 and hf <= 2^14 + 1) and the run constants
             * fca_land_phys_ok admits: hf in [2^-15, 2^15] for the cube
             * root, |dL|/d in [2^-96, 2^126] for the square root, RN(|dL|*A)
             * in [2^-80, 2^80] for Qcap, g*hf in [2^-100, 2^100] for the
             * critical speed */
            uint64_t mfix = den_fast ? 0ull : ~0ull;
            const float nf = 0.5f * ((float)na + (float)nb);
            const float q0 = adL * (float)rd;
            const float sl = __builtin_fmaf(__builtin_fmaf(-q0, (float)d, adL), (float)rd, q0);
            const double y1 = cbrt_sq_newton(hf, (double)hf * (double)hf);
            mfix |= lane_mask(near_float_midpoint(y1));
            float v = rcp_in_range(nf) * (float)y1 * sqrt_in_range(sl);
            float qmag, x, c0, qcap;
            {
                /* the critical speed only where some lane's v is not clearly
                 * below it, a uniform branch. crit_cap's test with the
                 * constant folded: gk = RN(g*(1-2^-16)), and fl(v*v) <
                 * fl(hf*gk) gives v^2 < g*hf*(1-2^-16)(1+2^-24)^3 <
                 * g*hf*(1-2^-17), so v < sqrt(g*hf)*(1-2^-18) < vcr */
                if ((lane_mask(!(v * v < hf * (float)p.gk)) & mlive) != 0) {
                    bool cfix = false;
                    const float vcr = crit_speed_cert((float)p.g, hf, cfix);
                    mfix |= lane_mask(cfix);
                    if (v > vcr) {
                        v = vcr;
                    }
                }
            }
            qmag = v * hf * (float)p.w;
            x = adL * (float)p.A;
            c0 = x * (float)rden;
            qcap = __builtin_fmaf(__builtin_fmaf(-c0, (float)den, x), (float)rden, c0);
            cap = live && qcap < qmag;
            qmag = (qcap < qmag) ? qcap : qmag;
            qmag = up ? qmag : -qmag;
            if ((mfix & mlive) != 0) {
                FCA_COLD_PATH();
                const uint64_t lane_bit = 1ull << __builtin_amdgcn_mbcnt_hi(
                                              ~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
                if ((mfix & mlive & lane_bit) != 0) {
                    int cx;
                    const R q = face_land_exact<R>(za, la, na, zb, lb, nb, d, rd, dt, den, rden,
                                                   den_fast, p, cx);
                    cap = cx != 0;
                    return q;
                }
            }
            return live ? (R)qmag : (R)0; /* a dead face is +0, as face_land_exact */
        }
    }
#endif
    int cx;
    const R q = face_land_exact<R>(za, la, na, zb, lb, nb, d, rd, dt, den, rden, den_fast, p, cx);
    cap = cx != 0;
    return q;
}

/* Two faces of the LAND form at once (the east face of row r and the face
 * from row r-1 to row r, or the two diagonals between them; both of length
 * d): on each face face_land's value, bit for bit, with the two faces'
 * arithmetic in the same basic blocks and one uniform branch where
 * face_land takes one per face (no live face, a critical speed to take, a
 * certificate to redo), so that the compiler can interleave the two faces'
 * independent chains of dependent instructions. */
template <class R>
FCA_HDI void face_land_pair(const R (&za)[2], const R (&la)[2], const R (&na)[2],
                            const R (&zb)[2], const R (&lb)[2], const R (&nb)[2], R d, R rd,
                            R dt, R den, R rden, int den_fast, const PhysK<R> &p, R (&q)[2],
                            bool (&cap)[2])
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(R) == 4) {
        float dL[2], adL[2], hf[2], v[2], qm[2];
        bool up[2], live[2];
        uint64_t mlive[2], mfix[2];
        for (int k = 0; k < 2; k++) {
            dL[k] = (float)la[k] - (float)lb[k];
            up[k] = dL[k] > 0.0f;
            adL[k] = __builtin_fabsf(dL[k]);
            hf[k] = (up[k] ? (float)la[k] : (float)lb[k]) -
                    (((float)za[k] > (float)zb[k]) ? (float)za[k] : (float)zb[k]);
            live[k] = !(adL[k] < (float)p.tau_l) && !(hf[k] < (float)p.tau_h);
            mlive[k] = lane_mask(!(adL[k] < (float)p.tau_l)) & lane_mask(!(hf[k] < (float)p.tau_h));
            cap[k] = false;
        }
        if ((mlive[0] | mlive[1]) == 0) {
            q[0] = q[1] = (R)0;
            return;
        }
        /* as in face_land: dead lanes compute on their own inputs, masked
         * at the end; mfix collects the lanes whose certificate fails */
        for (int k = 0; k < 2; k++) {
            mfix[k] = den_fast ? 0ull : ~0ull;
            const float nf = 0.5f * ((float)na[k] + (float)nb[k]);
            const float q0 = adL[k] * (float)rd;
            const float sl = __builtin_fmaf(__builtin_fmaf(-q0, (float)d, adL[k]), (float)rd, q0);
            const double y1 = cbrt_sq_newton(hf[k], (double)hf[k] * (double)hf[k]);
            mfix[k] |= lane_mask(near_float_midpoint(y1));
            v[k] = rcp_in_range(nf) * (float)y1 * sqrt_in_range(sl);
        }
        /* the critical speed of a face where some live lane's v is not
         * clearly below it (face_land's test), face by face */
        for (int k = 0; k < 2; k++) {
            if ((lane_mask(!(v[k] * v[k] < hf[k] * (float)p.gk)) & mlive[k]) != 0) {
                bool cfix = false;
                const float vcr = crit_speed_cert((float)p.g, hf[k], cfix);
                mfix[k] |= lane_mask(cfix);
                if (v[k] > vcr) {
                    v[k] = vcr;
                }
            }
        }
        for (int k = 0; k < 2; k++) {
            float qmag = v[k] * hf[k] * (float)p.w;
            const float x = adL[k] * (float)p.A;
            const float c0 = x * (float)rden;
            const float qcap = __builtin_fmaf(__builtin_fmaf(-c0, (float)den, x), (float)rden, c0);
            cap[k] = live[k] && qcap < qmag;
            qmag = (qcap < qmag) ? qcap : qmag;
            qm[k] = up[k] ? qmag : -qmag;
        }
        if (((mfix[0] & mlive[0]) | (mfix[1] & mlive[1])) != 0) {
            FCA_COLD_PATH();
            const uint64_t lane_bit = 1ull << __builtin_amdgcn_mbcnt_hi(
                                          ~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
            for (int k = 0; k < 2; k++) {
                if ((mfix[k] & mlive[k] & lane_bit) != 0) {
                    int cx;
                    qm[k] = (float)face_land_exact<R>(za[k], la[k], na[k], zb[k], lb[k], nb[k], d,
                                                      rd, dt, den, rden, den_fast, p, cx);
                    cap[k] = cx != 0;
                }
            }
        }
        for (int k = 0; k < 2; k++) {
            /* a dead face is +0, as face_land_exact (a lane the exact path
             * took is live) */
            q[k] = live[k] ? (R)qm[k] : (R)0;
        }
        return;
    }
#endif
    for (int k = 0; k < 2; k++) {
        q[k] = face_land<R>(za[k], la[k], na[k], zb[k], lb[k], nb[k], d, rd, dt, den, rden,
                            den_fast, p, cap[k]);
    }
}

/* Open-boundary virtual face of a LAND edge cell: z_b = z_a - S_bnd*dx,
 * h_b = 0, n_b = n_a, d = dx; inflow clipped to 0 (model.c:166-177). */
template <class R>
FCA_HD R open_flux(R za, R ha, R na, R dt, R den, R rden, const PhysK<R> &p, int &cap)
{
    R zb = za - p.S_bnd * p.dx;
    R q = face_flux<R>(za, ha, na, FCA_LAND, zb, (R)0, na, FCA_LAND, p.dx, p.rdx, p.fdx, dt,
                       den, rden, p, cap);
    return (q < (R)0) ? (R)0 : q;
}

/* Limiter factor (physics.h:127-134). */
template <class R>
FCA_HD R limiter_s(R out_vol, R h, R A)
{
    R hold = h * A;
    if (out_vol > hold && out_vol > (R)0) {
        FCA_COLD_PATH();
        return hold / out_vol;
    }
    return (R)1;
}

/* limiter_s with a uniform branch on the device: the division runs on every
 * lane when some lane of the wavefront limits, and not at all otherwise, so
 * no lane mask is saved and restored (the LAND form). Same values. */
template <class R>
FCA_HDI R limiter_s_wave(R out_vol, R h, R A)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const R hold = h * A;
    const bool lim = out_vol > hold && out_vol > (R)0;
    R s = (R)1;
    if ((lane_mask(out_vol > hold) & lane_mask(out_vol > (R)0)) != 0) {
        const R q = hold / out_vol;
        s = lim ? q : s;
    }
    return s;
#else
    return limiter_s<R>(out_vol, h, A);
#endif
}

/* div_c for the LAND form's net/A (the flag of A set, checked by the host):
 * the two-fma form on every lane, a zero passed through with its sign, and a
 * correctly rounded division, on the lanes that need it, only when some
 * dividend lies outside [2^-80, 2^80] (a uniform branch). Same values. */
FCA_HDI float div_c_wave(float a, float d, float rd)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float aa = __builtin_fabsf(a);
    const float q0 = a * rd;
    float q = __builtin_fmaf(__builtin_fmaf(-q0, d, a), rd, q0);
    const bool slow = !((aa >= 0x1p-80f && aa <= 0x1p80f) || aa == 0.0f);
    q = (aa == 0.0f) ? q0 : q;
    if (((lane_mask(!(aa >= 0x1p-80f)) | lane_mask(!(aa <= 0x1p80f))) & ~lane_mask(aa == 0.0f)) !=
        0) {
        const float qd = a / d;
        q = slow ? qd : q;
    }
    return q;
#else
    return div_c(a, d, rd, 1);
#endif
}
FCA_HDI double div_c_wave(double a, double d, double rd) { return div_c(a, d, rd, 1); }

/* |la - lb| / d (physics.h:147-154). */
template <class R>
FCA_HD R abs_slope(R la, R lb, R d)
{
    R dL = la - lb;
    if (dL < (R)0) {
        dL = -dL;
    }
    return dL / d;
}

/* Level a neighbour presents to the slope test (physics.h:136-145). */
template <class R>
FCA_HD R level_of(int m, R z, R h, const PhysK<R> &p)
{
    return (m == FCA_SEA) ? p.z_sea : ((m == FCA_WALL) ? z : z + h);
}

/* Manning speed of a wet LAND cell from its largest slope, capped at the
 * critical speed and floored at 1e-6 (physics.h:196-214). */
template <class R>
FCA_HD R cell_speed(R h, R n, R smax, const PhysK<R> &p)
{
    R v;
    if (n < (R)1e-12) {
        n = (R)1e-12;
    }
    if (smax < (R)0) {
        smax = (R)0;
    }
    v = rcp_cr(n) * cbrt_sq(h) * sqrt_via_double(smax);
    v = crit_cap<R>(v, h, p);
    if (v < (R)1e-6) {
        v = (R)1e-6;
    }
    return v;
}

/* Stable step of a wet LAND cell from its largest slope
 * (physics.h:193-217); the caller has checked h >= tau_h. */
template <class R>
FCA_HD R cell_dt_from_smax(R h, R n, R smax, const PhysK<R> &p)
{
    if (h < p.tau_h) {
        return p.dt_max;
    }
    return p.alpha * p.dx / cell_speed<R>(h, n, smax, p);
}

/* ---- the step controller (model.c:363-382, 599-601; rain.c:13-106) ------ */

/* What one launch does, derived from FcaDevCtrl and the time-step candidate. */
template <class R>
struct StepVals {
    double t, F, t_new;
    R dt, dt_base, tmax_val, rain_f, den, rden;
    R tmax_prev; /* tmax value of the previous step (valid when armed) */
    int active, hit, need_cand, step;
    int den_fast; /* den admits div_step_fast (step_den_fast) */
    int armed;    /* a step was taken since the state was uploaded */
};

/* rain.c:13-28 */
FCA_HD double tri_rate(double t, double T, double tp)
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

/* rain.c:30-67; D_h and phi arrive already widened from real. */
FCA_HD double hyeto_integral(const FcaCtlParams &c, double t, double dt)
{
    double t1, T, tp, sum, a, b;
    if (dt <= 0.0) {
        return 0.0;
    }
    t1 = t + dt;
    if (c.hyeto == FCA_HYETO_UNIFORM) {
        return dt / 86400.0;
    }
    T = c.D_h * 3600.0;
    tp = c.phi * T;
    if (t1 <= 0.0 || t >= T) {
        return 0.0;
    }
    if (t < 0.0) {
        t = 0.0;
    }
    if (t1 > T) {
        t1 = T;
    }
    sum = 0.0;
    a = t;
    while (a < t1 - 1e-15) {
        b = (a < tp && t1 > tp) ? tp : t1;
        sum += 0.5 * (tri_rate(a, T, tp) + tri_rate(b, T, tp)) * (b - a);
        a = b;
    }
    return sum;
}

/* rain.c:69-93 */
FCA_HD double next_event_time(const FcaCtlParams &c, double t)
{
    double t_next = c.T_sim, t_day, t_out;
    t_day = (floor(t / FCA_DAY_S) + 1.0) * FCA_DAY_S;
    if (t_day - t < 1e-9) {
        t_day += FCA_DAY_S;
    }
    if (t_day < t_next) {
        t_next = t_day;
    }
    if (c.out_every > 0.0) {
        t_out = (floor(t / c.out_every) + 1.0) * c.out_every;
        if (t_out - t < 1e-9) {
            t_out += c.out_every;
        }
        if (t_out < t_next) {
            t_next = t_out;
        }
    }
    return t_next;
}

/* rain.c:95-106 */
FCA_HD int forcing_day_index(double t)
{
    int d;
    if (t < 0.0) {
        return 0;
    }
    d = (int)floor(t / FCA_DAY_S);
    return d < 0 ? 0 : d;
}

/* One evaluation of fca_step's head (model.c:338-382) and tail (599-601).
 * cand is the smallest per-cell stable step of the current state (+inf when
 * no LAND cell is wet). Fills sv for this launch and nx for the next one. A
 * launch is inactive when the run is over, paused for a forcing reload, or
 * the event clamp leaves no time to step (model.c:379-381). */
template <class R>
FCA_HD void ctrl_eval(const FcaDevCtrl &in, R cand, const FcaCtlParams &c, const PhysK<R> &p,
                      StepVals<R> &sv, FcaDevCtrl &nx)
{
    R dtb, dt;
    double t_next, remain;
    int hit = 0;
    nx = in;
    sv.active = 0;
    sv.need_cand = 0;
    sv.step = in.step;
    sv.t = in.t;
    sv.armed = in.armed;
    sv.tmax_prev = (R)in.tmax_last;
    if (in.state != FCA_CTL_RUN) {
        return;
    }
    if (c.steps > 0 && in.step >= c.steps) {
        return;
    }
    if (c.steps <= 0 && in.t >= c.T_sim - 1e-15) {
        nx.state = FCA_CTL_DONE;
        return;
    }
    if (c.n_force_days > 0) {
        int d = forcing_day_index(in.t);
        if (d >= c.n_force_days) {
            d = c.n_force_days - 1;
        }
        if (d != in.force_day) {
            nx.state = FCA_CTL_PAUSE;
            nx.want_day = d;
            return;
        }
    }
    dtb = (R)in.dt_base;
    if (in.step % (c.U > 0 ? c.U : 1) == 0) {
        R dtmin = p.dt_max;
        if (cand < dtmin) {
            dtmin = cand;
        }
        if (dtmin < p.dt_min) {
            dtmin = p.dt_min;
            nx.dt_min_hits += 1;
        }
        if (dtmin > p.dt_max) {
            dtmin = p.dt_max;
        }
        dtb = dtmin;
        nx.dt_base = (double)dtb;
    }
    dt = dtb;
    t_next = next_event_time(c, in.t);
    remain = t_next - in.t;
    if (remain <= (double)dt) {
        if (remain < 0) {
            remain = 0;
        }
        dt = (R)remain;
        hit = 1;
    }
    sv.dt_base = dtb;
    if (dt <= (R)0) {
        if (c.steps <= 0) {
            nx.state = FCA_CTL_DONE; /* fca_run stops when t does not advance */
        }
        return;
    }
    sv.active = 1;
    sv.hit = hit;
    sv.dt = dt;
    sv.F = hyeto_integral(c, in.t, (double)dt);
    sv.rain_f = (R)sv.F;
    sv.tmax_val = (R)(in.t + (double)dt);
    sv.den = (R)(2 * p.K) * dt;
    sv.rden = (R)1 / sv.den;
    sv.den_fast = step_den_fast(sv.den);
    sv.t_new = hit ? t_next : in.t + (double)dt;
    sv.need_cand = ((in.step + 1) % (c.U > 0 ? c.U : 1)) == 0;
    nx.t = sv.t_new;
    nx.step = in.step + 1;
    nx.armed = 1;
    nx.tmax_last = (double)sv.tmax_val;
}

/* Order-preserving key of a time-step candidate for an unsigned atomic min
 * (NaN candidates are dropped by the caller, as the reference's '<' test
 * drops them). */
FCA_HD uint32_t cand_key(float v)
{
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
FCA_HD float cand_unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float v;
    __builtin_memcpy(&v, &b, 4);
    return v;
}
FCA_HD unsigned long long cand_key(double v)
{
    unsigned long long b;
    __builtin_memcpy(&b, &v, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
FCA_HD double cand_unkey(unsigned long long k)
{
    unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    __builtin_memcpy(&v, &b, 8);
    return v;
}

template <class R>
struct CandKey;
template <>
struct CandKey<float> {
    typedef uint32_t type;
};
template <>
struct CandKey<double> {
    typedef unsigned long long type;
};

/* Stable step of cell (r, c) read straight from grid arrays: dt_max unless a
 * wet LAND cell (model.c:123-137, physics.h:158-217). */
template <class R>
FCA_HD R cell_dt_at(const R *z, const R *h, const R *n, const uint8_t *mask, int r, int c,
                    int nr, int nc, const PhysK<R> &p)
{
    const int dr[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    const int dc[8] = {1, -1, 0, 0, 1, -1, 1, -1};
    size_t a = (size_t)r * (size_t)nc + (size_t)c;
    int k, nk = (p.K == 8) ? 8 : 4;
    R la, smax = (R)0, sl;
    if (mask[a] != FCA_LAND || h[a] < p.tau_h) {
        return p.dt_max;
    }
    la = z[a] + h[a];
    for (k = 0; k < nk; k++) {
        int rb = r + dr[k], cb = c + dc[k];
        size_t b;
        if (rb < 0 || rb >= nr || cb < 0 || cb >= nc) {
            continue;
        }
        b = (size_t)rb * (size_t)nc + (size_t)cb;
        if (mask[b] == FCA_WALL) {
            continue;
        }
        sl = abs_slope<R>(la, level_of<R>(mask[b], z[b], h[b], p), k < 4 ? p.dx : p.ddiag);
        if (sl > smax) {
            smax = sl;
        }
    }
    if (p.open && (c == 0 || c == nc - 1 || r == 0 || r == nr - 1)) {
        sl = abs_slope<R>(la, z[a] - p.S_bnd * p.dx, p.dx);
        if (sl > smax) {
            smax = sl;
        }
    }
    return cell_dt_from_smax<R>(h[a], n[a], smax, p);
}

/* PhysK from the ABI's FcaPhys (physics.h:37-48 for the widths). */
template <class R>
inline PhysK<R> make_phys(const FcaPhys &f)
{
    PhysK<R> p;
    p.dx = (R)f.dx, p.A = (R)f.A, p.g = (R)f.g, p.alpha = (R)f.alpha;
    p.tau_h = (R)f.tau_h, p.tau_l = (R)f.tau_l, p.z_sea = (R)f.z_sea;
    p.dt_max = (R)f.dt_max, p.dt_min = (R)f.dt_min, p.S_bnd = (R)f.S_bnd;
    p.w = (f.K == 8) ? ((R)sqrt(2.0) - (R)1) * p.dx : p.dx;
    p.ddiag = (R)sqrt(2.0) * p.dx;
    p.U = f.U, p.K = f.K, p.open = f.boundary == FCA_BND_OPEN;
    p.rdx = (R)1 / p.dx, p.rddiag = (R)1 / p.ddiag, p.rA = (R)1 / p.A;
    p.fdx = p.fddiag = p.fA = 0;
    p.fast = 0;
    p.gk = p.g * ((R)1 - (R)0x1p-16);
    return p;
}

#endif /* __cplusplus */

#endif
