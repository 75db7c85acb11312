/*
 * mathcheck.hip -- test-only kernels that evaluate the device scalar math of
 * fca_device.h and fca_walk.h over ranges of float bit patterns, so the GPU
 * suite can compare every input against libm or check a property on every
 * input (tests/test_gpu_math.py). Never part of the product.
 */
#include <hip/hip_runtime.h>

#include "fca_walk.h"

/* Parameters of one evaluation; x is the float whose bit pattern is the
 * work-item's index. */
struct McArgs {
    float g, d, rd, e, f;
    int dok, k;
};

__global__ void k_eval(int kind, uint32_t first, uint32_t count, McArgs a, float *out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    uint32_t b = first + (uint32_t)i;
    float x = __uint_as_float(b), y;
    PhysK<float> p;
    p.g = a.g;
    switch (kind) {
    case 0: y = cbrt_sq(x); break;
    case 1: y = sqrt_via_double(x); break;
    case 2: y = crit_speed(a.g, x); break;
    case 3: y = rcp_cr(x); break;
    case 4: y = div_c(x, a.d, a.rd, a.dok); break;
#if defined(__HIP_DEVICE_COMPILE__)
    case 6: { /* relative error of the Newton value inside the fast domain */
        double sq = (double)x * (double)x, y1 = cbrt_sq_newton(x, sq), z = cbrt(sq);
        y = (x >= FCA_CBRT_LO && x <= FCA_CBRT_HI) ? (float)__builtin_fabs(y1 / z - 1.0) : 0.0f;
        break;
    }
#endif
    case 7: y = sqrt_in_range(x); break;
    case 8: y = rcp_in_range(x); break;
    case 9: y = div_c_big(x, a.d, a.rd); break;
    case 10: { /* crit_cap near the critical speed: v = vcr stepped by k ulps */
        const float v = f32_step(crit_speed(a.g, x), a.k);
        y = crit_cap<float>(v, x, p);
        break;
    }
    case 11: { /* stage D bound: 1 where it would skip a cell whose exact
                * speed exceeds thr = (exact speed stepped by k ulps) */
        const float h = x, n = a.e, ax = a.f;
        const float s_hi = (ax * a.rd) * (1.0f + 0x1p-20f);
        const float smax = div_c(ax, a.d, a.rd, a.dok);
        const float v = cell_speed<float>(h, n, smax, p);
        const float thr = f32_step(v, a.k);
        const int need = speed_may_exceed<float>(h, n, s_hi, ax == 0.0f, thr, a.g);
        y = (!need && v > thr) ? 1.0f : (need ? 0.5f : 0.0f);
        break;
    }
    case 12: y = div_step_cr(x, a.d, a.rd); break;
#if defined(__HIP_DEVICE_COMPILE__)
    case 13: /* absolute error of the hardware log2 against log2 in double */
        y = (float)__builtin_fabs((double)__builtin_amdgcn_logf(x) - log2((double)x));
        break;
#endif
    case 14: { /* stage D bound in logarithms: as case 11 with speed_may_exceed_fast */
        const float h = x, n = a.e, ax = a.f;
        const float s_hi = (ax * a.rd) * (1.0f + 0x1p-20f);
        const float smax = div_c(ax, a.d, a.rd, a.dok);
        const float v = cell_speed<float>(h, n, smax, p);
        const float thr = f32_step(v, a.k);
        const int need = speed_may_exceed_fast<float>(h, n, s_hi, ax == 0.0f, thr,
                                                      speed_log_thr(thr),
                                                      speed_log_half_g<float>(a.g), a.g);
        y = (!need && v > thr) ? 1.0f : (need ? 0.5f : 0.0f);
        break;
    }
    case 16: y = div_step_fast(x, a.d, a.rd, step_den_fast(a.d)); break;
#if defined(__HIP_DEVICE_COMPILE__)
    case 17: { /* certified critical speed; -1 where it defers to crit_speed */
        bool fix = false;
        const float v = crit_speed_cert(a.g, x, fix);
        y = fix ? -1.0f : v;
        break;
    }
    case 18: { /* certified cube root; -1 where it defers */
        bool fix = false;
        const float v = cbrt_sq_cert(x, fix);
        y = fix ? -1.0f : v;
        break;
    }
#endif
    default: y = sqrt_rough(x); break;
    }
    out[i] = y;
}

/* For each divisor significand m = first + i (divisor 1.m in [1, 2)): the
 * number of dividend significands x in [1, 2) for which the two-fma
 * refinement of x*RN(1/d) is not the correctly rounded x/d. A refinement q
 * is correctly rounded iff its exact remainder r = x - q*d (one fma) is
 * below d times half the gap from q to its neighbour on r's side (no
 * quotient of two floats lies halfway between two floats, so the strict
 * test decides it). */
__global__ void k_div_pairs(uint32_t first, uint32_t count, uint32_t *fails)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    const float d = __uint_as_float(0x3f800000u | ((first + (uint32_t)i) & 0x7FFFFFu));
    const float r = 1.0f / d;
    uint32_t bad = 0;
    for (uint32_t m = 0; m < (1u << 23); m++) {
        const float x = __uint_as_float(0x3f800000u | m);
        const float q0 = x * r;
        const float q = __builtin_fmaf(__builtin_fmaf(-q0, d, x), r, q0);
        const float rem = __builtin_fmaf(-q, d, x);
        /* the rounding boundary on the side of the exact quotient: half the
         * gap to q's neighbour there (the gap below a power of two is half
         * the gap above it) */
        const float ulp = (rem < 0.0f) ? q - __uint_as_float(__float_as_uint(q) - 1u)
                                       : __uint_as_float(__float_as_uint(q) + 1u) - q;
        bad += !(__builtin_fabsf(rem) < 0.5f * ulp * d);
    }
    fails[i] = bad;
}

extern "C" int mc_div_pairs(uint32_t first, uint32_t count, uint32_t *out)
{
    uint32_t *dv = nullptr;
    hipError_t e = hipMalloc((void **)&dv, (size_t)count * sizeof(uint32_t));
    if (e != hipSuccess) {
        return (int)e;
    }
    hipLaunchKernelGGL(k_div_pairs, dim3((count + 255) / 256), dim3(256), 0, 0, first, count, dv);
    e = hipGetLastError();
    if (e == hipSuccess) {
        e = hipMemcpy(out, dv, (size_t)count * sizeof(uint32_t), hipMemcpyDeviceToHost);
    }
    (void)hipFree(dv);
    return (int)e;
}

/* Evaluates kind over [first, first+count) into host out. Returns 0 or a HIP error. */
extern "C" int mc_eval(int kind, uint32_t first, uint32_t count, const McArgs *a, float *out)
{
    float *dv = nullptr;
    hipError_t e = hipMalloc((void **)&dv, (size_t)count * sizeof(float));
    if (e != hipSuccess) {
        return (int)e;
    }
    hipLaunchKernelGGL(k_eval, dim3((count + 255) / 256), dim3(256), 0, 0, kind, first, count,
                       *a, dv);
    e = hipGetLastError();
    if (e == hipSuccess) {
        e = hipMemcpy(out, dv, (size_t)count * sizeof(float), hipMemcpyDeviceToHost);
    }
    (void)hipFree(dv);
    return (int)e;
}

/* Every work-item of a full wavefront writes the wavefront's maximum of in
 * (wave_max_nonneg, the stage-D threshold reduction). */
template <class T>
__global__ void k_wave_max(const T *in, T *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    out[i] = wave_max_nonneg(in[i]);
}

/* n values (a multiple of 256, every wavefront full) of float (dbl = 0) or
 * double (dbl = 1) from host in to host out. Returns 0 or a HIP error. */
extern "C" int mc_wave_max(int dbl, const void *in, void *out, uint32_t n)
{
    if (n == 0 || n % 256 != 0) {
        return (int)hipErrorInvalidValue;
    }
    const size_t bytes = (size_t)n * (dbl ? sizeof(double) : sizeof(float));
    void *dv = nullptr;
    hipError_t e = hipMalloc(&dv, 2 * bytes);
    if (e != hipSuccess) {
        return (int)e;
    }
    e = hipMemcpy(dv, in, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (dbl) {
            hipLaunchKernelGGL(k_wave_max<double>, dim3(n / 256), dim3(256), 0, 0,
                               (const double *)dv, (double *)dv + n);
        } else {
            hipLaunchKernelGGL(k_wave_max<float>, dim3(n / 256), dim3(256), 0, 0,
                               (const float *)dv, (float *)dv + n);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        e = hipMemcpy(out, (char *)dv + bytes, bytes, hipMemcpyDeviceToHost);
    }
    (void)hipFree(dv);
    return (int)e;
}
