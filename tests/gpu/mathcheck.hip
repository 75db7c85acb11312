/*
 * mathcheck.hip -- test-only kernels that evaluate the device scalar math of
 * fca_device.h over ranges of float bit patterns, so the GPU suite can compare
 * every input against libm (tests/test_gpu_math.py). Never part of the product.
 */
#include <hip/hip_runtime.h>

#include "fca_device.h"

__global__ void k_eval(int kind, uint32_t first, uint32_t count, float g, float d, float rd,
                       int dok, float *out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    uint32_t b = first + (uint32_t)i;
    float x = __uint_as_float(b), y;
    switch (kind) {
    case 0: y = cbrt_sq(x); break;
    case 1: y = sqrt_via_double(x); break;
    case 2: y = crit_speed(g, x); break;
    case 3: y = rcp_cr(x); break;
    case 4: y = div_c(x, d, rd, dok); break;
#if defined(__HIP_DEVICE_COMPILE__)
    case 6: { /* relative error of the Newton value inside the fast domain */
        double sq = (double)x * (double)x, y1 = cbrt_sq_newton(x, sq), z = cbrt(sq);
        y = (x >= FCA_CBRT_LO && x <= FCA_CBRT_HI) ? (float)__builtin_fabs(y1 / z - 1.0) : 0.0f;
        break;
    }
#endif
    default: y = sqrt_rough(x); break;
    }
    out[i] = y;
}

/* Evaluates kind over [first, first+count) into host out. Returns 0 or a HIP error. */
extern "C" int mc_eval(int kind, uint32_t first, uint32_t count, float g, float d, float rd,
                       int dok, float *out)
{
    float *dv = nullptr;
    hipError_t e = hipMalloc((void **)&dv, (size_t)count * sizeof(float));
    if (e != hipSuccess) {
        return (int)e;
    }
    hipLaunchKernelGGL(k_eval, dim3((count + 255) / 256), dim3(256), 0, 0, kind, first, count,
                       g, d, rd, dok, dv);
    e = hipGetLastError();
    if (e == hipSuccess) {
        e = hipMemcpy(out, dv, (size_t)count * sizeof(float), hipMemcpyDeviceToHost);
    }
    (void)hipFree(dv);
    return (int)e;
}
