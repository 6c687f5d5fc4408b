/*
 * fk_debug.hip — device evaluation of the shared FP64 primitives for the
 * function-by-function parity tests (include/spac_fk_testing.h).
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/spac_fk_testing.h"
#include "fk_device.h"
#include "fk_internal.h"

namespace spac {

__global__ void debug_math_kernel(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], b = y[i], r = 0.0;
    switch (fn) {
    case 0: r = spac_exp(a); break;
    case 1: r = spac_log(a); break;
    case 2: r = spac_sinpi(a); break;
    case 3: r = spac_cospi(a); break;
    case 4: r = spac_atan2(a, b); break;
    case 5: r = __builtin_sqrt(a); break;
    case 6: r = a / b; break;
    case 7: r = CoefFranke::u(a, b); break;
    case 8: {
        double n0, n1;
        spac_normal_pair(0, 0, (uint32_t)a, (uint32_t)b, &n0, &n1);
        r = n0;
        break;
    }
    case 9: r = spac_exp_finite(a); break;
    case 10: r = spac_log_normal(a); break;
    case 11: r = spac_gobet_inside(a, b, 1.7) ? 1.0 : 0.0; break;
    default: r = 0.0; break;
    }
    out[i] = r;
}

} // namespace spac

extern "C" spac_status spac_debug_math(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
    using namespace spac;
    if (!x || !y || !out || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    size_t bytes = (size_t)n * 8;
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
        hipMalloc(&dout, bytes) != hipSuccess) {
        (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dout);
        return set_error(SPAC_E_OUT_OF_MEMORY, "hipMalloc failed");
    }
    spac_status rc = SPAC_OK;
    if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = set_error(SPAC_E_HIP, "copy to device failed");
    } else {
        int64_t blocks = (n + 255) / 256;
        hipLaunchKernelGGL(debug_math_kernel, dim3((unsigned)blocks), dim3(256), 0, 0, fn, dx, dy, dout, n);
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = set_error(SPAC_E_HIP, "debug kernel failed: %s", hipGetErrorString(e));
        else if (hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_error(SPAC_E_HIP, "copy to host failed");
    }
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dout);
    return rc;
}

extern "C" spac_status spac_debug_device_alloc(size_t bytes, void** out) {
    using namespace spac;
    if (!out || bytes == 0) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) return set_error(SPAC_E_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SPAC_OK;
}

extern "C" spac_status spac_debug_device_free(void* ptr) {
    using namespace spac;
    if (!ptr) return SPAC_OK;
    hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "hipFree failed: %s", hipGetErrorString(e));
    return SPAC_OK;
}

extern "C" spac_status spac_debug_copy_to_device(void* dst, const void* src, size_t bytes) {
    using namespace spac;
    if (!dst || !src) return set_error(SPAC_E_INVALID_ARG, "null argument");
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "copy to device failed: %s", hipGetErrorString(e));
    return SPAC_OK;
}

extern "C" spac_status spac_debug_copy_to_host(void* dst, const void* src, size_t bytes) {
    using namespace spac;
    if (!dst || !src) return set_error(SPAC_E_INVALID_ARG, "null argument");
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "copy to host failed: %s", hipGetErrorString(e));
    return SPAC_OK;
}

extern "C" spac_status spac_debug_copy_device(void* dst, const void* src, size_t bytes) {
    using namespace spac;
    if (!dst || !src) return set_error(SPAC_E_INVALID_ARG, "null argument");
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "device copy failed: %s", hipGetErrorString(e));
    return SPAC_OK;
}
