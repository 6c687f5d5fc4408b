/*
 * spac_fk_testing.h — test hooks exported by libspac_fk.so.
 *
 * Not part of the reference's interface: these entry points let the parity
 * tests evaluate the device copies of the shared FP64 primitives (spac_math.h)
 * and of the coefficient/boundary device twins, so that CPU and GPU bits can be
 * compared function by function before whole paths are compared.
 */
#ifndef SPAC_FK_TESTING_H
#define SPAC_FK_TESTING_H

#include <stddef.h>
#include <stdint.h>

#include "spac_fk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fn: 0 exp(x), 1 log(x), 2 sinpi(x), 3 cospi(x), 4 atan2(x, y), 5 sqrt(x),
 * 6 x / y, 7 Franke u(x, y), 8 first Brownian normal of (seed 0, stream 0,
 * path x, step y).  Host arrays of n values; runs one device kernel. */
spac_status spac_debug_math(int32_t fn, const double* x, const double* y, double* out, int64_t n);

/* Device memory for tests of the device-resident entry points (a process that
 * loads this library's HIP runtime needs no second runtime for it): allocate
 * `bytes` on the current device, free, and synchronous copies to and from it. */
spac_status spac_debug_device_alloc(size_t bytes, void** out);
spac_status spac_debug_device_free(void* ptr);
spac_status spac_debug_copy_to_device(void* dst, const void* src, size_t bytes);
spac_status spac_debug_copy_to_host(void* dst, const void* src, size_t bytes);
/* device-to-device copy of `bytes` (a plain copy of known size: the calibration dispatch of
 * the HBM traffic counters) */
spac_status spac_debug_copy_device(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
