/*
 * nvshmem_device.h -- device API of the MI355X-native OpenSHMEM subset.
 *
 * Drop-in replacement for the prototype's device header (reference
 * nvshmem_device.h:34-75).  Names, argument order (target, source), the
 * element count `len` as a signed int, and C linkage are unchanged.  The
 * definitions live in libnvshmem.a as relocatable device code; compile callers
 * with `hipcc -fgpu-rdc` and link with `--hip-link -lnvshmem`.
 *
 * Semantics (reference README:100-174, nvshmem_device.cu:431-584):
 *  - every call is complete for the calling thread alone; no cooperation
 *    between threads is assumed (the cooperative forms are in nvshmemx_device.h);
 *  - p / put return once the stores are issued; nvshmem_quiet() orders and
 *    completes them before any later store of the calling thread;
 *  - nvshmem_int_wait blocks while *ivar == cmp_value;
 *    nvshmem_int_wait_until blocks until *ivar == cmp_value and ignores `cmp`.
 *  - nvshmem_trace_begin / nvshmem_trace_end are block-collective: every
 *    thread of the block calls them.
 */
#ifndef NVSHMEM_AMD_NVSHMEM_DEVICE_H
#define NVSHMEM_AMD_NVSHMEM_DEVICE_H

#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference nvshmem_device.h:38-40 */
__device__ void nvshmem_trace_begin(void);
__device__ void nvshmem_trace_end(void);
__device__ void nvshmem_quiet(void);

/* single-element put, reference nvshmem_device.h:42-47 */
__device__ void nvshmem_int_p(int *target, const int value, int pe);
__device__ void nvshmem_float_p(float *target, const float value, int pe);
__device__ void nvshmem_double_p(double *target, const double value, int pe);

/* single-element get, reference nvshmem_device.h:49-54 */
__device__ int nvshmem_int_g(int *source, int pe);
__device__ float nvshmem_float_g(float *source, int pe);
__device__ double nvshmem_double_g(double *source, int pe);

/* bulk put: target (symmetric, on `pe`) <- source (local); reference nvshmem_device.h:56-61 */
__device__ void nvshmem_int_put(int *target, int *source, int len, int pe);
__device__ void nvshmem_float_put(float *target, float *source, int len, int pe);
__device__ void nvshmem_double_put(double *target, double *source, int len, int pe);

/* bulk get: target (local) <- source (symmetric, on `pe`); reference nvshmem_device.h:63-68 */
__device__ void nvshmem_int_get(int *target, int *source, int len, int pe);
__device__ void nvshmem_float_get(float *target, float *source, int len, int pe);
__device__ void nvshmem_double_get(double *target, double *source, int len, int pe);

/* point-to-point synchronisation, reference nvshmem_device.h:70-71 */
__device__ void nvshmem_int_wait(int *ivar, int cmp_value);
__device__ void nvshmem_int_wait_until(int *ivar, int cmp, int cmp_value);

#ifdef __cplusplus
}
#endif

#endif /* NVSHMEM_AMD_NVSHMEM_DEVICE_H */
