/*
 * nvshmemx_device.h -- device-side extensions with no counterpart in the
 * reference's device header (reference nvshmem_device.h:38-71 declares only the
 * per-thread calls).  They add what a bandwidth-bound caller on MI355X needs:
 *
 *  - wavefront- and workgroup-cooperative put/get: every lane (or every thread
 *    of the workgroup) calls with identical arguments and the copy is split
 *    across them with 16-byte accesses;
 *  - the calling workgroup's PE in workgroup-partitioned (local) mode;
 *  - waits with the OpenSHMEM comparison set that the reference's
 *    nvshmem_int_wait_until ignores (reference nvshmem_device.cu:564-575);
 *  - a device barrier across every workgroup of a persistent local-mode kernel.
 */
#ifndef NVSHMEM_AMD_NVSHMEMX_DEVICE_H
#define NVSHMEM_AMD_NVSHMEMX_DEVICE_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include "nvshmem_device.h"

#define NVSHMEMX_CMP_EQ 0
#define NVSHMEMX_CMP_NE 1
#define NVSHMEMX_CMP_GT 2
#define NVSHMEMX_CMP_GE 3
#define NVSHMEMX_CMP_LT 4
#define NVSHMEMX_CMP_LE 5

#ifdef __cplusplus
extern "C" {
#endif

/* PE of the calling workgroup (process mode: the process PE; local mode:
 * blockIdx.x / workgroups-per-PE) and the PE count. */
__device__ int nvshmemx_my_pe_d(void);
__device__ int nvshmemx_n_pes_d(void);

/* Device-side nvshmem_ptr: address of symmetric `addr` on `pe`. */
__device__ void *nvshmemx_ptr_d(const void *addr, int pe);

/* Wavefront-cooperative copies: all active lanes of the wave call together. */
__device__ void nvshmemx_putmem_wave(void *target, const void *source, size_t nbytes, int pe);
__device__ void nvshmemx_getmem_wave(void *target, const void *source, size_t nbytes, int pe);
__device__ void nvshmemx_int_put_wave(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_put_wave(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_put_wave(double *target, const double *source, int len, int pe);
__device__ void nvshmemx_int_get_wave(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_get_wave(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_get_wave(double *target, const double *source, int len, int pe);

/* Non-blocking put (per thread, like nvshmem_T_put): posts the put to the
 * calling PE's device queue and returns; nvshmemx_nbi_flush_on_stream (host)
 * moves every posted put in one persistent launch, stream-ordered after the
 * kernels that posted.  A put is complete when that flush is; a post that
 * finds its PE's queue full (nvshmemx_nbi_capacity() posts since the last
 * flush) is copied at once.  Destinations of puts posted between two flushes
 * must not overlap.  Posts still waiting when the library stops are dropped
 * (a new start begins with empty queues).  No counterpart in the reference. */
__device__ void nvshmemx_putmem_nbi(void *target, const void *source, size_t nbytes, int pe);
/* the same for `len` elements (nothing for len <= 0) */
__device__ void nvshmemx_int_put_nbi(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_put_nbi(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_put_nbi(double *target, const double *source, int len, int pe);
/* Non-blocking get: posts the copy of `nbytes` of symmetric `source` on `pe`
 * into the caller's local `target` to the same queue; complete, like the
 * posted puts, when the flush is (copied at once if the queue is full).
 * Targets of gets and puts posted between two flushes must not overlap. */
__device__ void nvshmemx_getmem_nbi(void *target, const void *source, size_t nbytes, int pe);
__device__ void nvshmemx_int_get_nbi(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_get_nbi(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_get_nbi(double *target, const double *source, int len, int pe);

/* Workgroup-cooperative copies: every thread of the workgroup calls; the
 * call ends with each thread's stores released at agent scope and a
 * workgroup barrier, so any thread may signal completion afterwards. */
__device__ void nvshmemx_putmem_block(void *target, const void *source, size_t nbytes, int pe);
__device__ void nvshmemx_getmem_block(void *target, const void *source, size_t nbytes, int pe);
__device__ void nvshmemx_int_put_block(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_put_block(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_put_block(double *target, const double *source, int len, int pe);
__device__ void nvshmemx_int_get_block(int *target, const int *source, int len, int pe);
__device__ void nvshmemx_float_get_block(float *target, const float *source, int len, int pe);
__device__ void nvshmemx_double_get_block(double *target, const double *source, int len, int pe);

/* Strided and 2-D transfers (the reference API is contiguous only; halo faces
 * of a 3-D block are strided).  iput: target[i * tst] = source[i * sst] on
 * `pe` for i < nelems (strides in elements, >= 1), OpenSHMEM iput semantics;
 * iget: target[i * tst] = source[i * sst] read from `pe`.  putmem_2d:
 * `height` rows of `width` bytes, row r at target + r * tpitch on `pe` from
 * source + r * spitch.  _wave: all active lanes of a wavefront call
 * together; _block: every thread of the workgroup calls, and the block
 * variants end with a release and a workgroup barrier like the contiguous
 * ones. */
__device__ void nvshmemx_int_iput_wave(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                       int pe);
__device__ void nvshmemx_float_iput_wave(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst,
                                         size_t nelems, int pe);
__device__ void nvshmemx_double_iput_wave(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst,
                                          size_t nelems, int pe);
__device__ void nvshmemx_int_iget_wave(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                       int pe);
__device__ void nvshmemx_float_iget_wave(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst,
                                         size_t nelems, int pe);
__device__ void nvshmemx_double_iget_wave(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst,
                                          size_t nelems, int pe);
__device__ void nvshmemx_int_iput_block(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                        int pe);
__device__ void nvshmemx_float_iput_block(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst,
                                          size_t nelems, int pe);
__device__ void nvshmemx_double_iput_block(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst,
                                           size_t nelems, int pe);
__device__ void nvshmemx_int_iget_block(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                        int pe);
__device__ void nvshmemx_float_iget_block(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst,
                                          size_t nelems, int pe);
__device__ void nvshmemx_double_iget_block(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst,
                                           size_t nelems, int pe);
__device__ void nvshmemx_putmem_2d_block(void *target, size_t tpitch, const void *source, size_t spitch, size_t width,
                                         size_t height, int pe);
__device__ void nvshmemx_getmem_2d_block(void *target, size_t tpitch, const void *source, size_t spitch, size_t width,
                                         size_t height, int pe);

/* Block until (*ivar cmp value) holds; returns the value observed. */
__device__ int nvshmemx_int_wait_until_cmp(int *ivar, int cmp, int value);

/* Barrier across every workgroup of the running grid in local mode.  All
 * workgroups of the grid must be co-resident (launch with
 * nvshmemx_collective_launch) and every thread of each workgroup calls it. */
__device__ void nvshmemx_barrier_all_wg(void);

#ifdef __cplusplus
}
#endif

#endif /* NVSHMEM_AMD_NVSHMEMX_DEVICE_H */
