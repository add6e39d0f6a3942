/*
 * nvshmemx.h -- host-side extensions of the MI355X-native OpenSHMEM subset.
 *
 * None of these has a counterpart in the reference's host header
 * (reference nvshmem.h:44-52); they are what the MI355X design adds around the
 * same semantics: host-launched grid-wide put/get (the bandwidth path), stream
 * selection, workgroup-partitioned PEs in one process, sticky device errors in
 * place of the reference's print-and-continue time-outs
 * (nvshmem_kernels.cu:65-68), and the batched-put engine with its autotuner
 * and learned cost model.  Plain C ABI: pointers, sizes and ints only.
 */
#ifndef NVSHMEM_AMD_NVSHMEMX_H
#define NVSHMEM_AMD_NVSHMEMX_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "nvshmem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sticky device error bits returned by nvshmemx_get_last_error() */
#define NVSHMEMX_ERR_WAIT_TIMEOUT 0x1    /* a device wait outlived NVSHMEM_WAIT_TIMEOUT_MS */
#define NVSHMEMX_ERR_BARRIER_TIMEOUT 0x2 /* an offload barrier outlived the same budget */
#define NVSHMEMX_ERR_BAD_PE 0x4
#define NVSHMEMX_ERR_TRACE_OVERFLOW 0x8
#define NVSHMEMX_ERR_TRACE_CONFLICT 0x10

#define NVSHMEMX_MODE_PROCESS 0 /* one PE per process, peers' heaps IPC-mapped */
#define NVSHMEMX_MODE_LOCAL 1   /* all PEs in this process, workgroup-partitioned */

/* ---- layout and state ---- */

/* Start `npes` PEs inside this process: one heap allocation with a
 * power-of-two stride per PE; in kernels, workgroup b belongs to PE
 * b / wgs_per_pe (wgs_per_pe a power of two).  Same as nvstart_pes() with
 * NVSHMEM_LOCAL_PES / NVSHMEM_WGS_PER_PE set.  Returns 0. */
int nvshmemx_start_local_pes(int npes, int wgs_per_pe);
int nvshmemx_initialized(void);
int nvshmemx_mode(void);
int nvshmemx_wgs_per_pe(void);
size_t nvshmemx_heap_size(void);
const char *nvshmemx_bootstrap_name(void);
int nvshmemx_wallclock_khz(void);

/* Synchronise the device and return the sticky NVSHMEMX_ERR_* bits. */
int nvshmemx_get_last_error(void);
void nvshmemx_clear_errors(void);
/* Budget of every device wait and barrier before it records a time-out. */
void nvshmemx_set_wait_timeout_ms(unsigned long long ms);

/* Stream used by nvshmem_barrier_all_offload (default: the null stream). */
void nvshmemx_set_stream(hipStream_t stream);
hipStream_t nvshmemx_get_stream(void);

/* ---- host-launched bulk copies (grid-cooperative, 16-byte vectorised) ---- */

/* target is symmetric (resolved on `pe`), source is a local device pointer */
void nvshmemx_putmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t stream);
/* target is a local device pointer, source is symmetric (read on `pe`) */
void nvshmemx_getmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t stream);
void nvshmemx_int_put_on_stream(int *target, const int *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_float_put_on_stream(float *target, const float *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_double_put_on_stream(double *target, const double *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_int_get_on_stream(int *target, const int *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_float_get_on_stream(float *target, const float *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_double_get_on_stream(double *target, const double *source, size_t len, int pe, hipStream_t stream);

/* Launch shape of a host-launched copy.  Zero / negative fields keep the
 * library's choice for that field. */
typedef struct nvshmemx_copy_config {
    int workgroups;  /* grid size */
    int block;       /* threads per workgroup (64..1024, multiple of 64) */
    int vec_bytes;   /* widest access: 4, 8 or 16 */
    int unroll;      /* independent accesses in flight per lane: 1, 2, 4 or 8 */
    int nontemporal; /* bit 0: non-temporal loads, bit 1: non-temporal stores */
    int mode;        /* 0: lanes of the whole grid interleave; 1: one contiguous slice per workgroup */
} nvshmemx_copy_config_t;

void nvshmemx_putmem_config_on_stream(void *target, const void *source, size_t nbytes, int pe,
                                      const nvshmemx_copy_config_t *cfg, hipStream_t stream);

/* Stream-ordered device barrier on an explicit stream. */
void nvshmemx_barrier_all_on_stream(hipStream_t stream);

/* Tracing (library built with NVSHMEM_TRACING): drain the ring now; report
 * whether records are being kept.  Both are no-ops / 0 otherwise. */
void nvshmemx_trace_flush(void);
int nvshmemx_trace_enabled(void);

/* Launch a kernel whose workgroups must all be resident at once (kernels that
 * call nvshmemx_barrier_all_wg); exits with an error when the grid does not
 * fit the device. */
void nvshmemx_collective_launch(const void *func, dim3 grid, dim3 block, void **args, size_t shmem,
                                hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NVSHMEM_AMD_NVSHMEMX_H */
