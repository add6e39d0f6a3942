/*
 * nvshmem.h -- host API of the MI355X-native device-initiated OpenSHMEM subset.
 *
 * Drop-in replacement for the prototype's host header (reference nvshmem.h:34-55).
 * Every declaration below keeps the reference's name, argument list, return type
 * and C linkage, so existing callers recompile unchanged against this header and
 * link libnvshmem.a (relocatable device code, built with -fgpu-rdc).
 *
 * Errors follow the reference convention (nvshmem.cu:57-91): the call prints
 * "[file:line] [pe] ..." to stderr and exits.  Device-side failures that the
 * reference silently ignored (wait/barrier time-outs) are recorded as sticky
 * error bits, readable through nvshmemx_get_last_error() in nvshmemx.h.
 */
#ifndef NVSHMEM_AMD_NVSHMEM_H
#define NVSHMEM_AMD_NVSHMEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference nvshmem.h:38-42: cast helper kept for source compatibility. */
#ifndef NONVOLATILE_UPDATES
#define VOLATILE_POINTER(x) ((volatile __typeof__(*(x)) *)(x))
#else
#define VOLATILE_POINTER(x) (x)
#endif

/* Initialise the PE environment (reference nvshmem.cu:472-773).  Call after
 * MPI_Init in MPI programs; without MPI the PE layout comes from the
 * environment (NVSHMEM_BOOTSTRAP=shm, NVSHMEM_LOCAL_PES), see nvshmemx.h. */
void nvstart_pes(void);

/* Barrier, unmap peers, free the heap, finish tracing (reference nvshmem.cu:821-856). */
void nvstop_pes(void);

/* Calling PE id and PE count (reference nvshmem.cu:193-201). */
int nvshmem_my_pe(void);
int nvshmem_n_pes(void);

/* Symmetric bump allocation from the heap (reference nvshmem.cu:775-802).
 * Collective by contract: every PE makes the same call sequence, so equal
 * offsets land at equal addresses relative to each PE's heap base.  Returns a
 * 256-byte aligned block. */
void *nvshmalloc(size_t size);

/* Release every symmetric allocation made so far (reference nvshmem.cu:805-819). */
void nvshmcleanup(void);

/* Host barrier: drain this PE's GPU work, then meet every other PE
 * (reference nvshmem.cu:858-867). */
void nvshmem_barrier_all(void);

/* Stream-ordered device barrier; the host returns immediately
 * (reference nvshmem.cu:869-878, kernel nvshmem_kernels.cu:56-102).  With
 * local (workgroup-partitioned) PEs it does nothing: every PE lives in the
 * same kernels, which the stream already orders. */
void nvshmem_barrier_all_offload(void);

/* Address of the symmetric object `local` on PE `pe`, for direct loads and
 * stores from kernels (reference nvshmem.cu:880-883). */
void *nvshmem_ptr(void *local, int pe);

#ifdef __cplusplus
}
#endif

#endif /* NVSHMEM_AMD_NVSHMEM_H */
