/*
 * nvshmemi_host.h -- host-side runtime state and internal entry points.
 *
 * The reference keeps its runtime in file-scope globals of nvshmem.cu
 * (nvshmem.cu:119-191); here one struct holds the same facts plus what the
 * MI355X design adds (PE layout mode, heap granularity, wait budget, stream).
 * Like the reference, the host runtime is single-threaded and not thread-safe.
 */
#ifndef NVSHMEMI_HOST_H
#define NVSHMEMI_HOST_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string>
#include <vector>

#include "nvshmemi_state.h"
#include "nvshmemi_trace.h"
#include "nvshmemx.h"

namespace nvshmemi {

/* Process-group services used only at start-up and in host barriers
 * (reference: MPI_Allgather/MPI_Barrier/MPI_Bcast, nvshmem.cu:509,551,863,372). */
struct Bootstrap {
    virtual ~Bootstrap() {}
    virtual const char *name() const = 0;
    virtual int rank() const = 0;
    virtual int size() const = 0;
    virtual void allgather(const void *in, void *out, size_t bytes) = 0;
    virtual void barrier() = 0;
    virtual void bcast(void *buf, size_t bytes, int root) = 0;
    virtual void finalize() {}
};

/* MPI when the program initialised it, else the shared-memory bootstrap when
 * NVSHMEM_BOOTSTRAP=shm, else a single process. */
Bootstrap *bootstrap_create();
bool mpi_available();

struct Region {
    size_t offset;
    size_t size;
};

struct HostState {
    bool initialized = false;
    int mode = NVSHMEMI_MODE_PROCESS;
    int mype = 0;
    int npes = 1;
    int wgs_per_pe = 1;
    int device = 0;
    bool fine_grained = false;
    size_t heap_size = 0;   /* user bytes per PE */
    size_t heap_bytes = 0;  /* bytes per PE including the internal region */
    size_t heap_stride = 0; /* local mode: distance between PE heaps (power of two) */
    char *heap_alloc = nullptr; /* this process's allocation */
    char *heap_base[NVSHMEMI_MAX_PES] = {};
    size_t tail = 0; /* next free user offset */
    std::vector<Region> regions;
    hipStream_t stream = nullptr;
    int *d_errors = nullptr;
    Bootstrap *boot = nullptr;
    int wallclock_khz = 100000;
    uint64_t wait_timeout_ms = 10000;
    nvshmemi_device_state_t dstate = {};
    nvshmemi_trace_dev_t *trace_dev = nullptr; /* device copy of the ring descriptor */
};

extern HostState g;

[[noreturn]] void fatal(const char *file, int line, const char *fmt, ...);
void warn(const char *fmt, ...);

#define NVSHMEMI_FATAL(...) ::nvshmemi::fatal(__FILE__, __LINE__, __VA_ARGS__)
#define NVSHMEMI_HIP(stmt)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (stmt);                                                            \
        if (e_ != hipSuccess) NVSHMEMI_FATAL("hip failed with %s (%s)", hipGetErrorString(e_), #stmt); \
    } while (0)

void require_init(const char *fn);
void upload_device_state();
size_t parse_size(const char *s, size_t dflt);

/* kernels.hip */
void launch_offload_barrier(hipStream_t s);
struct CopyConfig {
    int workgroups; /* 0: derive from size */
    int block;      /* threads per workgroup */
    int vec_bytes;  /* 4, 8 or 16 */
    int unroll;     /* 1, 2, 4 or 8 */
    int nontemporal; /* NT_* bits: 1 loads, 2 stores */
    int mode;        /* 0 grid-interleaved, 1 contiguous slice per workgroup, 2 tiles over interleaved streams */
    int streams;     /* mode 2: address streams the tiles alternate between */
};
CopyConfig default_copy_config(size_t nbytes);
/* host_runtime.hip: fatal unless [addr, addr + nbytes) lies in one PE's heap */
void check_symmetric(const void *addr, size_t nbytes, const char *fn);
void launch_copy(void *dst, const void *src, size_t nbytes, const CopyConfig &cfg, hipStream_t s);
void launch_copy_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                    hipStream_t s);
template <typename T>
void launch_strided(T *dst, const T *src, ptrdiff_t dst_st, ptrdiff_t src_st, size_t n, hipStream_t s);

/* batch.hip: plans over requests whose targets are already resolved to this
 * process's address of the target heap (also used by the autotuner, which
 * rebuilds one plan per candidate over scratch buffers) */
struct resolved_req {
    const char *src;
    char *dst;
    size_t nbytes;
    int set;
};
/* Default launch of a fused batch: 8192 workgroups of 1024 threads with
 * non-temporal stores.  Far more workgroups than fit on the device at once
 * (512 of this size), so the hardware dispatcher hands the per-workgroup
 * piece lists to whichever slot frees first, which evens out the tail that a
 * static assignment alone leaves; cached loads, because the puts of a batch
 * may read overlapping sources.  On the 4096-put step this measured 5.5 TB/s
 * against 5.1 TB/s for 2048 x 256 with non-temporal loads and stores. */
constexpr int kBatchWorkgroups = 8192;
constexpr int kBatchBlock = 1024;
constexpr int kBatchNontemporal = 2; /* NT_STORE */

/* Batch pieces start on multiples of 4 KiB from their request's start.  With
 * 256-byte boundaries, pieces of 8-24 KiB made the 4096-put step 2-4% slower
 * than with 4 KiB boundaries (a piece that straddles DRAM page interleave
 * units spreads its accesses over more pages). */
constexpr size_t kPieceQuantum = 4096;

enum { ASSIGN_LPT = 0, ASSIGN_ROUND_ROBIN = 1, ASSIGN_SCATTERED = 2, ASSIGN_SOURCE_ORDER = 3 };
/* autotune.hip: largest balanced piece of a batch (0: no cap) and the
 * constrained choice over the 54 candidates' costs (times or log-times) */
size_t balance_cap(size_t batch_bytes);
int pick_candidate(const float *cost, size_t nbytes, size_t cap, bool log_cost);
/* threads the device runs at once (compute units x threads per unit) */
size_t resident_threads();
void plan_fill(nvshmemx_batch_plan *p, const std::vector<resolved_req> &reqs, const nvshmemx_batch_choice_t *choices,
               const nvshmemx_batch_choice_t *one_choice, int nsets, int wgs_per_set, int block, int nontemporal,
               int assign);
void plan_launch(nvshmemx_batch_plan *p, hipStream_t s);
nvshmemx_batch_plan *plan_new();

/* persistent.hip: per-device set-up of the persistent batch kernel (called
 * once the PE layout is known, with the device current) */
void persistent_init();

/* trace_host.hip (no-ops unless built with NVSHMEM_TRACING); trace_init
 * returns the device-side ring descriptor, or nullptr when tracing is off */
nvshmemi_trace_dev_t *trace_init();
void trace_flush();
void trace_finalize();

} // namespace nvshmemi

#endif /* NVSHMEMI_HOST_H */
