/*
 * nvshmemi_trace.h -- per-PE operation tracing (reference nvshmem_device.cu:49-204,
 * 331-421; host side nvshmem.cu:207-458, 610-770).
 *
 * Built only when NVSHMEM_TRACING is defined (the reference's TRACING=1 /
 * _ENABLE_TRACING_ switch, Makefile:7-12); otherwise every hook below compiles
 * to nothing and the device API pays no branch.  When built, the runtime
 * switch NVSHMEM_ENABLE_TRACING selects whether records are kept.
 *
 * Design: one ring of fixed 64-byte records per PE in pinned, host-coherent
 * memory, so a host thread can read it while kernels run (the role of the
 * reference's log_backup thread, nvshmem.cu:206-330).  The claim counter
 * (head) lives in device memory.  A wavefront claims one slot per recording
 * lane with a single compare-and-swap, which refuses to pass the host's
 * consumed count (tail, in host memory; the device keeps its last reading in
 * device memory and only reads the host again when the ring looks full).
 * When the ring is full the wavefront waits for the host to drain it, up to
 * full_wait_ticks, and then drops its records and raises the overflow flag
 * instead of overwriting unread ones.  A record is stamped with s_memrealtime
 * (wall_clock64, constant rate) and published by a system-scope release store
 * of its sequence number, written last.  As in the reference's block-level
 * mode, only operations issued between nvshmem_trace_begin and
 * nvshmem_trace_end are recorded; the active table slot of a block is
 * (grid id mod G) * B + (block id mod B).
 */
#ifndef NVSHMEMI_TRACE_H
#define NVSHMEMI_TRACE_H

#include <stdint.h>

/* operation codes; names in the CSV follow the reference's table (nvshmem.cu:117) */
#define NVSHMEMI_OP_PUT_INT 0
#define NVSHMEMI_OP_PUT_FLOAT 1
#define NVSHMEMI_OP_PUT_DOUBLE 2
#define NVSHMEMI_OP_GET_INT 3
#define NVSHMEMI_OP_GET_FLOAT 4
#define NVSHMEMI_OP_GET_DOUBLE 5
#define NVSHMEMI_OP_WAIT_INT 6
#define NVSHMEMI_OP_WAIT_UNTIL_INT 7
#define NVSHMEMI_NUM_OPS 8

typedef struct nvshmemi_trace_record {
    int64_t tstamp;       /* wall_clock64 ticks at the call */
    uint64_t gridid;      /* AQL dispatch packet address of the kernel */
    uint64_t local_addr;
    uint64_t remote_addr;
    int64_t tspent;       /* wall_clock64 ticks spent (waits) */
    int32_t blockid;
    int32_t tid;
    int32_t type;
    int32_t size;
    int32_t remote_pe;
    uint32_t seq;         /* claim index + 1, written last */
} nvshmemi_trace_record_t;

typedef struct nvshmemi_trace_dev {
    nvshmemi_trace_record_t *ring; /* capacity records, pinned host memory (device address) */
    unsigned int *head;            /* claim counter (device memory) */
    unsigned int *tail_seen;       /* the device's last reading of *host_tail (device memory) */
    unsigned int *host_tail;       /* records consumed by the host (pinned host memory) */
    unsigned long long *active;    /* [grids * blocks] grid id + 1 of the owning block */
    unsigned int *flags;           /* [0] overflow, [1] slot conflict */
    unsigned long long full_wait_ticks; /* how long a claim waits for the host when the ring is full */
    unsigned int capacity;         /* power of two */
    unsigned int grid_mask;
    unsigned int block_mask;
    unsigned int block_shift;
} nvshmemi_trace_dev_t;

#if defined(__HIPCC__) && defined(NVSHMEM_TRACING)
#include "nvshmemi_device.h"
namespace nvshmemi {

/* Identity of the running grid: the address of its AQL dispatch packet, unique
 * among grids that run at the same time (the role of the reference's %gridid). */
__device__ __forceinline__ uint64_t grid_id() { return (uint64_t)(uintptr_t)__builtin_amdgcn_dispatch_ptr(); }

__device__ __forceinline__ unsigned int trace_slot(const nvshmemi_trace_dev_t *t, uint64_t gid) {
    /* packets are 64-byte aligned: drop the always-zero low bits first */
    return (unsigned int)((((gid >> 6) & t->grid_mask) << t->block_shift) | (flat_block() & t->block_mask));
}

/* Claim n consecutive slots (one lane of the wavefront calls it; defined in
 * trace.hip); 0 when they were dropped. */
__device__ int trace_claim(nvshmemi_trace_dev_t *t, unsigned int n, unsigned int *base);

__device__ __forceinline__ void trace_record(int type, const void *local, const void *remote, int size,
                                             int pe, int64_t tspent) {
    if (!nvshmemi_dstate.trace_enabled) return;
    nvshmemi_trace_dev_t *t = nvshmemi_dstate.trace;
    const uint64_t gid = grid_id();
    const unsigned int slot = trace_slot(t, gid);
    if (__hip_atomic_load(&t->active[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gid + 1) return;
    /* the lanes recording now take consecutive slots, in lane order */
    const unsigned long long live = __ballot(1);
    const int leader = __ffsll((long long)live) - 1;
    const unsigned int rank = (unsigned int)__popcll(live & __lanemask_lt());
    unsigned int base = 0;
    int ok = 0;
    if ((int)__lane_id() == leader) ok = trace_claim(t, (unsigned int)__popcll(live), &base);
    ok = __builtin_amdgcn_readlane(ok, leader);
    if (!ok) return;
    const unsigned int h = (unsigned int)__builtin_amdgcn_readlane((int)base, leader) + rank;
    nvshmemi_trace_record_t *r = &t->ring[h & (t->capacity - 1)];
    r->tstamp = (int64_t)wall_clock64();
    r->gridid = gid;
    r->local_addr = (uint64_t)(uintptr_t)local;
    r->remote_addr = (uint64_t)(uintptr_t)remote;
    r->tspent = tspent;
    r->blockid = (int32_t)flat_block();
    r->tid = (int32_t)flat_tid();
    r->type = type;
    r->size = size;
    r->remote_pe = pe;
    __hip_atomic_store(&r->seq, h + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
} // namespace nvshmemi

#define NVSHMEMI_TRACE(type, local, remote, size, pe, tspent) \
    nvshmemi::trace_record((type), (local), (remote), (size), (pe), (tspent))
#define NVSHMEMI_TRACE_TIMED_BEGIN() const int64_t nvshmemi_t0_ = (int64_t)wall_clock64()
#define NVSHMEMI_TRACE_TIMED_END(type, local, remote, size, pe) \
    nvshmemi::trace_record((type), (local), (remote), (size), (pe), (int64_t)wall_clock64() - nvshmemi_t0_)

#else
#define NVSHMEMI_TRACE(type, local, remote, size, pe, tspent) \
    do {                                                      \
    } while (0)
#define NVSHMEMI_TRACE_TIMED_BEGIN() \
    do {                             \
    } while (0)
#define NVSHMEMI_TRACE_TIMED_END(type, local, remote, size, pe) \
    do {                                                        \
    } while (0)
#endif

#endif /* NVSHMEMI_TRACE_H */
