/*
 * nvshmemi_trace.h -- per-PE operation tracing (reference nvshmem_device.cu:49-204,
 * 331-421; host side nvshmem.cu:207-458, 610-770).
 *
 * Built only when NVSHMEM_TRACING is defined (the reference's TRACING=1 /
 * _ENABLE_TRACING_ switch, Makefile:7-12); otherwise every hook below compiles
 * to nothing and the device API pays no branch.  When built, the runtime
 * switch NVSHMEM_ENABLE_TRACING selects whether records are kept.
 *
 * Design: one ring of fixed 64-byte records per PE in device memory.  A record
 * is claimed with a compare-and-swap on the ring head that refuses to pass the
 * host's consumed count (so a full ring drops records and raises the overflow
 * flag instead of overwriting unread ones), stamped with s_memrealtime
 * (wall_clock64, constant rate) and published by writing its sequence number
 * last.  As in the reference's block-level mode, only operations issued
 * between nvshmem_trace_begin and nvshmem_trace_end are recorded; the active
 * table slot of a block is (grid id mod G) * B + (block id mod B).
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
    nvshmemi_trace_record_t *ring; /* capacity records, device memory */
    unsigned int *head;            /* claim counter (device memory) */
    unsigned int *tail;            /* records consumed by the host (device memory) */
    unsigned long long *active;    /* [grids * blocks] grid id + 1 of the owning block */
    unsigned int *flags;           /* [0] overflow, [1] slot conflict */
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

__device__ __forceinline__ void trace_record(int type, const void *local, const void *remote, int size,
                                             int pe, int64_t tspent) {
    if (!nvshmemi_dstate.trace_enabled) return;
    nvshmemi_trace_dev_t *t = nvshmemi_dstate.trace;
    const uint64_t gid = grid_id();
    const unsigned int slot = trace_slot(t, gid);
    if (__hip_atomic_load(&t->active[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gid + 1) return;
    unsigned int h = __hip_atomic_load(t->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const unsigned int tl = __hip_atomic_load(t->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h - tl >= t->capacity) {
            __hip_atomic_store(&t->flags[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        unsigned int expect = h;
        if (__hip_atomic_compare_exchange_strong(t->head, &expect, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            break;
        h = expect;
    }
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
    __hip_atomic_store(&r->seq, h + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
