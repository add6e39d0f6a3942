/*
 * nvshmemi_state.h -- state shared by the host runtime and the device library.
 *
 * The reference keeps four separate __constant__ symbols plus a host table of
 * at most 16 heap pointers (nvshmem_device.cu:39-47, nvshmem_constants.h:38-50).
 * Here one struct describes the whole PE layout, uploaded once with
 * hipMemcpyToSymbol, and the PE limit is 128.
 *
 * Symmetric translation (reference nvshmem_device.cu:445-446, nvshmem.cu:880-883)
 * is written so that one formula serves both PE layouts:
 *
 *     remote = heap_base[pe] + ((addr - anchor) & offset_mask)
 *
 *  - process mode (one PE per process, peers IPC-mapped): anchor is this
 *    process's heap base and offset_mask is all ones, which is exactly the
 *    reference formula base[pe] + (addr - my_base);
 *  - local mode (all PEs in one process, one allocation with a power-of-two
 *    stride per PE): anchor is PE 0's base and offset_mask is stride-1, so an
 *    address inside ANY PE's heap maps to the same offset on `pe`.
 */
#ifndef NVSHMEMI_STATE_H
#define NVSHMEMI_STATE_H

#include <stddef.h>
#include <stdint.h>

#define NVSHMEMI_MAX_PES 128

/* bytes at the start of every PE heap reserved for the library (barrier
 * words); user allocations start after it */
#define NVSHMEMI_INTERNAL_BYTES (64u * 1024u)
/* offsets inside the internal region, on distinct cache lines.  A barrier
 * word holds the arrival count in its low 16 bits and the generation (the
 * barrier's sense) in its high 16 bits. */
#define NVSHMEMI_OFF_BAR_WORD 0u     /* offload barrier among PEs */
#define NVSHMEMI_OFF_WGBAR_WORD 256u /* barrier among the workgroups of a grid */
#define NVSHMEMI_MAX_BARRIER_WGS 8192u

#define NVSHMEMI_ALLOC_ALIGN 256u

enum nvshmemi_mode { NVSHMEMI_MODE_PROCESS = 0, NVSHMEMI_MODE_LOCAL = 1 };

/* sticky error bits (also exported through nvshmemx.h) */
#define NVSHMEMI_ERR_WAIT_TIMEOUT 0x1
#define NVSHMEMI_ERR_BARRIER_TIMEOUT 0x2
#define NVSHMEMI_ERR_BAD_PE 0x4
#define NVSHMEMI_ERR_TRACE_OVERFLOW 0x8
#define NVSHMEMI_ERR_TRACE_CONFLICT 0x10
#define NVSHMEMI_ERR_BAD_ADDRESS 0x20

struct nvshmemi_trace_dev; /* tracing ring, see nvshmemi_trace.h */

typedef struct nvshmemi_device_state {
    char *heap_base[NVSHMEMI_MAX_PES]; /* every PE's heap, in this process's VA */
    char *anchor;                      /* see translation note above */
    uint64_t offset_mask;
    int mype;     /* process mode: this PE; local mode: 0 */
    int npes;
    int wg_shift; /* local mode: log2(workgroups per PE); process mode: 31 */
    int mode;
    uint64_t wait_timeout_ticks; /* wall_clock64 ticks before a wait gives up */
    int *errors;                 /* this process's sticky error word (device memory) */
    int trace_enabled;
    int pad0;
    struct nvshmemi_trace_dev *trace; /* NULL when tracing is off */
    uint64_t heap_bytes;              /* bytes of each PE's heap (internal region included) */
    uint64_t anchor_span;             /* a symmetric address lies in [anchor, anchor + anchor_span) */
} nvshmemi_device_state_t;

#endif /* NVSHMEMI_STATE_H */
