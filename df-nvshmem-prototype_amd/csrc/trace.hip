/*
 * trace.hip -- tracing: the block-collective nvshmem_trace_begin/end device
 * calls (reference nvshmem_device.cu:331-421) and the host side that drains the
 * per-PE ring and writes the reference's CSV (nvshmem.cu:207-458, 610-770).
 *
 * Host side: the ring and the host's consumed count are in pinned,
 * host-coherent memory.  A drain thread (the role of the reference's
 * log_backup thread, nvshmem.cu:206-330) polls the ring while kernels run,
 * copies every published record out in claim order, hands its slot back by
 * advancing the consumed count and appends the records to the binary spill
 * file nvshmem_tracefile_<pid>; a kernel that emits more records than the
 * ring holds therefore loses none as long as the thread keeps up, and waits
 * for it (up to NVSHMEM_LOG_FULL_WAIT_MS) when it does not.  At every host
 * synchronisation point (nvshmem_barrier_all, nvstop_pes,
 * nvshmemx_trace_flush) every record claimed so far is drained.  nvstop_pes
 * converts the spill file into
 * nvshmem_processed_tracefile_<pid of PE 0>_<pe> with the reference's header
 * and row format (nvshmem.cu:377, 390, 411), timestamps converted from
 * wall_clock64 ticks to nanoseconds.
 *
 * Environment (names from the reference, nvshmem.cu:611-638):
 * NVSHMEM_ENABLE_TRACING (default 1), NVSHMEM_LOG_BUFFER_SIZE (bytes, default
 * 2 MiB) x NVSHMEM_LOG_BUFFER_COUNT (default 32) = ring bytes,
 * NVSHMEM_LOG_ACTIVE_GRIDS (128) and NVSHMEM_LOG_ACTIVE_BLOCKS (256), both
 * powers of two.  NVSHMEM_TRACE_DIR selects the output directory,
 * NVSHMEM_LOG_LIVE_DRAIN=0 turns the drain thread off (the ring then drains
 * only at synchronisation points and a full ring drops records at once) and
 * NVSHMEM_LOG_FULL_WAIT_MS (default 1000) bounds how long a claim waits for
 * room in a full ring before it drops its records and raises the overflow
 * flag.
 */
#include "nvshmem_device.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"
#include "nvshmemi_trace.h"
#include "trace_ring.h"

#ifdef NVSHMEM_TRACING
#include <atomic>
#include <chrono>
#include <mutex>
#include <string.h>
#include <thread>
#include <unistd.h>
#endif

/* ------------------------------------------------------------- device -- */

extern "C" __device__ void nvshmem_trace_begin(void) {
#ifdef NVSHMEM_TRACING
    using namespace nvshmemi;
    if (!nvshmemi_dstate.trace_enabled) return;
    __syncthreads();
    if (flat_tid() == 0) {
        nvshmemi_trace_dev_t *t = nvshmemi_dstate.trace;
        const uint64_t gid = grid_id();
        unsigned long long expect = 0;
        if (!__hip_atomic_compare_exchange_strong(&t->active[trace_slot(t, gid)], &expect,
                                                  (unsigned long long)gid + 1, __ATOMIC_RELAXED,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            __hip_atomic_store(&t->flags[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#endif
}

extern "C" __device__ void nvshmem_trace_end(void) {
#ifdef NVSHMEM_TRACING
    using namespace nvshmemi;
    if (!nvshmemi_dstate.trace_enabled) return;
    __syncthreads();
    if (flat_tid() == 0) {
        nvshmemi_trace_dev_t *t = nvshmemi_dstate.trace;
        const uint64_t gid = grid_id();
        unsigned long long expect = (unsigned long long)gid + 1;
        __hip_atomic_compare_exchange_strong(&t->active[trace_slot(t, gid)], &expect, 0ull, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#endif
}

#ifdef NVSHMEM_TRACING
namespace nvshmemi {
__device__ int trace_claim(nvshmemi_trace_dev_t *t, unsigned int n, unsigned int *base) {
    if (n > t->capacity) {
        __hip_atomic_store(&t->flags[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    unsigned int h = __hip_atomic_load(t->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t t0 = 0;
    for (;;) {
        unsigned int tl = __hip_atomic_load(t->tail_seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h + n - tl > t->capacity) {
            tl = __hip_atomic_load(t->host_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            /* a stale value stored here only sends a later claim to the host again */
            __hip_atomic_store(t->tail_seen, tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (h + n - tl > t->capacity) {
                const uint64_t now = wall_clock64();
                if (t0 == 0) t0 = now;
                if (now - t0 >= t->full_wait_ticks) {
                    __hip_atomic_store(&t->flags[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return 0;
                }
                __builtin_amdgcn_s_sleep(64);
                h = __hip_atomic_load(t->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                continue;
            }
        }
        unsigned int expect = h;
        if (__hip_atomic_compare_exchange_strong(t->head, &expect, h + n, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            *base = h;
            return 1;
        }
        h = expect;
    }
}

} // namespace nvshmemi
#endif

/* --------------------------------------------------------------- host -- */

namespace nvshmemi {

#ifdef NVSHMEM_TRACING

static const char *const kOpNames[NVSHMEMI_NUM_OPS] = {"put_int", "put_float", "put_double", "get_int",
                                                       "get_float", "get_double", "wait_int", "wait_until_int"};

namespace {
struct TraceHost {
    nvshmemi_trace_dev_t desc = {};     /* host copy of the device descriptor (device addresses) */
    nvshmemi_trace_dev_t *d_desc = nullptr;
    nvshmemi_trace_record_t *ring = nullptr; /* host address of the ring */
    unsigned int *host_tail = nullptr;       /* host address of the consumed count */
    unsigned int consumed = 0;               /* == *host_tail, owned by whoever holds mu */
    std::string dir = ".";
    std::string spill_path;
    FILE *spill = nullptr;
    unsigned long long records = 0, seq_errors = 0;
    std::vector<nvshmemi_trace_record_t> batch;
    std::mutex mu;
    std::thread drainer;
    std::atomic<bool> stop{false};
};
TraceHost *th = nullptr;

/* Copy out the published records that follow the consumed count, in claim
 * order, hand their slots back and append them to the spill file.  Stops at
 * the first slot whose sequence number is not yet the next one.  Caller
 * holds t->mu.  Returns the number of records taken. */
unsigned int drain_published(TraceHost *t) {
    const unsigned int cap = t->desc.capacity;
    const unsigned int n = ring_published(t->ring, cap, t->consumed);
    if (n == 0) return 0;
    t->batch.resize(n);
    ring_copy(t->batch.data(), t->ring, cap, t->consumed, n);
    t->consumed += n;
    t->records += n;
    __atomic_store_n(t->host_tail, t->consumed, __ATOMIC_RELEASE); /* the slots are free again */
    if (fwrite(t->batch.data(), sizeof(nvshmemi_trace_record_t), n, t->spill) != n)
        warn("short write to trace spill file %s", t->spill_path.c_str());
    return n;
}

void drain_loop(TraceHost *t) {
    while (!t->stop.load(std::memory_order_relaxed)) {
        unsigned int n;
        {
            std::lock_guard<std::mutex> lk(t->mu);
            n = drain_published(t);
        }
        if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

int pow2_env(const char *name, int dflt) {
    const char *v = getenv(name);
    int x = v ? atoi(v) : dflt;
    if (x < 1 || (x & (x - 1)) != 0) NVSHMEMI_FATAL("%s must be a power of two, got %d", name, x);
    return x;
}
} // namespace

nvshmemi_trace_dev_t *trace_init() {
    const char *en = getenv("NVSHMEM_ENABLE_TRACING");
    if (en && atoi(en) == 0) return nullptr;
    th = new TraceHost();
    const size_t buf = parse_size(getenv("NVSHMEM_LOG_BUFFER_SIZE"), (size_t)2 << 20);
    const char *cs = getenv("NVSHMEM_LOG_BUFFER_COUNT");
    const size_t count = cs ? (size_t)atoi(cs) : 32;
    size_t recs = buf * count / sizeof(nvshmemi_trace_record_t);
    if (recs < 1) recs = 1;
    unsigned int cap = 1;
    while ((size_t)cap * 2 <= recs && cap < (1u << 30)) cap <<= 1;
    const int grids = pow2_env("NVSHMEM_LOG_ACTIVE_GRIDS", 128);
    const int blocks = pow2_env("NVSHMEM_LOG_ACTIVE_BLOCKS", 256);
    const char *lv = getenv("NVSHMEM_LOG_LIVE_DRAIN");
    const bool live = !lv || atoi(lv) != 0;
    const size_t wait_ms = parse_size(getenv("NVSHMEM_LOG_FULL_WAIT_MS"), 1000);
    nvshmemi_trace_dev_t &d = th->desc;
    d.capacity = cap;
    d.grid_mask = (unsigned)grids - 1;
    d.block_mask = (unsigned)blocks - 1;
    d.block_shift = 0;
    while ((1 << d.block_shift) < blocks) ++d.block_shift;
    /* nobody drains a full ring while a kernel runs without the thread */
    d.full_wait_ticks = live ? (unsigned long long)wait_ms * (unsigned long long)g.wallclock_khz : 0ull;
    const unsigned int coherent = hipHostMallocMapped | hipHostMallocCoherent;
    NVSHMEMI_HIP(hipHostMalloc((void **)&th->ring, (size_t)cap * sizeof(nvshmemi_trace_record_t), coherent));
    NVSHMEMI_HIP(hipHostMalloc((void **)&th->host_tail, 64, coherent));
    memset(th->ring, 0, (size_t)cap * sizeof(nvshmemi_trace_record_t));
    memset(th->host_tail, 0, 64);
    NVSHMEMI_HIP(hipHostGetDevicePointer((void **)&d.ring, th->ring, 0));
    NVSHMEMI_HIP(hipHostGetDevicePointer((void **)&d.host_tail, th->host_tail, 0));
    NVSHMEMI_HIP(hipMalloc((void **)&d.head, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.tail_seen, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.flags, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.active, (size_t)grids * blocks * sizeof(unsigned long long)));
    NVSHMEMI_HIP(hipMemset(d.head, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.tail_seen, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.flags, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.active, 0, (size_t)grids * blocks * sizeof(unsigned long long)));
    NVSHMEMI_HIP(hipMalloc((void **)&th->d_desc, sizeof(d)));
    NVSHMEMI_HIP(hipMemcpy(th->d_desc, &d, sizeof(d), hipMemcpyHostToDevice));
    if (const char *dir = getenv("NVSHMEM_TRACE_DIR")) th->dir = dir;
    th->spill_path = th->dir + "/nvshmem_tracefile_" + std::to_string((long)getpid());
    th->spill = fopen(th->spill_path.c_str(), "w+b");
    if (!th->spill) NVSHMEMI_FATAL("cannot create trace spill file %s", th->spill_path.c_str());
    if (live) th->drainer = std::thread(drain_loop, th);
    return th->d_desc;
}

/* Every record claimed so far is drained: the kernels that claimed them have
 * finished, so each is published; a slot that is not counts as a sequence
 * error and is skipped. */
void trace_flush() {
    if (!th) return;
    NVSHMEMI_HIP(hipDeviceSynchronize());
    unsigned int head = 0;
    NVSHMEMI_HIP(hipMemcpy(&head, th->desc.head, sizeof(head), hipMemcpyDeviceToHost));
    std::lock_guard<std::mutex> lk(th->mu);
    while (th->consumed != head) {
        if (drain_published(th) == 0) {
            ++th->seq_errors;
            ++th->consumed;
            __atomic_store_n(th->host_tail, th->consumed, __ATOMIC_RELEASE);
        }
    }
    fflush(th->spill);
}

void trace_finalize() {
    if (!th) return;
    trace_flush();
    th->stop.store(true);
    if (th->drainer.joinable()) th->drainer.join();
    unsigned int flags[2] = {0, 0};
    NVSHMEMI_HIP(hipMemcpy(flags, th->desc.flags, sizeof(flags), hipMemcpyDeviceToHost));
    fprintf(stdout, "[%d] LOG STATUS: overflow: %u conflict: %u records: %llu sequence errors: %llu\n", g.mype,
            flags[0], flags[1], th->records, th->seq_errors);
    if (flags[0])
        fprintf(stderr, "[%d] trace ring overflowed; raise NVSHMEM_LOG_BUFFER_SIZE / NVSHMEM_LOG_BUFFER_COUNT\n",
                g.mype);
    if (flags[1])
        fprintf(stderr, "[%d] two blocks mapped to one trace slot; raise NVSHMEM_LOG_ACTIVE_BLOCKS / "
                        "NVSHMEM_LOG_ACTIVE_GRIDS\n", g.mype);
    fflush(stdout);

    long root_pid = (long)getpid();
    g.boot->bcast(&root_pid, sizeof(root_pid), 0);
    const std::string out = th->dir + "/nvshmem_processed_tracefile_" + std::to_string(root_pid) + "_" +
                            std::to_string(g.mype);
    FILE *w = fopen(out.c_str(), "w");
    if (!w) NVSHMEMI_FATAL("cannot create processed trace file %s", out.c_str());
    fprintf(w, "# timestamp, gridid, blockid, tid, type, local_address, remote_address, size, remote_pe, timespent\n");
    rewind(th->spill);
    const double ns_per_tick = 1.0e6 / (double)g.wallclock_khz;
    nvshmemi_trace_record_t r;
    while (fread(&r, sizeof(r), 1, th->spill) == 1) {
        const char *op = (r.type >= 0 && r.type < NVSHMEMI_NUM_OPS) ? kOpNames[r.type] : "unknown";
        fprintf(w, "%lld, %llu, %d, %d, %s, %p, %p, %d, %d, %lld\n", (long long)(r.tstamp * ns_per_tick),
                (unsigned long long)r.gridid, r.blockid, r.tid, op, (void *)(uintptr_t)r.local_addr,
                (void *)(uintptr_t)r.remote_addr, r.size, r.remote_pe, (long long)(r.tspent * ns_per_tick));
    }
    fclose(w);
    fclose(th->spill);
    remove(th->spill_path.c_str());
    fprintf(stdout, "[%d] LOGGED TO FILE : %s\n", g.mype, out.c_str());
    fflush(stdout);
    (void)hipHostFree(th->ring);
    (void)hipHostFree(th->host_tail);
    (void)hipFree(th->desc.head);
    (void)hipFree(th->desc.tail_seen);
    (void)hipFree(th->desc.flags);
    (void)hipFree(th->desc.active);
    (void)hipFree(th->d_desc);
    delete th;
    th = nullptr;
}

#else

nvshmemi_trace_dev_t *trace_init() { return nullptr; }
void trace_flush() {}
void trace_finalize() {}

#endif

} // namespace nvshmemi

/* Drain every record claimed so far (a host synchronisation point); no-op
 * without tracing. */
extern "C" void nvshmemx_trace_flush(void) { nvshmemi::trace_flush(); }

/* 1 when the library was built with tracing and NVSHMEM_ENABLE_TRACING is on. */
extern "C" int nvshmemx_trace_enabled(void) { return nvshmemi::g.trace_dev != nullptr ? 1 : 0; }
