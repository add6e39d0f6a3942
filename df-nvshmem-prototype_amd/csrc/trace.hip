/*
 * trace.hip -- tracing: the block-collective nvshmem_trace_begin/end device
 * calls (reference nvshmem_device.cu:331-421) and the host side that drains the
 * per-PE ring and writes the reference's CSV (nvshmem.cu:207-458, 610-770).
 *
 * Host side: at every host synchronisation point (nvshmem_barrier_all,
 * nvstop_pes, nvshmemx_trace_flush) the unread part of the ring is copied to
 * the host and handed to a writer thread that appends it to the binary spill
 * file nvshmem_tracefile_<pid>; the device keeps writing while the file I/O
 * runs.  nvstop_pes converts the spill file into
 * nvshmem_processed_tracefile_<pid of PE 0>_<pe> with the reference's header
 * and row format (nvshmem.cu:377, 390, 411), timestamps converted from
 * wall_clock64 ticks to nanoseconds.
 *
 * Environment (names from the reference, nvshmem.cu:611-638):
 * NVSHMEM_ENABLE_TRACING (default 1), NVSHMEM_LOG_BUFFER_SIZE (bytes, default
 * 2 MiB) x NVSHMEM_LOG_BUFFER_COUNT (default 32) = ring bytes,
 * NVSHMEM_LOG_ACTIVE_GRIDS (128) and NVSHMEM_LOG_ACTIVE_BLOCKS (256), both
 * powers of two.  NVSHMEM_TRACE_DIR selects the output directory.
 */
#include "nvshmem_device.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"
#include "nvshmemi_trace.h"

#ifdef NVSHMEM_TRACING
#include <condition_variable>
#include <deque>
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

/* --------------------------------------------------------------- host -- */

namespace nvshmemi {

#ifdef NVSHMEM_TRACING

static const char *const kOpNames[NVSHMEMI_NUM_OPS] = {"put_int", "put_float", "put_double", "get_int",
                                                       "get_float", "get_double", "wait_int", "wait_until_int"};

namespace {
struct TraceHost {
    nvshmemi_trace_dev_t desc = {};     /* host copy of the device descriptor */
    nvshmemi_trace_dev_t *d_desc = nullptr;
    unsigned int tail = 0;
    std::string dir = ".";
    std::string spill_path;
    FILE *spill = nullptr;
    unsigned long long records = 0, seq_errors = 0;
    /* writer thread */
    std::thread writer;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::vector<nvshmemi_trace_record_t>> queue;
    bool stop = false;
};
TraceHost *th = nullptr;

void writer_loop(TraceHost *t) {
    std::unique_lock<std::mutex> lk(t->mu);
    for (;;) {
        t->cv.wait(lk, [t] { return t->stop || !t->queue.empty(); });
        if (t->queue.empty() && t->stop) return;
        std::vector<nvshmemi_trace_record_t> batch = std::move(t->queue.front());
        t->queue.pop_front();
        lk.unlock();
        if (!batch.empty() &&
            fwrite(batch.data(), sizeof(nvshmemi_trace_record_t), batch.size(), t->spill) != batch.size())
            warn("short write to trace spill file %s", t->spill_path.c_str());
        lk.lock();
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
    nvshmemi_trace_dev_t &d = th->desc;
    d.capacity = cap;
    d.grid_mask = (unsigned)grids - 1;
    d.block_mask = (unsigned)blocks - 1;
    d.block_shift = 0;
    while ((1 << d.block_shift) < blocks) ++d.block_shift;
    NVSHMEMI_HIP(hipMalloc((void **)&d.ring, (size_t)cap * sizeof(nvshmemi_trace_record_t)));
    NVSHMEMI_HIP(hipMalloc((void **)&d.head, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.tail, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.flags, 64));
    NVSHMEMI_HIP(hipMalloc((void **)&d.active, (size_t)grids * blocks * sizeof(unsigned long long)));
    NVSHMEMI_HIP(hipMemset(d.ring, 0, (size_t)cap * sizeof(nvshmemi_trace_record_t)));
    NVSHMEMI_HIP(hipMemset(d.head, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.tail, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.flags, 0, 64));
    NVSHMEMI_HIP(hipMemset(d.active, 0, (size_t)grids * blocks * sizeof(unsigned long long)));
    NVSHMEMI_HIP(hipMalloc((void **)&th->d_desc, sizeof(d)));
    NVSHMEMI_HIP(hipMemcpy(th->d_desc, &d, sizeof(d), hipMemcpyHostToDevice));
    if (const char *dir = getenv("NVSHMEM_TRACE_DIR")) th->dir = dir;
    th->spill_path = th->dir + "/nvshmem_tracefile_" + std::to_string((long)getpid());
    th->spill = fopen(th->spill_path.c_str(), "w+b");
    if (!th->spill) NVSHMEMI_FATAL("cannot create trace spill file %s", th->spill_path.c_str());
    th->writer = std::thread(writer_loop, th);
    return th->d_desc;
}

void trace_flush() {
    if (!th) return;
    NVSHMEMI_HIP(hipDeviceSynchronize());
    unsigned int head = 0;
    NVSHMEMI_HIP(hipMemcpy(&head, th->desc.head, sizeof(head), hipMemcpyDeviceToHost));
    const unsigned int n = head - th->tail;
    if (n == 0) return;
    std::vector<nvshmemi_trace_record_t> batch(n);
    const unsigned int cap = th->desc.capacity;
    const unsigned int first = th->tail & (cap - 1);
    const unsigned int n1 = (first + n <= cap) ? n : cap - first;
    NVSHMEMI_HIP(hipMemcpy(batch.data(), th->desc.ring + first, (size_t)n1 * sizeof(batch[0]), hipMemcpyDeviceToHost));
    if (n1 < n)
        NVSHMEMI_HIP(hipMemcpy(batch.data() + n1, th->desc.ring, (size_t)(n - n1) * sizeof(batch[0]),
                               hipMemcpyDeviceToHost));
    for (unsigned int i = 0; i < n; ++i)
        if (batch[i].seq != th->tail + i + 1) ++th->seq_errors;
    th->tail = head;
    th->records += n;
    NVSHMEMI_HIP(hipMemcpy(th->desc.tail, &th->tail, sizeof(th->tail), hipMemcpyHostToDevice));
    {
        std::lock_guard<std::mutex> lk(th->mu);
        th->queue.push_back(std::move(batch));
    }
    th->cv.notify_one();
}

void trace_finalize() {
    if (!th) return;
    trace_flush();
    {
        std::lock_guard<std::mutex> lk(th->mu);
        th->stop = true;
    }
    th->cv.notify_one();
    th->writer.join();
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
    (void)hipFree(th->desc.ring);
    (void)hipFree(th->desc.head);
    (void)hipFree(th->desc.tail);
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

/* Drain the trace ring now (a host synchronisation point); no-op without tracing. */
extern "C" void nvshmemx_trace_flush(void) { nvshmemi::trace_flush(); }

/* 1 when the library was built with tracing and NVSHMEM_ENABLE_TRACING is on. */
extern "C" int nvshmemx_trace_enabled(void) { return nvshmemi::g.trace_dev != nullptr ? 1 : 0; }
