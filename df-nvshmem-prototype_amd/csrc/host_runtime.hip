/*
 * host_runtime.hip -- the host API (reference nvshmem.cu:193-201, 472-883).
 *
 * Start-up, in order (reference nvstart_pes, nvshmem.cu:472-773, with the
 * defects listed below fixed rather than copied):
 *   1. pick the bootstrap (MPI if initialised, else shm, else single process);
 *   2. decide the PE layout: one PE per process (peers' heaps IPC-mapped), or
 *      NVSHMEM_LOCAL_PES PEs inside this process (one allocation, workgroup-
 *      partitioned PEs in kernels);
 *   3. check npes against the table size BEFORE touching the table (the
 *      reference checks after writing it, nvshmem.cu:569 vs 593);
 *   4. read NVSHMEM_HEAP_SIZE BEFORE allocating (the reference allocates
 *      first, so the variable never sized the heap, nvshmem.cu:504 vs 532);
 *   5. allocate the heap (coarse-grained hipMalloc by default, fine-grained
 *      with NVSHMEM_HEAP_KIND=fine), zero the library's internal region,
 *      exchange IPC handles, open peers' heaps;
 *   6. upload the device state in one hipMemcpyToSymbol.
 * nvstop_pes meets every PE BEFORE unmapping peers' heaps (the reference closes
 * the mappings first, nvshmem.cu:825-836).
 */
#include <math.h>
#include <stdarg.h>
#include <string.h>
#include <unistd.h>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_host.h"

extern __constant__ nvshmemi_device_state_t nvshmemi_dstate;

namespace nvshmemi {

HostState g;

void fatal(const char *file, int line, const char *fmt, ...) {
    fprintf(stderr, "[%s:%d] [%d] ", file, line, g.mype);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    fflush(stderr);
    exit(-1);
}

void warn(const char *fmt, ...) {
    fprintf(stderr, "[nvshmem] [%d] ", g.mype);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    fflush(stderr);
}

void require_init(const char *fn) {
    if (!g.initialized) NVSHMEMI_FATAL("%s called before nvstart_pes", fn);
}

size_t parse_size(const char *s, size_t dflt) {
    if (!s || !*s) return dflt;
    char *end = nullptr;
    double v = strtod(s, &end);
    if (end == s || v < 0) NVSHMEMI_FATAL("cannot parse size '%s'", s);
    switch (*end) {
    case 'k': case 'K': v *= 1024.0; break;
    case 'm': case 'M': v *= 1024.0 * 1024.0; break;
    case 'g': case 'G': v *= 1024.0 * 1024.0 * 1024.0; break;
    case '\0': break;
    default: NVSHMEMI_FATAL("cannot parse size '%s'", s);
    }
    return (size_t)v;
}

static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
static size_t next_pow2(size_t x) {
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
static int ilog2(size_t x) {
    int r = 0;
    while ((size_t)1 << (r + 1) <= x) ++r;
    return r;
}

void upload_device_state() {
    nvshmemi_device_state_t &d = g.dstate;
    memset(&d, 0, sizeof(d));
    for (int i = 0; i < g.npes; ++i) d.heap_base[i] = g.heap_base[i];
    d.npes = g.npes;
    d.mode = g.mode;
    if (g.mode == NVSHMEMI_MODE_LOCAL) {
        d.anchor = g.heap_base[0];
        d.offset_mask = g.heap_stride - 1;
        d.mype = 0;
        d.wg_shift = ilog2((size_t)g.wgs_per_pe);
    } else {
        d.anchor = g.heap_alloc;
        d.offset_mask = ~(uint64_t)0;
        d.mype = g.mype;
        d.wg_shift = 31;
    }
    d.wait_timeout_ticks = g.wait_timeout_ms * (uint64_t)g.wallclock_khz;
    d.errors = g.d_errors;
    d.trace = g.trace_dev;
    d.trace_enabled = g.trace_dev != nullptr;
    d.heap_bytes = g.heap_bytes;
    /* local mode: every PE's slot of the one allocation (the offset mask alone
     * would let any address alias into some PE's heap); process mode: this
     * PE's heap */
    d.anchor_span = g.mode == NVSHMEMI_MODE_LOCAL ? g.heap_stride * (uint64_t)g.npes : g.heap_bytes;
    NVSHMEMI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(nvshmemi_dstate), &d, sizeof(d)));
}

static void alloc_heap(char **p, size_t bytes) {
    if (g.fine_grained)
        NVSHMEMI_HIP(hipExtMallocWithFlags((void **)p, bytes, hipDeviceMallocFinegrained));
    else
        NVSHMEMI_HIP(hipMalloc((void **)p, bytes));
}

struct PeInfo {
    hipIpcMemHandle_t heap;
    char bus_id[32];
    uint64_t heap_bytes;
    int32_t pid;
    int32_t fine;
};

static void common_setup() {
    NVSHMEMI_HIP(hipGetDevice(&g.device));
    NVSHMEMI_HIP(hipFree(nullptr)); /* make sure the runtime is up */
    g.heap_size = parse_size(getenv("NVSHMEM_HEAP_SIZE"), (size_t)1 << 30);
    const char *kind = getenv("NVSHMEM_HEAP_KIND");
    g.fine_grained = kind && strcmp(kind, "fine") == 0;
    if (kind && !g.fine_grained && strcmp(kind, "coarse") != 0)
        NVSHMEMI_FATAL("NVSHMEM_HEAP_KIND must be 'coarse' or 'fine', got '%s'", kind);
    g.wait_timeout_ms = parse_size(getenv("NVSHMEM_WAIT_TIMEOUT_MS"), 10000);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g.device) == hipSuccess && khz > 0)
        g.wallclock_khz = khz;
    g.heap_bytes = round_up((size_t)NVSHMEMI_INTERNAL_BYTES + g.heap_size, (size_t)2 << 20);
    NVSHMEMI_HIP(hipMalloc((void **)&g.d_errors, sizeof(int)));
    NVSHMEMI_HIP(hipMemset(g.d_errors, 0, sizeof(int)));
    g.stream = nullptr;
    g.tail = 0;
    g.regions.clear();
}

static void start_process_mode() {
    g.mode = NVSHMEMI_MODE_PROCESS;
    g.mype = g.boot->rank();
    g.npes = g.boot->size();
    if (g.npes > NVSHMEMI_MAX_PES)
        NVSHMEMI_FATAL("%d PEs requested, at most %d are supported", g.npes, NVSHMEMI_MAX_PES);
    common_setup();
    if (g.npes > 1 && g.heap_bytes >= ((size_t)2 << 30))
        NVSHMEMI_FATAL("a heap shared between processes must stay below 2 GiB per PE (NVSHMEM_HEAP_SIZE=%zu)",
                       g.heap_size);
    alloc_heap(&g.heap_alloc, g.heap_bytes);
    NVSHMEMI_HIP(hipMemset(g.heap_alloc, 0, NVSHMEMI_INTERNAL_BYTES));
    NVSHMEMI_HIP(hipDeviceSynchronize());

    std::vector<PeInfo> all(g.npes);
    PeInfo me;
    memset(&me, 0, sizeof(me));
    if (g.npes > 1) NVSHMEMI_HIP(hipIpcGetMemHandle(&me.heap, g.heap_alloc));
    NVSHMEMI_HIP(hipDeviceGetPCIBusId(me.bus_id, sizeof(me.bus_id), g.device));
    me.heap_bytes = g.heap_bytes;
    me.pid = (int32_t)getpid();
    me.fine = g.fine_grained;
    g.boot->allgather(&me, all.data(), sizeof(PeInfo));

    for (int i = 0; i < g.npes; ++i) {
        if (i == g.mype) {
            g.heap_base[i] = g.heap_alloc;
            continue;
        }
        if (all[i].heap_bytes != g.heap_bytes || all[i].fine != me.fine)
            NVSHMEMI_FATAL("PE %d heap (%llu bytes, %s) differs from PE %d's; every PE must use the same "
                           "NVSHMEM_HEAP_SIZE and NVSHMEM_HEAP_KIND", i, (unsigned long long)all[i].heap_bytes,
                           all[i].fine ? "fine" : "coarse", g.mype);
        if (strcmp(all[i].bus_id, me.bus_id) != 0) {
            /* reference nvshmem.cu:517-530: peers on another device need P2P */
            int peer_dev = -1, can = 0;
            if (hipDeviceGetByPCIBusId(&peer_dev, all[i].bus_id) != hipSuccess ||
                hipDeviceCanAccessPeer(&can, g.device, peer_dev) != hipSuccess || !can)
                NVSHMEMI_FATAL("PE %d is on device %s without peer access from %s", i, all[i].bus_id, me.bus_id);
        }
        void *p = nullptr;
        NVSHMEMI_HIP(hipIpcOpenMemHandle(&p, all[i].heap, hipIpcMemLazyEnablePeerAccess));
        g.heap_base[i] = (char *)p;
    }
}

static void start_local_mode(int npes, int wgs_per_pe) {
    g.mode = NVSHMEMI_MODE_LOCAL;
    g.mype = 0;
    if (npes < 1 || npes > NVSHMEMI_MAX_PES)
        NVSHMEMI_FATAL("%d local PEs requested, between 1 and %d are supported", npes, NVSHMEMI_MAX_PES);
    if (wgs_per_pe < 1 || (wgs_per_pe & (wgs_per_pe - 1)) != 0)
        NVSHMEMI_FATAL("workgroups per PE must be a power of two, got %d", wgs_per_pe);
    g.npes = npes;
    g.wgs_per_pe = wgs_per_pe;
    common_setup();
    g.heap_stride = next_pow2(g.heap_bytes);
    alloc_heap(&g.heap_alloc, g.heap_stride * (size_t)npes);
    for (int i = 0; i < npes; ++i) {
        g.heap_base[i] = g.heap_alloc + (size_t)i * g.heap_stride;
        NVSHMEMI_HIP(hipMemset(g.heap_base[i], 0, NVSHMEMI_INTERNAL_BYTES));
    }
    NVSHMEMI_HIP(hipDeviceSynchronize());
}

static void finish_start() {
    g.trace_dev = trace_init();
    upload_device_state();
    persistent_init();
    NVSHMEMI_HIP(hipDeviceSynchronize());
    g.initialized = true;
    g.boot->barrier();
}

static char *user_base() {
    return (g.mode == NVSHMEMI_MODE_LOCAL ? g.heap_base[0] : g.heap_alloc) + NVSHMEMI_INTERNAL_BYTES;
}

} // namespace nvshmemi

using namespace nvshmemi;

/* ------------------------------------------------------------ host API -- */

extern "C" void nvstart_pes(void) {
    if (g.initialized) {
        warn("nvstart_pes called twice; ignoring the second call");
        return;
    }
    g.boot = bootstrap_create();
    const char *lp = getenv("NVSHMEM_LOCAL_PES");
    if (lp && atoi(lp) > 0 && g.boot->size() == 1) {
        const char *w = getenv("NVSHMEM_WGS_PER_PE");
        start_local_mode(atoi(lp), w ? atoi(w) : 1);
    } else {
        start_process_mode();
    }
    finish_start();
}

extern "C" int nvshmemx_start_local_pes(int npes, int wgs_per_pe) {
    if (g.initialized) {
        warn("already started");
        return -1;
    }
    g.boot = bootstrap_create();
    if (g.boot->size() != 1) NVSHMEMI_FATAL("local PEs need a single-process job, found %d ranks", g.boot->size());
    start_local_mode(npes, wgs_per_pe);
    finish_start();
    return 0;
}

extern "C" int nvshmem_my_pe(void) { return g.mype; }
extern "C" int nvshmem_n_pes(void) { return g.npes; }

extern "C" void *nvshmalloc(size_t size) {
    require_init("nvshmalloc");
    const size_t off = round_up(g.tail, NVSHMEMI_ALLOC_ALIGN);
    if (off > g.heap_size || size > g.heap_size - off) {
        fprintf(stderr,
                "[%d] symmetric heap exhausted: %zu bytes requested at offset %zu of a %zu-byte heap "
                "(raise NVSHMEM_HEAP_SIZE), exiting\n",
                g.mype, size, off, g.heap_size);
        fflush(stderr);
        exit(-1);
    }
    g.tail = off + size;
    g.regions.push_back(Region{off, size});
    return user_base() + off;
}

extern "C" void nvshmcleanup(void) {
    require_init("nvshmcleanup");
    g.regions.clear();
    g.tail = 0;
}

extern "C" void *nvshmem_ptr(void *local, int pe) {
    require_init("nvshmem_ptr");
    if (pe < 0 || pe >= g.npes) return nullptr;
    const uint64_t off = ((uint64_t)((char *)local - g.dstate.anchor)) & g.dstate.offset_mask;
    return g.heap_base[pe] + off;
}

namespace nvshmemi {
/* A device wait or offload barrier that outlived NVSHMEM_WAIT_TIMEOUT_MS gave
 * up with its condition still false, so whatever ran after it may have used
 * stale data.  The reference's barrier prints and carries on
 * (nvshmem_kernels.cu:65-68); here the host's next synchronising call ends the
 * process with the reference's error convention instead.  Expects a
 * synchronised device. */
static void fail_on_recorded_timeout(const char *fn) {
    int err = 0;
    NVSHMEMI_HIP(hipMemcpy(&err, g.d_errors, sizeof(int), hipMemcpyDeviceToHost));
    if (err & (NVSHMEMI_ERR_WAIT_TIMEOUT | NVSHMEMI_ERR_BARRIER_TIMEOUT))
        NVSHMEMI_FATAL("%s: a device %s outlived its %llu ms budget (errors 0x%x); results after it cannot be "
                       "trusted", fn, (err & NVSHMEMI_ERR_BARRIER_TIMEOUT) ? "barrier" : "wait",
                       (unsigned long long)g.wait_timeout_ms, err);
}
} // namespace nvshmemi

extern "C" void nvshmem_barrier_all(void) {
    require_init("nvshmem_barrier_all");
    NVSHMEMI_HIP(hipDeviceSynchronize());
    fail_on_recorded_timeout("nvshmem_barrier_all");
    trace_flush();
    g.boot->barrier();
}

extern "C" void nvshmem_barrier_all_offload(void) {
    require_init("nvshmem_barrier_all_offload");
    if (g.mode == NVSHMEMI_MODE_PROCESS && g.npes > 1) launch_offload_barrier(g.stream);
}

extern "C" void nvstop_pes(void) {
    require_init("nvstop_pes");
    NVSHMEMI_HIP(hipDeviceSynchronize());
    fail_on_recorded_timeout("nvstop_pes");
    trace_flush();
    g.boot->barrier(); /* every PE is done with every heap before any is unmapped */
    int err = 0;
    NVSHMEMI_HIP(hipMemcpy(&err, g.d_errors, sizeof(int), hipMemcpyDeviceToHost));
    if (err) warn("device errors recorded during the run: 0x%x (see nvshmemx.h NVSHMEMX_ERR_*)", err);
    if (g.mode == NVSHMEMI_MODE_PROCESS)
        for (int i = 0; i < g.npes; ++i)
            if (i != g.mype) NVSHMEMI_HIP(hipIpcCloseMemHandle(g.heap_base[i]));
    g.boot->barrier();
    NVSHMEMI_HIP(hipFree(g.heap_alloc));
    trace_finalize();
    NVSHMEMI_HIP(hipFree(g.d_errors));
    g.boot->finalize();
    delete g.boot;
    g = HostState();
}

/* --------------------------------------------------------- extensions -- */

extern "C" int nvshmemx_initialized(void) { return g.initialized ? 1 : 0; }
extern "C" int nvshmemx_mode(void) { return g.mode; }
extern "C" int nvshmemx_wgs_per_pe(void) { return g.wgs_per_pe; }
extern "C" size_t nvshmemx_heap_size(void) { return g.heap_size; }
extern "C" int nvshmemx_heap_kind(void) { return g.fine_grained ? NVSHMEMX_HEAP_FINE : NVSHMEMX_HEAP_COARSE; }
extern "C" const char *nvshmemx_bootstrap_name(void) { return g.boot ? g.boot->name() : "none"; }

extern "C" int nvshmemx_get_last_error(void) {
    require_init("nvshmemx_get_last_error");
    int err = 0;
    NVSHMEMI_HIP(hipDeviceSynchronize());
    NVSHMEMI_HIP(hipMemcpy(&err, g.d_errors, sizeof(int), hipMemcpyDeviceToHost));
    return err;
}

extern "C" void nvshmemx_clear_errors(void) {
    require_init("nvshmemx_clear_errors");
    NVSHMEMI_HIP(hipDeviceSynchronize());
    NVSHMEMI_HIP(hipMemset(g.d_errors, 0, sizeof(int)));
    NVSHMEMI_HIP(hipDeviceSynchronize());
}

extern "C" void nvshmemx_set_wait_timeout_ms(unsigned long long ms) {
    require_init("nvshmemx_set_wait_timeout_ms");
    NVSHMEMI_HIP(hipDeviceSynchronize());
    g.wait_timeout_ms = ms;
    g.dstate.wait_timeout_ticks = ms * (uint64_t)g.wallclock_khz;
    NVSHMEMI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(nvshmemi_dstate), &g.dstate, sizeof(g.dstate)));
    NVSHMEMI_HIP(hipDeviceSynchronize());
}

extern "C" int nvshmemx_wallclock_khz(void) { return g.wallclock_khz; }

extern "C" void nvshmemx_set_stream(hipStream_t s) { g.stream = s; }
extern "C" hipStream_t nvshmemx_get_stream(void) { return g.stream; }

static void check_pe(int pe, const char *fn) {
    if (pe < 0 || pe >= g.npes) NVSHMEMI_FATAL("%s: PE %d out of range [0, %d)", fn, pe, g.npes);
}

namespace nvshmemi {
/* A symmetric range must lie inside one PE's heap: in process mode this
 * process's heap, in local mode any PE's slot of the one allocation.  A
 * pointer from outside (a private hipMalloc, a host pointer) would otherwise
 * translate to an address outside every peer's mapping. */
void check_symmetric(const void *addr, size_t nbytes, const char *fn) {
    const char *a = (const char *)addr;
    const char *lo = g.mode == NVSHMEMI_MODE_LOCAL ? g.heap_base[0] : g.heap_base[g.mype];
    const size_t span = g.mode == NVSHMEMI_MODE_LOCAL ? g.heap_stride * (size_t)g.npes : g.heap_bytes;
    if (a < lo || (size_t)(a - lo) >= span)
        NVSHMEMI_FATAL("%s: %p is not a symmetric address (heap %p + %zu bytes)", fn, addr, (const void *)lo, span);
    /* in local mode a PE's slot is heap_stride bytes, of which the first
     * heap_bytes are its heap: an offset in the padding is not symmetric */
    const size_t off = g.mode == NVSHMEMI_MODE_LOCAL ? (size_t)(a - lo) & (g.heap_stride - 1) : (size_t)(a - lo);
    if (off >= g.heap_bytes || nbytes > g.heap_bytes - off)
        NVSHMEMI_FATAL("%s: %zu bytes at symmetric offset %zu run past the %zu-byte heap", fn, nbytes, off,
                       g.heap_bytes);
}
} // namespace nvshmemi

extern "C" void nvshmemx_putmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t s) {
    require_init("nvshmemx_putmem_on_stream");
    check_pe(pe, "nvshmemx_putmem_on_stream");
    if (nbytes == 0) return;
    check_symmetric(target, nbytes, "nvshmemx_putmem_on_stream");
    launch_copy(nvshmem_ptr(target, pe), source, nbytes, default_copy_config(nbytes), s);
}

extern "C" void nvshmemx_getmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t s) {
    require_init("nvshmemx_getmem_on_stream");
    check_pe(pe, "nvshmemx_getmem_on_stream");
    if (nbytes == 0) return;
    check_symmetric(source, nbytes, "nvshmemx_getmem_on_stream");
    launch_copy(target, nvshmem_ptr((void *)source, pe), nbytes, default_copy_config(nbytes), s);
}

extern "C" void nvshmemx_putmem_config_on_stream(void *target, const void *source, size_t nbytes, int pe,
                                                 const nvshmemx_copy_config_t *cfg, hipStream_t s) {
    require_init("nvshmemx_putmem_config_on_stream");
    check_pe(pe, "nvshmemx_putmem_config_on_stream");
    if (nbytes == 0) return;
    check_symmetric(target, nbytes, "nvshmemx_putmem_config_on_stream");
    CopyConfig c = default_copy_config(nbytes);
    if (cfg) {
        if (cfg->workgroups > 0) c.workgroups = cfg->workgroups;
        if (cfg->block > 0) c.block = cfg->block;
        if (cfg->vec_bytes > 0) c.vec_bytes = cfg->vec_bytes;
        if (cfg->unroll > 0) c.unroll = cfg->unroll;
        if (cfg->nontemporal >= 0) c.nontemporal = cfg->nontemporal;
        if (cfg->mode >= 0) {
            if (cfg->mode == 2 && c.mode != 2 && cfg->workgroups <= 0) c.workgroups = 0; /* one tile per workgroup */
            c.mode = cfg->mode;
        }
        if (cfg->streams > 0) c.streams = cfg->streams;
    }
    launch_copy(nvshmem_ptr(target, pe), source, nbytes, c, s);
}

static void check_2d(size_t pitch_a, size_t pitch_b, size_t width, const char *fn) {
    if (pitch_a < width || pitch_b < width) NVSHMEMI_FATAL("%s: pitches must be at least the row width", fn);
}
static size_t extent_2d(size_t pitch, size_t width, size_t height) { return height ? (height - 1) * pitch + width : 0; }

extern "C" void nvshmemx_putmem_2d_on_stream(void *target, size_t tpitch, const void *source, size_t spitch,
                                             size_t width, size_t height, int pe, hipStream_t s) {
    require_init("nvshmemx_putmem_2d_on_stream");
    check_pe(pe, "nvshmemx_putmem_2d_on_stream");
    check_2d(tpitch, spitch, width, "nvshmemx_putmem_2d_on_stream");
    if (!width || !height) return;
    check_symmetric(target, extent_2d(tpitch, width, height), "nvshmemx_putmem_2d_on_stream");
    launch_copy_2d(nvshmem_ptr(target, pe), tpitch, source, spitch, width, height, s);
}

extern "C" void nvshmemx_getmem_2d_on_stream(void *target, size_t tpitch, const void *source, size_t spitch,
                                             size_t width, size_t height, int pe, hipStream_t s) {
    require_init("nvshmemx_getmem_2d_on_stream");
    check_pe(pe, "nvshmemx_getmem_2d_on_stream");
    check_2d(tpitch, spitch, width, "nvshmemx_getmem_2d_on_stream");
    if (!width || !height) return;
    check_symmetric(source, extent_2d(spitch, width, height), "nvshmemx_getmem_2d_on_stream");
    launch_copy_2d(target, tpitch, nvshmem_ptr((void *)source, pe), spitch, width, height, s);
}

template <typename T>
static void strided_on_stream(bool put, T *target, const T *source, ptrdiff_t tst, ptrdiff_t sst, size_t n, int pe,
                              hipStream_t s, const char *fn) {
    require_init(fn);
    check_pe(pe, fn);
    if (tst < 1 || sst < 1) NVSHMEMI_FATAL("%s: strides must be >= 1 (got %td, %td)", fn, tst, sst);
    if (!n) return;
    const size_t ext = ((n - 1) * (size_t)(put ? tst : sst) + 1) * sizeof(T);
    if (put) {
        check_symmetric(target, ext, fn);
        launch_strided((T *)nvshmem_ptr(target, pe), source, tst, sst, n, s);
    } else {
        check_symmetric(source, ext, fn);
        launch_strided(target, (const T *)nvshmem_ptr((void *)source, pe), tst, sst, n, s);
    }
}

#define NVSHMEMI_DEF_HOST_STRIDED(Name, Type)                                                                      \
    extern "C" void nvshmemx_##Name##_iput_on_stream(Type *target, const Type *source, ptrdiff_t tst, ptrdiff_t sst, \
                                                     size_t n, int pe, hipStream_t s) {                            \
        strided_on_stream<Type>(true, target, source, tst, sst, n, pe, s, "nvshmemx_" #Name "_iput_on_stream");   \
    }                                                                                                              \
    extern "C" void nvshmemx_##Name##_iget_on_stream(Type *target, const Type *source, ptrdiff_t tst, ptrdiff_t sst, \
                                                     size_t n, int pe, hipStream_t s) {                            \
        strided_on_stream<Type>(false, target, source, tst, sst, n, pe, s, "nvshmemx_" #Name "_iget_on_stream");  \
    }
NVSHMEMI_DEF_HOST_STRIDED(int, int)
NVSHMEMI_DEF_HOST_STRIDED(float, float)
NVSHMEMI_DEF_HOST_STRIDED(double, double)

extern "C" void nvshmemx_default_copy_config(size_t nbytes, nvshmemx_copy_config_t *out) {
    if (!out) NVSHMEMI_FATAL("nvshmemx_default_copy_config: null output");
    const CopyConfig c = default_copy_config(nbytes);
    out->workgroups = c.workgroups;
    out->block = c.block;
    out->vec_bytes = c.vec_bytes;
    out->unroll = c.unroll;
    out->nontemporal = c.nontemporal;
    out->mode = c.mode;
    out->streams = c.streams;
}

#define NVSHMEMI_DEF_HOST_TYPED(Name, Type)                                                              \
    extern "C" void nvshmemx_##Name##_put_on_stream(Type *target, const Type *source, size_t len, int pe, \
                                                    hipStream_t s) {                                     \
        nvshmemx_putmem_on_stream(target, source, len * sizeof(Type), pe, s);                            \
    }                                                                                                    \
    extern "C" void nvshmemx_##Name##_get_on_stream(Type *target, const Type *source, size_t len, int pe, \
                                                    hipStream_t s) {                                     \
        nvshmemx_getmem_on_stream(target, source, len * sizeof(Type), pe, s);                            \
    }
NVSHMEMI_DEF_HOST_TYPED(int, int)
NVSHMEMI_DEF_HOST_TYPED(float, float)
NVSHMEMI_DEF_HOST_TYPED(double, double)

extern "C" void nvshmemx_barrier_all_on_stream(hipStream_t s) {
    require_init("nvshmemx_barrier_all_on_stream");
    if (g.mode == NVSHMEMI_MODE_PROCESS && g.npes > 1) launch_offload_barrier(s);
}

extern "C" void nvshmemx_collective_launch(const void *func, dim3 grid, dim3 block, void **args, size_t shmem,
                                           hipStream_t s) {
    require_init("nvshmemx_collective_launch");
    int dev_blocks = 0, per_cu = 0, cus = 0;
    NVSHMEMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g.device));
    NVSHMEMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, (int)(block.x * block.y * block.z), shmem));
    dev_blocks = per_cu * cus;
    const long long want = (long long)grid.x * grid.y * grid.z;
    if (want > dev_blocks)
        NVSHMEMI_FATAL("collective launch of %lld workgroups exceeds the %d that can be resident at once", want,
                       dev_blocks);
    if (want > (long long)NVSHMEMI_MAX_BARRIER_WGS)
        NVSHMEMI_FATAL("collective launch of %lld workgroups exceeds the barrier limit %u", want,
                       NVSHMEMI_MAX_BARRIER_WGS);
    NVSHMEMI_HIP(hipLaunchCooperativeKernel(func, grid, block, args, (unsigned)shmem, s));
}
