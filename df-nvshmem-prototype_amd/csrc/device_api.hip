/*
 * device_api.hip -- the device API (reference nvshmem_device.cu:431-584) and its
 * cooperative extensions, compiled as relocatable device code (-fgpu-rdc) into
 * libnvshmem.a / libnvshmem_amd.so.
 *
 * Differences from the reference that keep its contract but not its mechanism:
 *  - p/g use one agent-scope atomic store/load instead of a volatile access;
 *  - put/get copy with the widest access the relative alignment allows
 *    (16 B when target and source agree mod 16), four loads in flight, instead
 *    of one volatile 4/8-byte access per element (nvshmem_device.cu:497-500);
 *  - quiet is an agent-scope fence: all PEs share one agent;
 *  - waits spin, then poll with s_sleep back-off, and give up after a wall-clock budget,
 *    recording NVSHMEMI_ERR_WAIT_TIMEOUT, where the reference spins forever
 *    reading %globaltimer (nvshmem_device.cu:551-553).
 */
#include "nvshmem_device.h"
#include "nvshmemx_device.h"
#include "nvshmemi_device.h"
#include "nvshmemi_trace.h"

/* the single device-state symbol, uploaded by the host runtime at start-up */
__constant__ nvshmemi_device_state_t nvshmemi_dstate;

using namespace nvshmemi;

/* NVSHMEM_IDEALIZE (the reference's IDEALIZE=1 / _ENABLE_IDEALIZE_COMM_,
 * Makefile:14-16): the reference API's p, g, put, get, quiet and waits compile
 * to nothing (g returns 0), to time an application's compute alone.  The
 * nvshmemx_ extensions keep working. */
#ifdef NVSHMEM_IDEALIZE
#define NVSHMEMI_IDEAL(...) return __VA_ARGS__
#else
#define NVSHMEMI_IDEAL(...) \
    do {                    \
    } while (0)
#endif

/* A PE outside [0, npes), or a symmetric range that leaves the heap (or an
 * address outside every PE's heap, such as a private allocation), makes
 * the call do nothing and records NVSHMEMI_ERR_BAD_PE or
 * NVSHMEMI_ERR_BAD_ADDRESS (read by nvshmemx_get_last_error) instead of
 * touching an address outside every PE's mapping (the reference does not
 * check, nvshmem_device.cu:445-446). */
namespace {
__device__ __forceinline__ bool bad_pe(int pe) {
    if ((unsigned)pe < (unsigned)nvshmemi_dstate.npes) return false;
    set_error(NVSHMEMI_ERR_BAD_PE);
    return true;
}
__device__ __forceinline__ bool bad_target(const void *sym, size_t nbytes, int pe) {
    if (bad_pe(pe)) return true;
    const nvshmemi_device_state_t &s = nvshmemi_dstate;
    const uint64_t rel = (uint64_t)((const char *)sym - s.anchor);
    const uint64_t off = rel & s.offset_mask;
    if (rel < s.anchor_span && off < s.heap_bytes && nbytes <= s.heap_bytes - off) return false;
    set_error(NVSHMEMI_ERR_BAD_ADDRESS);
    return true;
}
/* bytes a strided transfer of n elements of `elem` bytes spans */
__device__ __forceinline__ size_t strided_extent(size_t n, ptrdiff_t stride, size_t elem) {
    return n ? ((n - 1) * (size_t)stride + 1) * elem : 0;
}
/* translated address for a valid pe, the untranslated one otherwise (the
 * caller then moves nothing) */
__device__ __forceinline__ char *bad_pe_target(const void *addr, int pe) {
    return (unsigned)pe < (unsigned)nvshmemi_dstate.npes ? translate(addr, pe) : (char *)addr;
}
} // namespace

/* ------------------------------------------------------------ quiet -- */

extern "C" __device__ void nvshmem_quiet(void) {
    /* reference nvshmem_device.cu:431-436 */
    NVSHMEMI_IDEAL();
    fence_acq_rel();
}

/* ------------------------------------------------------------- p / g -- */

#define NVSHMEMI_DEF_P(Name, Type, OpType)                                            \
    extern "C" __device__ void nvshmem_##Name##_p(Type *target, const Type value, int pe) { \
        NVSHMEMI_IDEAL();                                                             \
        if (bad_target(target, sizeof(Type), pe)) return;                             \
        NVSHMEMI_TRACE(OpType, (const void *)0, target, 1, pe, 0);                   \
        store_relaxed((Type *)translate(target, pe), value);                          \
    }
NVSHMEMI_DEF_P(int, int, NVSHMEMI_OP_PUT_INT)
NVSHMEMI_DEF_P(float, float, NVSHMEMI_OP_PUT_FLOAT)
NVSHMEMI_DEF_P(double, double, NVSHMEMI_OP_PUT_DOUBLE)

#define NVSHMEMI_DEF_G(Name, Type, OpType)                                            \
    extern "C" __device__ Type nvshmem_##Name##_g(Type *source, int pe) {               \
        NVSHMEMI_IDEAL((Type)0);                                                      \
        if (bad_target(source, sizeof(Type), pe)) return (Type)0;                     \
        NVSHMEMI_TRACE(OpType, (const void *)0, source, 1, pe, 0);                   \
        return load_relaxed((const Type *)translate(source, pe));                     \
    }
NVSHMEMI_DEF_G(int, int, NVSHMEMI_OP_GET_INT)
NVSHMEMI_DEF_G(float, float, NVSHMEMI_OP_GET_FLOAT)
NVSHMEMI_DEF_G(double, double, NVSHMEMI_OP_GET_DOUBLE)

/* --------------------------------------------------- per-thread bulk -- */

/* A single lane's copy is bound by round trips, not bandwidth.  When target
 * and source agree modulo 16, the 16-byte vectors go in blocks of 16, then 8,
 * then 4 loads in flight, and the last few one by one, so neither long nor
 * short puts wait on more round trips than they need; the unaligned head and
 * the tail (under 16 bytes each) are each one round trip: all their loads are
 * issued before the first store.  Other alignments take the general copy. */
namespace {
/* n < 16 bytes, any alignment: byte accesses */
__device__ __forceinline__ void lane_bytes(char *d, const char *s, size_t n) {
    unsigned char r[15];
#pragma unroll
    for (int i = 0; i < 15; ++i)
        if ((size_t)i < n) r[i] = (unsigned char)s[i];
#pragma unroll
    for (int i = 0; i < 15; ++i)
        if ((size_t)i < n) d[i] = (char)r[i];
}
/* n < 16 bytes that end (`up`) or start (!up) on a 16-byte boundary of both
 * sides: pieces of 1, 2, 4, 8 bytes (ascending) or 8, 4, 2, 1 (descending),
 * each naturally aligned, so a 4-byte element moves as one dword */
template <bool UP>
__device__ __forceinline__ void lane_edge(char *d, const char *s, size_t n) {
    const size_t o1 = UP ? 0 : (n & 14), o2 = UP ? (n & 1) : (n & 12), o4 = UP ? (n & 3) : (n & 8),
                 o8 = UP ? (n & 7) : 0;
    unsigned long long a8 = 0;
    unsigned int a4 = 0;
    unsigned short a2 = 0;
    unsigned char a1 = 0;
    if (n & 8) a8 = *(const unsigned long long *)(s + o8);
    if (n & 4) a4 = *(const unsigned int *)(s + o4);
    if (n & 2) a2 = *(const unsigned short *)(s + o2);
    if (n & 1) a1 = *(const unsigned char *)(s + o1);
    if (n & 8) *(unsigned long long *)(d + o8) = a8;
    if (n & 4) *(unsigned int *)(d + o4) = a4;
    if (n & 2) *(unsigned short *)(d + o2) = a2;
    if (n & 1) *(unsigned char *)(d + o1) = a1;
}
template <int K>
__device__ __forceinline__ void lane_block(u32x4 *d, const u32x4 *s) {
    u32x4 r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = s[k];
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = r[k];
}
__device__ __forceinline__ void lane_copy(void *dstv, const void *srcv, size_t nbytes) {
    char *dst = (char *)dstv;
    const char *src = (const char *)srcv;
    if (((uintptr_t)dst ^ (uintptr_t)src) & 15) {
        copy<16, 4>(dst, src, nbytes, 0, 1);
        return;
    }
    const split<16> sp(dst, nbytes);
    if (sp.head) {
        if (((uintptr_t)(dst + sp.head) & 15) == 0) lane_edge<true>(dst, src, sp.head);
        else lane_bytes(dst, src, sp.head); /* the whole copy ends short of a boundary */
    }
    u32x4 *d = (u32x4 *)(dst + sp.head);
    const u32x4 *s = (const u32x4 *)(src + sp.head);
    size_t i = 0;
    for (; i + 16 <= sp.nvec; i += 16) lane_block<16>(d + i, s + i);
    if (i + 8 <= sp.nvec) lane_block<8>(d + i, s + i), i += 8;
    if (i + 4 <= sp.nvec) lane_block<4>(d + i, s + i), i += 4;
    for (; i < sp.nvec; ++i) d[i] = s[i];
    if (sp.tail) lane_edge<false>(dst + sp.tail_start, src + sp.tail_start, sp.tail);
}
} // namespace

#define NVSHMEMI_DEF_PUT(Name, Type, OpType)                                                \
    extern "C" __device__ void nvshmem_##Name##_put(Type *target, Type *source, int len, int pe) { \
        NVSHMEMI_IDEAL();                                                                   \
        if (bad_target(target, len > 0 ? (size_t)len * sizeof(Type) : 0, pe)) return;      \
        NVSHMEMI_TRACE(OpType, source, target, len, pe, 0);                                \
        if (len <= 0) return;                                                               \
        lane_copy(translate(target, pe), source, (size_t)len * sizeof(Type));             \
    }
NVSHMEMI_DEF_PUT(int, int, NVSHMEMI_OP_PUT_INT)
NVSHMEMI_DEF_PUT(float, float, NVSHMEMI_OP_PUT_FLOAT)
NVSHMEMI_DEF_PUT(double, double, NVSHMEMI_OP_PUT_DOUBLE)

#define NVSHMEMI_DEF_GET(Name, Type, OpType)                                                \
    extern "C" __device__ void nvshmem_##Name##_get(Type *target, Type *source, int len, int pe) { \
        NVSHMEMI_IDEAL();                                                                   \
        if (bad_target(source, len > 0 ? (size_t)len * sizeof(Type) : 0, pe)) return;      \
        NVSHMEMI_TRACE(OpType, target, source, len, pe, 0);                                \
        if (len <= 0) return;                                                               \
        lane_copy(target, translate(source, pe), (size_t)len * sizeof(Type));             \
    }
NVSHMEMI_DEF_GET(int, int, NVSHMEMI_OP_GET_INT)
NVSHMEMI_DEF_GET(float, float, NVSHMEMI_OP_GET_FLOAT)
NVSHMEMI_DEF_GET(double, double, NVSHMEMI_OP_GET_DOUBLE)

/* ---------------------------------------------- non-blocking posts -- */

/* Post a put (or get) for the next nvshmemx_nbi_flush_on_stream: one slot
 * of the calling PE's queue (the workgroup's PE in local mode, the
 * process's in process mode), claimed with one atomic.  A post that finds
 * the queue's share full (or a workgroup of no PE's) is copied at once by
 * the calling lane instead, which a non-blocking call may do.  Invalid
 * symmetric ranges are refused here, as for the other device calls;
 * zero-byte posts do nothing. */
namespace {
__device__ __forceinline__ void post_nbi(void *target, const void *source, size_t nbytes, int pe, bool get) {
    if (bad_target(get ? source : target, nbytes, pe) || nbytes == 0) return;
    const nvshmemi_device_state_t &s = nvshmemi_dstate;
    const bool local = s.mode == NVSHMEMI_MODE_LOCAL;
    const unsigned queues = local ? (unsigned)s.npes : 1u, q = local ? (unsigned)my_pe() : 0u;
    const unsigned share = kNbiSlots / queues;
    /* a workgroup past the last PE's (a grid larger than the PEs') has no queue */
    const unsigned slot = q < queues ? __hip_atomic_fetch_add(&nbi_count[q], 1u, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                     : share;
    if (slot < share) {
        nvshmemx_put_request_t &r = nbi_queue[(size_t)q * share + slot];
        r.target = target;
        r.source = source;
        r.nbytes = nbytes;
        r.pe = pe;
        r.elem_bytes = get ? kNbiGet : 1;
    } else if (get) {
        lane_copy(target, translate(source, pe), nbytes);
    } else {
        lane_copy(translate(target, pe), source, nbytes);
    }
}
} // namespace

extern "C" __device__ void nvshmemx_putmem_nbi(void *target, const void *source, size_t nbytes, int pe) {
    post_nbi(target, source, nbytes, pe, false);
}
extern "C" __device__ void nvshmemx_getmem_nbi(void *target, const void *source, size_t nbytes, int pe) {
    post_nbi(target, source, nbytes, pe, true);
}

#define NVSHMEMI_DEF_PUT_NBI(Name, Type)                                                          \
    extern "C" __device__ void nvshmemx_##Name##_put_nbi(Type *target, const Type *source, int len, int pe) { \
        if (len > 0) nvshmemx_putmem_nbi(target, source, (size_t)len * sizeof(Type), pe);             \
    }
NVSHMEMI_DEF_PUT_NBI(int, int)
NVSHMEMI_DEF_PUT_NBI(float, float)
NVSHMEMI_DEF_PUT_NBI(double, double)
#define NVSHMEMI_DEF_GET_NBI(Name, Type)                                                          \
    extern "C" __device__ void nvshmemx_##Name##_get_nbi(Type *target, const Type *source, int len, int pe) { \
        if (len > 0) nvshmemx_getmem_nbi(target, source, (size_t)len * sizeof(Type), pe);             \
    }
NVSHMEMI_DEF_GET_NBI(int, int)
NVSHMEMI_DEF_GET_NBI(float, float)
NVSHMEMI_DEF_GET_NBI(double, double)

/* --------------------------------------------------------------- waits -- */

extern "C" __device__ void nvshmem_int_wait(int *ivar, int cmp_value) {
    /* reference nvshmem_device.cu:542-562: block while *ivar == cmp_value */
    NVSHMEMI_IDEAL();
    NVSHMEMI_TRACE_TIMED_BEGIN();
    wait_until(ivar, CMP_NE, cmp_value);
    NVSHMEMI_TRACE_TIMED_END(NVSHMEMI_OP_WAIT_INT, ivar, (const void *)0, 1, my_pe());
}

extern "C" __device__ void nvshmem_int_wait_until(int *ivar, int cmp, int cmp_value) {
    /* reference nvshmem_device.cu:564-584: block until *ivar == cmp_value;
     * the comparison operator `cmp` is ignored, as in the reference */
    (void)cmp;
    NVSHMEMI_IDEAL();
    NVSHMEMI_TRACE_TIMED_BEGIN();
    wait_until(ivar, CMP_EQ, cmp_value);
    NVSHMEMI_TRACE_TIMED_END(NVSHMEMI_OP_WAIT_UNTIL_INT, ivar, (const void *)0, 1, my_pe());
}

extern "C" __device__ int nvshmemx_int_wait_until_cmp(int *ivar, int cmp, int value) {
    return wait_until(ivar, cmp, value);
}

/* ---------------------------------------------------------- extensions -- */

extern "C" __device__ int nvshmemx_my_pe_d(void) { return my_pe(); }
extern "C" __device__ int nvshmemx_n_pes_d(void) { return nvshmemi_dstate.npes; }
extern "C" __device__ void *nvshmemx_ptr_d(const void *addr, int pe) {
    return (unsigned)pe < (unsigned)nvshmemi_dstate.npes ? (void *)translate(addr, pe) : nullptr;
}

namespace {
/* rank of the calling lane among the wave's active lanes, and their count */
__device__ __forceinline__ void active_lanes(unsigned &rank, unsigned &count) {
    const uint64_t exec = __builtin_amdgcn_read_exec();
    rank = __builtin_amdgcn_mbcnt_hi((unsigned)(exec >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)exec, 0u));
    count = (unsigned)__builtin_popcountll(exec);
}
__device__ __forceinline__ void wave_copy(void *dst, const void *src, size_t nbytes) {
    unsigned r, n;
    active_lanes(r, n);
    copy<16, 4>(dst, src, nbytes, r, n);
}
__device__ __forceinline__ void block_copy(void *dst, const void *src, size_t nbytes) {
    copy<16, 4>(dst, src, nbytes, flat_tid(), block_threads());
    fence_release();
    __syncthreads();
}
} // namespace

extern "C" __device__ void nvshmemx_putmem_wave(void *target, const void *source, size_t nbytes, int pe) {
    if (!bad_target(target, nbytes, pe)) wave_copy(translate(target, pe), source, nbytes);
}
extern "C" __device__ void nvshmemx_getmem_wave(void *target, const void *source, size_t nbytes, int pe) {
    if (!bad_target(source, nbytes, pe)) wave_copy(target, translate(source, pe), nbytes);
}
extern "C" __device__ void nvshmemx_putmem_block(void *target, const void *source, size_t nbytes, int pe) {
    if (!bad_target(target, nbytes, pe)) block_copy(translate(target, pe), source, nbytes);
}
extern "C" __device__ void nvshmemx_getmem_block(void *target, const void *source, size_t nbytes, int pe) {
    if (!bad_target(source, nbytes, pe)) block_copy(target, translate(source, pe), nbytes);
}

#define NVSHMEMI_DEF_COOP(Name, Type)                                                              \
    extern "C" __device__ void nvshmemx_##Name##_put_wave(Type *t, const Type *s, int len, int pe) { \
        if (len > 0 && !bad_target(t, (size_t)len * sizeof(Type), pe))                             \
            wave_copy(translate(t, pe), s, (size_t)len * sizeof(Type));                            \
    }                                                                                              \
    extern "C" __device__ void nvshmemx_##Name##_get_wave(Type *t, const Type *s, int len, int pe) { \
        if (len > 0 && !bad_target(s, (size_t)len * sizeof(Type), pe))                             \
            wave_copy(t, translate(s, pe), (size_t)len * sizeof(Type));                            \
    }                                                                                              \
    extern "C" __device__ void nvshmemx_##Name##_put_block(Type *t, const Type *s, int len, int pe) { \
        if (!bad_target(t, len > 0 ? (size_t)len * sizeof(Type) : 0, pe))                          \
            block_copy(translate(t, pe), s, len > 0 ? (size_t)len * sizeof(Type) : 0);              \
    }                                                                                              \
    extern "C" __device__ void nvshmemx_##Name##_get_block(Type *t, const Type *s, int len, int pe) { \
        if (!bad_target(s, len > 0 ? (size_t)len * sizeof(Type) : 0, pe))                          \
            block_copy(t, translate(s, pe), len > 0 ? (size_t)len * sizeof(Type) : 0);              \
    }
NVSHMEMI_DEF_COOP(int, int)
NVSHMEMI_DEF_COOP(float, float)
NVSHMEMI_DEF_COOP(double, double)

/* ------------------------------------------ strided and 2-D transfers -- */

namespace {
template <typename T>
__device__ __forceinline__ void strided(T *dst, const T *src, ptrdiff_t dst_st, ptrdiff_t src_st, size_t n,
                                        size_t lane, size_t lanes) {
    for (size_t i = lane; i < n; i += lanes) dst[(ptrdiff_t)i * dst_st] = src[(ptrdiff_t)i * src_st];
}
/* rows go to the workgroup's wavefronts in turn; each row is a wavefront copy.
 * A workgroup whose size is not a multiple of 64 ends with a partial
 * wavefront, which copies its rows with the lanes it has. */
__device__ __forceinline__ void rows_block(char *dst, size_t dpitch, const char *src, size_t spitch, size_t width,
                                           size_t height) {
    const unsigned threads = block_threads();
    const unsigned waves = (threads + 63u) / 64u, w = flat_tid() / 64u, lane = flat_tid() & 63u;
    const unsigned lanes = threads - w * 64u < 64u ? threads - w * 64u : 64u;
    for (size_t r = w; r < height; r += waves) copy<16, 4>(dst + r * dpitch, src + r * spitch, width, lane, lanes);
    fence_release();
    __syncthreads();
}
} // namespace

#define NVSHMEMI_DEF_STRIDED(Name, Type)                                                                          \
    extern "C" __device__ void nvshmemx_##Name##_iput_wave(Type *t, const Type *s, ptrdiff_t tst, ptrdiff_t sst,  \
                                                           size_t n, int pe) {                                   \
        unsigned r, c;                                                                                            \
        active_lanes(r, c);                                                                                       \
        if (!bad_target(t, strided_extent(n, tst, sizeof(Type)), pe))                                             \
            strided((Type *)translate(t, pe), s, tst, sst, n, r, c);                                              \
    }                                                                                                             \
    extern "C" __device__ void nvshmemx_##Name##_iget_wave(Type *t, const Type *s, ptrdiff_t tst, ptrdiff_t sst,  \
                                                           size_t n, int pe) {                                   \
        unsigned r, c;                                                                                            \
        active_lanes(r, c);                                                                                       \
        if (!bad_target(s, strided_extent(n, sst, sizeof(Type)), pe))                                             \
            strided(t, (const Type *)translate(s, pe), tst, sst, n, r, c);                                        \
    }                                                                                                             \
    extern "C" __device__ void nvshmemx_##Name##_iput_block(Type *t, const Type *s, ptrdiff_t tst, ptrdiff_t sst, \
                                                            size_t n, int pe) {                                  \
        if (!bad_target(t, strided_extent(n, tst, sizeof(Type)), pe))                                             \
            strided((Type *)translate(t, pe), s, tst, sst, n, flat_tid(), block_threads());                       \
        fence_release();                                                                                          \
        __syncthreads();                                                                                          \
    }                                                                                                             \
    extern "C" __device__ void nvshmemx_##Name##_iget_block(Type *t, const Type *s, ptrdiff_t tst, ptrdiff_t sst, \
                                                            size_t n, int pe) {                                  \
        if (!bad_target(s, strided_extent(n, sst, sizeof(Type)), pe))                                             \
            strided(t, (const Type *)translate(s, pe), tst, sst, n, flat_tid(), block_threads());                 \
        fence_release();                                                                                          \
        __syncthreads();                                                                                          \
    }
NVSHMEMI_DEF_STRIDED(int, int)
NVSHMEMI_DEF_STRIDED(float, float)
NVSHMEMI_DEF_STRIDED(double, double)

extern "C" __device__ void nvshmemx_putmem_2d_block(void *target, size_t tpitch, const void *source, size_t spitch,
                                                    size_t width, size_t height, int pe) {
    if (bad_target(target, height ? (height - 1) * tpitch + width : 0, pe))
        height = 0; /* every thread still meets the workgroup barrier */
    rows_block(bad_pe_target(target, pe), tpitch, (const char *)source, spitch, width, height);
}
extern "C" __device__ void nvshmemx_getmem_2d_block(void *target, size_t tpitch, const void *source, size_t spitch,
                                                    size_t width, size_t height, int pe) {
    if (bad_target(source, height ? (height - 1) * spitch + width : 0, pe)) height = 0;
    rows_block((char *)target, tpitch, bad_pe_target(source, pe), spitch, width, height);
}

/* ------------------------------------------------ workgroup barrier -- */

/* Count-and-generation barrier over every workgroup of the grid (the
 * protocol of the offload barrier, one arrival per workgroup, on a word of
 * its own).  No per-workgroup state survives between launches, so grids of
 * different sizes may reuse it. */
extern "C" __device__ void nvshmemx_barrier_all_wg(void) {
    fence_release();
    __syncthreads();
    if (flat_tid() == 0) {
        unsigned int *word = (unsigned int *)(nvshmemi_dstate.heap_base[0] + NVSHMEMI_OFF_WGBAR_WORD);
        gen_barrier(word, (int)(gridDim.x * gridDim.y * gridDim.z), NVSHMEMI_ERR_WAIT_TIMEOUT);
    }
    __syncthreads();
}
