/*
 * nvshmemi_device.h -- inline device primitives shared by the device API, the
 * host-launched copy kernels, the batched-put engine and the harness kernels.
 *
 * Memory model (gfx950, LLVM AMDGPU memory model for GFX942/GFX950):
 *  - every PE lives on the same agent (one MI355X), so agent scope is the
 *    widest scope any ordering here needs; the reference's system-scope fence
 *    (__threadfence_system, nvshmem_device.cu:434) would also write back and
 *    invalidate for the host, which no PE is;
 *  - the eight XCDs each have an L2 that is not coherent with the others for
 *    ordinary (coarse-grained) device memory.  Agent-scope release writes the
 *    local L2 back (buffer_wbl2 sc1) and agent-scope acquire invalidates it
 *    (buffer_inv sc1); flag words are accessed with agent-scope atomic
 *    loads/stores (sc1), which bypass the non-coherent L2 state;
 *  - bulk payload is moved with plain (or non-temporal) 16-byte vector
 *    accesses and published by one release, never by per-element volatile
 *    accesses as in the reference (nvshmem_device.cu:497-500).
 */
#ifndef NVSHMEMI_DEVICE_H
#define NVSHMEMI_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nvshmemi_state.h"
#include "nvshmemx.h"

extern __constant__ nvshmemi_device_state_t nvshmemi_dstate;

namespace nvshmemi {

/* ---------------------------------------------------------------- state -- */

/* Puts posted with nvshmemx_putmem_nbi wait here for
 * nvshmemx_nbi_flush_on_stream: kNbiSlots requests split evenly between the
 * queues of the PEs this process holds (all of them in local mode, its own
 * in process mode); nbi_count[q] counts queue q's posts since its last flush
 * (posts past the queue's share were copied at once and are not in it). */
constexpr unsigned kNbiSlots = 65536;
/* a posted get is queued as a request whose target is the local address and
 * whose source is the symmetric one, marked by this element width */
constexpr int kNbiGet = -1;
extern __device__ nvshmemx_put_request_t nbi_queue[kNbiSlots];
extern __device__ unsigned nbi_count[NVSHMEMI_MAX_PES];

__device__ __forceinline__ char *translate(const void *addr, int pe) {
    const nvshmemi_device_state_t &s = nvshmemi_dstate;
    uint64_t off = ((uint64_t)((const char *)addr - s.anchor)) & s.offset_mask;
    return s.heap_base[pe] + off;
}

/* PE of the calling workgroup: the process PE, or blockIdx.x >> wg_shift in
 * local (workgroup-partitioned) mode */
__device__ __forceinline__ int my_pe() {
    return nvshmemi_dstate.mype + (int)(blockIdx.x >> nvshmemi_dstate.wg_shift);
}

__device__ __forceinline__ void set_error(int bits) {
    __hip_atomic_fetch_or(nvshmemi_dstate.errors, bits, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_SYSTEM);
}

/* --------------------------------------------------------- memory model -- */

__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_acq_rel() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }

template <typename T> struct bits_of;
template <> struct bits_of<int> { typedef int type; };
template <> struct bits_of<float> { typedef int type; };
template <> struct bits_of<double> { typedef long long type; };
template <> struct bits_of<long long> { typedef long long type; };
template <> struct bits_of<unsigned int> { typedef unsigned int type; };

template <typename T> __device__ __forceinline__ void store_relaxed(T *p, T v) {
    typedef typename bits_of<T>::type B;
    __hip_atomic_store((B *)p, __builtin_bit_cast(B, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T load_relaxed(const T *p) {
    typedef typename bits_of<T>::type B;
    B b = __hip_atomic_load((B *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, b);
}
template <typename T> __device__ __forceinline__ void store_release(T *p, T v) {
    typedef typename bits_of<T>::type B;
    __hip_atomic_store((B *)p, __builtin_bit_cast(B, v), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T load_acquire(const T *p) {
    typedef typename bits_of<T>::type B;
    B b = __hip_atomic_load((B *)p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, b);
}

/* --------------------------------------------------------------- waits -- */

enum cmp_op { CMP_EQ = 0, CMP_NE = 1, CMP_GT = 2, CMP_GE = 3, CMP_LT = 4, CMP_LE = 5 };

__device__ __forceinline__ bool cmp_holds(int v, int op, int ref) {
    switch (op) {
    case CMP_EQ: return v == ref;
    case CMP_NE: return v != ref;
    case CMP_GT: return v > ref;
    case CMP_GE: return v >= ref;
    case CMP_LT: return v < ref;
    default: return v <= ref;
    }
}

/* Polling pace shared by every wait: the first NVSHMEMI_SPIN_POLLS polls go
 * back to back (a partner that is already running answers within a
 * microsecond or two, and sleeping then only adds latency) and read no clock:
 * the wall-clock read is a scalar memory access whose result the loop would
 * wait for after every poll, which measured 40-80 ns per poll.  After that an
 * s_sleep ladder (64 clocks per step, doubling to 8 steps) lets a long wait
 * yield issue slots, and the wall-clock budget starts. */
#ifndef NVSHMEMI_SPIN_POLLS
#define NVSHMEMI_SPIN_POLLS 256
#endif
struct sleep_ladder {
    int steps = 1;
    __device__ __forceinline__ void pause() {
        for (int i = 0; i < steps; ++i) __builtin_amdgcn_s_sleep(1);
        if (steps < 8) steps <<= 1;
    }
};

/* Poll *ivar with agent-scope loads until `v op ref` holds: back to back
 * first, then paced by the sleep ladder.  A wait that outlives the configured
 * wall-clock budget records NVSHMEMI_ERR_WAIT_TIMEOUT and returns instead of
 * hanging the device.  Returns the last value read.  Ends with an agent-scope
 * acquire so payload published before the flag is visible to the caller. */
__device__ __forceinline__ int wait_until(int *ivar, int op, int ref) {
    int v = load_relaxed(ivar);
    for (int i = 0; i < NVSHMEMI_SPIN_POLLS && !cmp_holds(v, op, ref); ++i) v = load_relaxed(ivar);
    if (!cmp_holds(v, op, ref)) {
        const uint64_t budget = nvshmemi_dstate.wait_timeout_ticks;
        const uint64_t t0 = wall_clock64();
        sleep_ladder pace;
        for (;;) {
            pace.pause();
            v = load_relaxed(ivar);
            if (cmp_holds(v, op, ref)) break;
            if ((uint64_t)(wall_clock64() - t0) > budget) {
                set_error(NVSHMEMI_ERR_WAIT_TIMEOUT);
                break;
            }
        }
    }
    fence_acquire();
    return v;
}

/* One episode of the generation barrier on `word` among `n` arrivals (n <
 * 65536), called by one lane per arrival.  The word holds the arrival count
 * in its low 16 bits and the generation in its high 16 bits, so one
 * agent-scope fetch-add both arrives and reports the generation of this
 * episode (it cannot advance before this arrival).  The last arriver stores
 * the next generation with a zero count, with release; the others poll the
 * generation (back to back, then with the sleep ladder) until it moves or the
 * wait budget runs out (then `timeout_bit` is set and the episode is
 * abandoned).  No arrival of the next episode can happen before that store:
 * it needs the new generation.  Ends with an acquire, so what every arrival
 * published before arriving is visible. */
__device__ __forceinline__ void gen_barrier(unsigned int *word, int n, int timeout_bit) {
    const unsigned int old = __hip_atomic_fetch_add(word, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int g0 = old >> 16;
    if ((old & 0xFFFFu) == (unsigned int)(n - 1)) {
        store_release(word, ((g0 + 1u) & 0xFFFFu) << 16);
        return;
    }
    bool moved = false;
    for (int i = 0; i < NVSHMEMI_SPIN_POLLS && !moved; ++i) moved = (load_relaxed(word) >> 16) != g0;
    if (!moved) {
        const uint64_t budget = nvshmemi_dstate.wait_timeout_ticks;
        const uint64_t t0 = wall_clock64();
        sleep_ladder pace;
        while ((load_relaxed(word) >> 16) == g0) {
            pace.pause();
            if ((uint64_t)(wall_clock64() - t0) > budget) {
                set_error(timeout_bit);
                break;
            }
        }
    }
    fence_acquire();
}

/* ---------------------------------------------------------- bulk copies -- */

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <int W> struct vec;
template <> struct vec<16> { typedef u32x4 t; };
template <> struct vec<8> { typedef u32x2 t; };
template <> struct vec<4> { typedef unsigned int t; };
template <> struct vec<2> { typedef unsigned short t; };
template <> struct vec<1> { typedef unsigned char t; };

template <bool NT, typename V> __device__ __forceinline__ V vload(const V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename V> __device__ __forceinline__ void vstore(V *p, V v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

/* Cache policy bits: NT_LOAD streams loads past the caches' normal
 * retention, NT_STORE does the same for stores (global_load/store ... nt);
 * ST_SC1 writes stores through the L2 at device scope (global_store ... sc1)
 * instead, so a kernel leaves no dirty lines for its end-of-kernel release
 * to write back. */
enum { NT_NONE = 0, NT_LOAD = 1, NT_STORE = 2, NT_BOTH = 3, ST_SC1 = 4 };

/* a store under the policy bits of P (ST_SC1 takes precedence over NT_STORE;
 * the write-through form exists for 4-, 8- and 16-byte vectors) */
template <int P, typename V> __device__ __forceinline__ void vstore_p(V *p, V v) {
    if constexpr ((P & ST_SC1) != 0 && sizeof(V) == 16)
        asm volatile("global_store_dwordx4 %0, %1, off sc1" : : "v"(p), "v"(v) : "memory");
    else if constexpr ((P & ST_SC1) != 0 && sizeof(V) == 8)
        asm volatile("global_store_dwordx2 %0, %1, off sc1" : : "v"(p), "v"(v) : "memory");
    else if constexpr ((P & ST_SC1) != 0 && sizeof(V) == 4)
        asm volatile("global_store_dword %0, %1, off sc1" : : "v"(p), "v"(v) : "memory");
    else vstore<(P & NT_STORE) != 0>(p, v);
}

/* Stream vectors [lo, hi) with `n` cooperating lanes (lane index `lane`),
 * UNROLL independent loads in flight per lane; lanes are interleaved so each
 * wavefront access covers 64 * sizeof(V) contiguous bytes. */
template <int UNROLL, int NT, typename V>
__device__ __forceinline__ void stream_vectors(V *d, const V *s, size_t lo, size_t hi, size_t lane, size_t n) {
    size_t i = lo + lane;
    if (UNROLL > 1) {
        const size_t step = n * UNROLL;
        for (; i + (UNROLL - 1) * n < hi; i += step) {
            V r[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) r[u] = vload<(NT & NT_LOAD) != 0>(s + i + u * n);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) vstore_p<NT>(d + i + u * n, r[u]);
        }
    }
    for (; i < hi; i += n) vstore_p<NT>(d + i, vload<(NT & NT_LOAD) != 0>(s + i));
}

/* Split of a W-aligned copy: `head` unaligned leading bytes, `nvec` W-byte
 * vectors, then the trailing bytes from `tail_start`.  Requires
 * (dst ^ src) % W == 0 so that one head aligns both sides. */
template <int W> struct split {
    size_t head, nvec, tail_start, tail;
    __device__ __forceinline__ split(const char *dst, size_t nbytes) {
        head = (size_t)((W - ((uintptr_t)dst & (W - 1))) & (W - 1));
        if (head > nbytes) head = nbytes;
        nvec = (nbytes - head) / W;
        tail_start = head + nvec * W;
        tail = nbytes - tail_start;
    }
};

/* Interleaved copy: `n` lanes with lane index `lane` share the whole range. */
template <int W, int UNROLL, int NT>
__device__ __forceinline__ void copy_w(char *dst, const char *src, size_t nbytes, size_t lane, size_t n) {
    typedef typename vec<W>::t V;
    const split<W> sp(dst, nbytes);
    if (W > 1) {
        for (size_t i = lane; i < sp.head; i += n) dst[i] = src[i];
        for (size_t i = lane; i < sp.tail; i += n) dst[sp.tail_start + i] = src[sp.tail_start + i];
    }
    stream_vectors<UNROLL, NT>((V *)(dst + sp.head), (const V *)(src + sp.head), 0, sp.nvec, lane, n);
}

/* Chunked copy: workgroup `part` of `parts` streams one contiguous slice of
 * the vectors with its `n` threads; part 0 also copies head and tail bytes. */
template <int W, int UNROLL, int NT>
__device__ __forceinline__ void copy_w_chunked(char *dst, const char *src, size_t nbytes, size_t part, size_t parts,
                                               size_t lane, size_t n) {
    typedef typename vec<W>::t V;
    const split<W> sp(dst, nbytes);
    if (W > 1 && part == 0) {
        for (size_t i = lane; i < sp.head; i += n) dst[i] = src[i];
        for (size_t i = lane; i < sp.tail; i += n) dst[sp.tail_start + i] = src[sp.tail_start + i];
    }
    const size_t per = (sp.nvec + parts - 1) / parts;
    const size_t lo = part * per;
    const size_t hi = lo + per < sp.nvec ? lo + per : sp.nvec;
    if (lo < hi) stream_vectors<UNROLL, NT>((V *)(dst + sp.head), (const V *)(src + sp.head), lo, hi, lane, n);
}

/* Tiled copy: the vectors are cut into tiles of UNROLL * n (one load per lane
 * per UNROLL step), and the tiles into `streams` equal runs; workgroup-sized
 * worker `part` of `parts` takes tiles part, part + parts, ...  in an order
 * that alternates between the runs (tile t -> run t % streams, position
 * t / streams).  With one tile per workgroup the hardware dispatches the grid
 * in tile order, so the device sweeps `streams` address windows at once.
 * Part 0 also copies the unaligned head and tail bytes. */
template <int W, int UNROLL, int NT>
__device__ __forceinline__ void copy_w_tiled(char *dst, const char *src, size_t nbytes, size_t streams, size_t part,
                                             size_t parts, size_t lane, size_t n) {
    typedef typename vec<W>::t V;
    const split<W> sp(dst, nbytes);
    if (W > 1 && part == 0) {
        for (size_t i = lane; i < sp.head; i += n) dst[i] = src[i];
        for (size_t i = lane; i < sp.tail; i += n) dst[sp.tail_start + i] = src[sp.tail_start + i];
    }
    V *d = (V *)(dst + sp.head);
    const V *s = (const V *)(src + sp.head);
    const size_t tile = (size_t)UNROLL * n;
    const size_t tiles = (sp.nvec + tile - 1) / tile;
    const size_t run = (tiles + streams - 1) / streams;
    for (size_t t = part; t < run * streams; t += parts) {
        const size_t base = ((t % streams) * run + t / streams) * tile + lane;
        V r[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (base + u * n < sp.nvec) r[u] = vload<(NT & NT_LOAD) != 0>(s + base + u * n);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (base + u * n < sp.nvec) vstore_p<NT>(d + base + u * n, r[u]);
    }
}

template <int MAXW, int UNROLL, int NT = NT_NONE>
__device__ __forceinline__ void copy_tiled(void *dstv, const void *srcv, size_t nbytes, size_t streams, size_t part,
                                           size_t parts, size_t lane, size_t n) {
    char *dst = (char *)dstv;
    const char *src = (const char *)srcv;
    const uintptr_t rel = (uintptr_t)dst ^ (uintptr_t)src;
    if (MAXW >= 16 && (rel & 15) == 0)
        copy_w_tiled<(MAXW >= 16 ? 16 : 1), UNROLL, NT>(dst, src, nbytes, streams, part, parts, lane, n);
    else if (MAXW >= 8 && (rel & 7) == 0)
        copy_w_tiled<(MAXW >= 8 ? 8 : 1), UNROLL, NT>(dst, src, nbytes, streams, part, parts, lane, n);
    else if (MAXW >= 4 && (rel & 3) == 0)
        copy_w_tiled<(MAXW >= 4 ? 4 : 1), UNROLL, NT>(dst, src, nbytes, streams, part, parts, lane, n);
    else if ((rel & 1) == 0) copy_w_tiled<2, UNROLL, NT>(dst, src, nbytes, streams, part, parts, lane, n);
    else copy_w_tiled<1, UNROLL, NT>(dst, src, nbytes, streams, part, parts, lane, n);
}

/* Pick the widest access width <= MAXW that the relative alignment of dst and
 * src allows, then copy.  Every lane of the cooperating group must call this
 * with identical arguments. */
template <int MAXW, int UNROLL, int NT = NT_NONE>
__device__ __forceinline__ void copy(void *dstv, const void *srcv, size_t nbytes, size_t lane, size_t n) {
    char *dst = (char *)dstv;
    const char *src = (const char *)srcv;
    const uintptr_t rel = (uintptr_t)dst ^ (uintptr_t)src;
    if (MAXW >= 16 && (rel & 15) == 0) copy_w<(MAXW >= 16 ? 16 : 1), UNROLL, NT>(dst, src, nbytes, lane, n);
    else if (MAXW >= 8 && (rel & 7) == 0) copy_w<(MAXW >= 8 ? 8 : 1), UNROLL, NT>(dst, src, nbytes, lane, n);
    else if (MAXW >= 4 && (rel & 3) == 0) copy_w<(MAXW >= 4 ? 4 : 1), UNROLL, NT>(dst, src, nbytes, lane, n);
    else if ((rel & 1) == 0) copy_w<2, UNROLL, NT>(dst, src, nbytes, lane, n);
    else copy_w<1, UNROLL, NT>(dst, src, nbytes, lane, n);
}

template <int MAXW, int UNROLL, int NT = NT_NONE>
__device__ __forceinline__ void copy_chunked(void *dstv, const void *srcv, size_t nbytes, size_t part, size_t parts,
                                             size_t lane, size_t n) {
    char *dst = (char *)dstv;
    const char *src = (const char *)srcv;
    const uintptr_t rel = (uintptr_t)dst ^ (uintptr_t)src;
    if (MAXW >= 16 && (rel & 15) == 0)
        copy_w_chunked<(MAXW >= 16 ? 16 : 1), UNROLL, NT>(dst, src, nbytes, part, parts, lane, n);
    else if (MAXW >= 8 && (rel & 7) == 0)
        copy_w_chunked<(MAXW >= 8 ? 8 : 1), UNROLL, NT>(dst, src, nbytes, part, parts, lane, n);
    else if (MAXW >= 4 && (rel & 3) == 0)
        copy_w_chunked<(MAXW >= 4 ? 4 : 1), UNROLL, NT>(dst, src, nbytes, part, parts, lane, n);
    else if ((rel & 1) == 0) copy_w_chunked<2, UNROLL, NT>(dst, src, nbytes, part, parts, lane, n);
    else copy_w_chunked<1, UNROLL, NT>(dst, src, nbytes, part, parts, lane, n);
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned flat_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ unsigned flat_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ unsigned block_threads() { return blockDim.x * blockDim.y * blockDim.z; }

} // namespace nvshmemi

#endif /* NVSHMEMI_DEVICE_H */
