/*
 * harness.hip -- caller-side kernels that exercise the device API the way the
 * reference's microbenchmarks do (microbench/shmem_latency.cu:48-81,
 * microbench/shmem_sync_latency.cu:48-78), plus the seeded data generator and
 * the ordering checks.  They are ordinary callers of the public device API and
 * are exported through plain C entry points (prefix nvshmemh_) so the test and
 * benchmark drivers can launch them without a compiler.
 *
 * Seeded data: PCG32 (O'Neill 2014, pcg32_random_r / pcg32_srandom_r of the
 * PCG "minimal C" reference), word i of a buffer = the i-th output after
 * seeding with (seed, seq).  Each lane jumps ahead with the LCG advance
 * formula (O(log n) multiply-adds), so a wavefront writes 64 consecutive words
 * per store.
 */
#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <vector>

#include "nvshmem.h"
#include "nvshmem_device.h"
#include "nvshmemx.h"
#include "nvshmemx_device.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

namespace {

constexpr uint64_t kPcgMult = 6364136223846793005ULL;

__host__ __device__ inline uint32_t pcg_output(uint64_t old) {
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

/* state after `delta` LCG steps from `state` (Brown, "Random number generation
 * with arbitrary strides", 1994) */
__host__ __device__ inline uint64_t pcg_advance(uint64_t state, uint64_t delta, uint64_t inc) {
    uint64_t acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
    while (delta) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

__host__ __device__ inline uint64_t pcg_seeded_state(uint64_t seed, uint64_t seq, uint64_t &inc) {
    inc = (seq << 1u) | 1u;
    uint64_t s = 0;
    s = s * kPcgMult + inc;
    s += seed;
    s = s * kPcgMult + inc;
    return s;
}

constexpr int kFillWordsPerLane = 64;

/* harness loops stop at the first recorded device error (a timed-out wait),
 * so a broken protocol costs one wait budget, not one per iteration */
__device__ __forceinline__ bool failed() { return nvshmemi::load_relaxed(nvshmemi_dstate.errors) != 0; }

__global__ void k_fill_pcg32(uint32_t *dst, size_t nwords, uint64_t s0, uint64_t inc, uint64_t m64, uint64_t a64) {
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const unsigned lane = threadIdx.x & 63u;
    const size_t base = wave * 64 * kFillWordsPerLane;
    if (base >= nwords) return;
    uint64_t st = pcg_advance(s0, base + lane, inc);
    for (int j = 0; j < kFillWordsPerLane; ++j) {
        const size_t idx = base + (size_t)j * 64 + lane;
        if (idx < nwords) dst[idx] = pcg_output(st);
        st = m64 * st + a64;
    }
}

/* ---- per-thread API drivers ---- */

template <typename T>
__device__ void thread_put(T *target, T *source, int len, int pe);
template <> __device__ void thread_put<int>(int *t, int *s, int l, int pe) { nvshmem_int_put(t, s, l, pe); }
template <> __device__ void thread_put<float>(float *t, float *s, int l, int pe) { nvshmem_float_put(t, s, l, pe); }
template <> __device__ void thread_put<double>(double *t, double *s, int l, int pe) { nvshmem_double_put(t, s, l, pe); }
template <typename T>
__device__ void thread_get(T *target, T *source, int len, int pe);
template <> __device__ void thread_get<int>(int *t, int *s, int l, int pe) { nvshmem_int_get(t, s, l, pe); }
template <> __device__ void thread_get<float>(float *t, float *s, int l, int pe) { nvshmem_float_get(t, s, l, pe); }
template <> __device__ void thread_get<double>(double *t, double *s, int l, int pe) { nvshmem_double_get(t, s, l, pe); }

/* thread i copies elements [i*chunk, min((i+1)*chunk, len)) with one
 * per-thread call, so every thread's copy is complete on its own */
template <typename T>
__global__ void k_thread_putget(int op, T *target, T *source, int len, int chunk, int pe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const long long lo = (long long)i * chunk;
    if (lo >= len) return;
    const int n = (int)((lo + chunk <= len) ? chunk : len - lo);
    if (op == 0) thread_put<T>(target + lo, source + lo, n, pe);
    else thread_get<T>(target + lo, source + lo, n, pe);
    nvshmem_quiet();
}

/* scope 0: every wave copies its slice with the wave-cooperative call;
 * scope 1: every workgroup copies its slice with the block-cooperative call */
__global__ void k_coop_putget(int op, int scope, char *target, const char *source, size_t nbytes, size_t slice,
                              int pe) {
    size_t unit = scope == 0 ? ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64 : (size_t)blockIdx.x;
    const size_t lo = unit * slice;
    const size_t n = lo >= nbytes ? 0 : (lo + slice <= nbytes ? slice : nbytes - lo);
    if (scope == 0) {
        if (n == 0) return;
        if (op == 0) nvshmemx_putmem_wave(target + lo, source + lo, n, pe);
        else nvshmemx_getmem_wave(target + lo, source + lo, n, pe);
    } else {
        if (op == 0) nvshmemx_putmem_block(target + lo, source + lo, n, pe);
        else nvshmemx_getmem_block(target + lo, source + lo, n, pe);
    }
}

/* thread i: p(target + i, values[i], pe) or out[i] = g(target + i, pe) */
template <typename T>
__global__ void k_pg(int op, T *target, const T *values, T *out, int n, int pe);
template <>
__global__ void k_pg<int>(int op, int *t, const int *v, int *o, int n, int pe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) nvshmem_int_p(t + i, v[i], pe);
    else o[i] = nvshmem_int_g(t + i, pe);
}
template <>
__global__ void k_pg<float>(int op, float *t, const float *v, float *o, int n, int pe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) nvshmem_float_p(t + i, v[i], pe);
    else o[i] = nvshmem_float_g(t + i, pe);
}
template <>
__global__ void k_pg<double>(int op, double *t, const double *v, double *o, int n, int pe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) nvshmem_double_p(t + i, v[i], pe);
    else o[i] = nvshmem_double_g(t + i, pe);
}

/* ---- ping-pong (p + wait round trips, optional payload put) ----
 * Initiator (PE `a`): for i = 1..iters: [put payload; quiet;] p(flag, i, b);
 * wait(flag, i-1) until b's reply arrives.  Responder: wait(flag, i-1);
 * [put payload back; quiet;] p(flag, i, a).  Waiting with "while equal to the
 * previous value" cannot miss an update the way the reference's
 * wait_until(flag, i+1) equality can (microbench/shmem_sync_latency.cu:60-68).
 * samples[i] = wall_clock64 ticks of round trip i (initiator only). */
__device__ void pingpong_body(int *flag_sym, int me, int peer, int initiator, int iters, int *payload_sym,
                              int payload_len, long long *samples) {
    int *myflag = (int *)nvshmemx_ptr_d(flag_sym, me);
    int *mypayload = payload_len > 0 ? (int *)nvshmemx_ptr_d(payload_sym, me) : nullptr;
    for (int i = 1; i <= iters; ++i) {
        if (initiator) {
            const long long t0 = wall_clock64();
            if (payload_len > 0) {
                nvshmem_int_put(payload_sym, mypayload, payload_len, peer);
                nvshmem_quiet();
            }
            nvshmem_int_p(flag_sym, i, peer);
            nvshmem_quiet();
            nvshmem_int_wait(myflag, i - 1);
            const long long t1 = wall_clock64();
            if (samples) samples[i - 1] = t1 - t0;
            if (failed()) break;
        } else {
            nvshmem_int_wait(myflag, i - 1);
            if (payload_len > 0) {
                nvshmem_int_put(payload_sym, mypayload, payload_len, peer);
                nvshmem_quiet();
            }
            nvshmem_int_p(flag_sym, i, peer);
            nvshmem_quiet();
            /* checked after the reply: the check is a load that misses L2
             * (the wait ended with an acquire), and before the reply it would
             * sit inside the initiator's timed round trip */
            if (failed()) break;
        }
    }
}

__global__ void k_pingpong_process(int *flag_sym, int me, int peer, int initiator, int iters, int *payload_sym,
                                   int payload_len, long long *samples) {
    if (threadIdx.x != 0) return;
    pingpong_body(flag_sym, me, peer, initiator, iters, payload_sym, payload_len, samples);
}

/* Barrier latency: `iters` episodes of the offload barrier's protocol (the
 * count-and-generation word in PE 0's internal region), one lane per
 * workgroup, each workgroup one arrival: launch one workgroup per process in
 * process mode, one per PE in local mode.  samples[i] = wall-clock ticks of
 * episode i as seen by workgroup 0 of this launch. */
__global__ void k_barrier_latency(int iters, long long *samples) {
    if (threadIdx.x != 0) return;
    char *root = nvshmemi_dstate.heap_base[0];
    unsigned int *word = (unsigned int *)(root + NVSHMEMI_OFF_BAR_WORD);
    for (int i = 0; i < iters; ++i) {
        if (failed()) break;
        const long long t0 = (long long)wall_clock64();
        nvshmemi::gen_barrier(word, nvshmemi_dstate.npes, NVSHMEMI_ERR_BARRIER_TIMEOUT);
        if (blockIdx.x == 0 && samples) samples[i] = (long long)wall_clock64() - t0;
    }
}

/* One workgroup calls one strided / 2-D device transfer: op 0 iput_block,
 * 1 iget_block, 2 iput_wave (wavefront 0 only), 3 putmem_2d_block,
 * 4 getmem_2d_block, 5 iget_wave (wavefront 0 only); dtype 0 int, 1 float,
 * 2 double (ops 0-2 and 5). */
template <typename T>
__device__ void strided_op(int op, T *t, const T *s, ptrdiff_t tst, ptrdiff_t sst, size_t n, int pe);
template <>
__device__ void strided_op<int>(int op, int *t, const int *s, ptrdiff_t tst, ptrdiff_t sst, size_t n, int pe) {
    if (op == 0) nvshmemx_int_iput_block(t, s, tst, sst, n, pe);
    else if (op == 1) nvshmemx_int_iget_block(t, s, tst, sst, n, pe);
    else if (op == 2 && threadIdx.x < 64) nvshmemx_int_iput_wave(t, s, tst, sst, n, pe);
    else if (op == 5 && threadIdx.x < 64) nvshmemx_int_iget_wave(t, s, tst, sst, n, pe);
}
template <>
__device__ void strided_op<float>(int op, float *t, const float *s, ptrdiff_t tst, ptrdiff_t sst, size_t n, int pe) {
    if (op == 0) nvshmemx_float_iput_block(t, s, tst, sst, n, pe);
    else if (op == 1) nvshmemx_float_iget_block(t, s, tst, sst, n, pe);
    else if (op == 2 && threadIdx.x < 64) nvshmemx_float_iput_wave(t, s, tst, sst, n, pe);
    else if (op == 5 && threadIdx.x < 64) nvshmemx_float_iget_wave(t, s, tst, sst, n, pe);
}
template <>
__device__ void strided_op<double>(int op, double *t, const double *s, ptrdiff_t tst, ptrdiff_t sst, size_t n,
                                   int pe) {
    if (op == 0) nvshmemx_double_iput_block(t, s, tst, sst, n, pe);
    else if (op == 1) nvshmemx_double_iget_block(t, s, tst, sst, n, pe);
    else if (op == 2 && threadIdx.x < 64) nvshmemx_double_iput_wave(t, s, tst, sst, n, pe);
    else if (op == 5 && threadIdx.x < 64) nvshmemx_double_iget_wave(t, s, tst, sst, n, pe);
}

__global__ void k_strided_device(int op, int dtype, void *target, const void *source, long long tst, long long sst,
                                 size_t n, size_t tpitch, size_t spitch, size_t width, size_t height, int pe) {
    if (op == 3) {
        nvshmemx_putmem_2d_block(target, tpitch, source, spitch, width, height, pe);
    } else if (op == 4) {
        nvshmemx_getmem_2d_block(target, tpitch, source, spitch, width, height, pe);
    } else if (dtype == 0) {
        strided_op<int>(op, (int *)target, (const int *)source, tst, sst, n, pe);
    } else if (dtype == 1) {
        strided_op<float>(op, (float *)target, (const float *)source, tst, sst, n, pe);
    } else {
        strided_op<double>(op, (double *)target, (const double *)source, tst, sst, n, pe);
    }
}

/* local mode: workgroup 0 is PE a (initiator), workgroup 1 is PE b */
__global__ void k_pingpong_local(int *flag_sym, int pe_a, int pe_b, int iters, int *payload_sym, int payload_len,
                                 long long *samples) {
    if (threadIdx.x != 0) return;
    const int me = blockIdx.x == 0 ? pe_a : pe_b;
    const int peer = blockIdx.x == 0 ? pe_b : pe_a;
    pingpong_body(flag_sym, me, peer, blockIdx.x == 0, iters, payload_sym, payload_len, blockIdx.x == 0 ? samples : nullptr);
}

/* ---- ordering check: payload published by quiet + flag ----
 * Producer workgroup (PE a): round k: fill its local staging buffer with k,
 * block-put it into PE b's `data`, then one thread does quiet + p(flag, k, b);
 * then it waits for b's ack == k before reusing the buffer.
 * Consumer workgroup (PE b): waits for flag != k-1, reads every element of its
 * own `data`, counts elements != k, acks.  violations[0] = total mismatches,
 * violations[1] = flag values observed out of order (flag < k). */
__global__ void k_order_local(int *data_sym, int *stage_sym, int *flag_sym, int *ack_sym, int n, int rounds,
                              int pe_a, int pe_b, unsigned long long *violations) {
    const bool producer = blockIdx.x == 0;
    const int me = producer ? pe_a : pe_b;
    const int peer = producer ? pe_b : pe_a;
    int *mydata = (int *)nvshmemx_ptr_d(data_sym, me);
    int *mystage = (int *)nvshmemx_ptr_d(stage_sym, me);
    int *myflag = (int *)nvshmemx_ptr_d(flag_sym, me);
    int *myack = (int *)nvshmemx_ptr_d(ack_sym, me);
    __shared__ int seen;
    __shared__ int stop;
    for (int k = 1; k <= rounds; ++k) {
        if (threadIdx.x == 0) stop = failed();
        __syncthreads();
        if (stop) break;
        if (producer) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) mystage[i] = k;
            nvshmemi::fence_release();
            __syncthreads();
            nvshmemx_int_put_block(data_sym, mystage, n, peer);
            if (threadIdx.x == 0) {
                nvshmem_quiet();
                nvshmem_int_p(flag_sym, k, peer);
                nvshmem_quiet();
                nvshmem_int_wait(myack, k - 1);
            }
            __syncthreads();
        } else {
            if (threadIdx.x == 0) seen = nvshmemx_int_wait_until_cmp(myflag, NVSHMEMX_CMP_NE, k - 1);
            __syncthreads();
            if (threadIdx.x == 0 && seen < k) atomicAdd(&violations[1], 1ull);
            nvshmemi::fence_acquire();
            unsigned long long bad = 0;
            for (int i = threadIdx.x; i < n; i += blockDim.x) bad += (mydata[i] != k);
            if (bad) atomicAdd(&violations[0], bad);
            __syncthreads();
            if (threadIdx.x == 0) {
                nvshmem_int_p(ack_sym, k, peer);
                nvshmem_quiet();
            }
        }
    }
}

/* ---- workgroup-PE barrier check (persistent local-mode kernel) ----
 * Each PE (its first workgroup) writes round r into slot[me] of every PE with
 * p, then all workgroups meet in nvshmemx_barrier_all_wg; afterwards every
 * slot of the PE's own copy must read >= r.  violations counts slots < r. */
__global__ void k_wg_barrier_check(int *slots_sym, int rounds, unsigned long long *violations) {
    const int me = nvshmemx_my_pe_d();
    const int npes = nvshmemx_n_pes_d();
    const bool leader = (blockIdx.x & ((1u << nvshmemi_dstate.wg_shift) - 1u)) == 0;
    int *mine = (int *)nvshmemx_ptr_d(slots_sym, me);
    __shared__ int stop;
    for (int r = 1; r <= rounds; ++r) {
        if (threadIdx.x == 0) stop = failed();
        __syncthreads();
        if (stop) break;
        if (leader) {
            for (int p = threadIdx.x; p < npes; p += blockDim.x) nvshmem_int_p(slots_sym + me, r, p);
            nvshmem_quiet();
        }
        nvshmemx_barrier_all_wg();
        if (leader)
            for (int p = threadIdx.x; p < npes; p += blockDim.x)
                if (nvshmemi::load_relaxed(mine + p) < r) atomicAdd(violations, 1ull);
        nvshmemx_barrier_all_wg();
    }
}

/* ---- offload-barrier check across processes (one kernel per round step) ---- */
__global__ void k_publish_round(int *slots_sym, int me, int npes, int r) {
    for (int p = threadIdx.x; p < npes; p += blockDim.x) nvshmem_int_p(slots_sym + me, r, p);
    nvshmem_quiet();
}
__global__ void k_check_round(const int *slots_local, int npes, int r, unsigned long long *violations) {
    for (int p = threadIdx.x; p < npes; p += blockDim.x)
        if (nvshmemi::load_relaxed(slots_local + p) < r) atomicAdd(violations, 1ull);
}

/* one thread waits once: until == 0 uses nvshmem_int_wait (while equal),
 * until == 1 uses nvshmem_int_wait_until (until equal); out[0] = value seen */
__global__ void k_wait_once(int *ivar, int cmp_value, int until, int *out) {
    if (threadIdx.x != 0) return;
    if (until) nvshmem_int_wait_until(ivar, 0, cmp_value);
    else nvshmem_int_wait(ivar, cmp_value);
    out[0] = nvshmemi::load_relaxed(ivar);
}

/* traced sequence (one workgroup, PE `me` talks to `pe`): one put outside the
 * begin/end bracket (must not be recorded), then `reps` x {int put(16),
 * float put(8), double get(4), int p, int g, int wait, int wait_until}, then
 * one more put after trace_end (not recorded either) */
__global__ void k_traced_ops(int *ints_sym, float *floats_sym, double *doubles_sym, int *flag_sym, int me, int pe,
                             int reps) {
    int *mine = (int *)nvshmemx_ptr_d(ints_sym, me);
    int *myflag = (int *)nvshmemx_ptr_d(flag_sym, me);
    if (threadIdx.x == 0) nvshmem_int_put(ints_sym, mine, 4, pe);
    nvshmem_trace_begin();
    if (threadIdx.x == 0) {
        for (int r = 0; r < reps; ++r) {
            nvshmem_int_put(ints_sym, mine, 16, pe);
            nvshmem_float_put(floats_sym, (float *)mine, 8, pe);
            nvshmem_double_get((double *)nvshmemx_ptr_d(doubles_sym, me), doubles_sym, 4, pe);
            nvshmem_int_p(flag_sym + 1, r, pe);
            (void)nvshmem_int_g(flag_sym + 1, pe);
            nvshmem_quiet();
            nvshmem_int_wait(myflag, 12345);            /* *myflag == 7: returns at once */
            nvshmem_int_wait_until(myflag, 0, 7);
        }
    }
    nvshmem_trace_end();
    if (threadIdx.x == 0) nvshmem_int_put(ints_sym, mine, 4, pe);
}

/* traced element puts from many lanes at once: threads below `lanes` each
 * put one int (rep r, thread t: value r * lanes + t into element t) to PE
 * `pe`, `reps` times, between trace_begin and trace_end */
__global__ void k_traced_lanes(int *ints_sym, int pe, int lanes, int reps) {
    nvshmem_trace_begin();
    if ((int)threadIdx.x < lanes)
        for (int r = 0; r < reps; ++r) nvshmem_int_p(ints_sym + threadIdx.x, r * lanes + (int)threadIdx.x, pe);
    nvshmem_trace_end();
}

/* bench self-check: workgroup i compares request i's target bytes with its
 * source bytes (the copy is bitwise), counting mismatching bytes */
struct verify_item {
    const unsigned char *src;
    const unsigned char *dst;
    unsigned long long nbytes;
};
__global__ void k_verify(const verify_item *items, int n, unsigned long long *bad) {
    for (int r = blockIdx.x; r < n; r += gridDim.x) {
        const verify_item it = items[r];
        unsigned long long local = 0;
        for (unsigned long long b = threadIdx.x; b < it.nbytes; b += blockDim.x) local += it.src[b] != it.dst[b];
        if (local) atomicAdd(bad, local);
    }
}

/* HBM ceilings for the copy kernel's roofline: a read-only and a write-only
 * stream of 16-byte non-temporal accesses, one tile of 2 x 256 vectors per
 * workgroup (the large-copy kernel's shape).  A copy moves every byte once in
 * each direction, so it can run no faster than the two streams back to back. */
__global__ void __launch_bounds__(256) k_stream_read(const nvshmemi::u32x4 *a, size_t n, unsigned int *sink) {
    nvshmemi::u32x4 acc = {0u, 0u, 0u, 0u};
    const size_t base = (size_t)blockIdx.x * 512 + threadIdx.x;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) acc ^= __builtin_nontemporal_load(&a[i]);
    }
    /* never true for a buffer of one repeated word; keeps the loads live */
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u && acc.x == 0x7F4A7C15u) sink[0] = 1u;
}
__global__ void __launch_bounds__(256) k_stream_write(nvshmemi::u32x4 *a, size_t n) {
    const size_t base = (size_t)blockIdx.x * 512 + threadIdx.x;
    const nvshmemi::u32x4 v = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) __builtin_nontemporal_store(v, &a[i]);
    }
}
/* the same stream with the system-coherence store policy (sc1), which a
 * write-only stream on this part drains faster than non-temporal stores */
__global__ void __launch_bounds__(256) k_stream_write_sc1(nvshmemi::u32x4 *a, size_t n) {
    const size_t base = (size_t)blockIdx.x * 512 + threadIdx.x;
    const nvshmemi::u32x4 v = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) asm volatile("global_store_dwordx4 %0, %1, off sc1" : : "v"(&a[i]), "v"(v) : "memory");
    }
}

/* The library's copy as it runs for puts from 32 MiB up (tiles of 256
 * 16-byte vectors, one access per lane, non-temporal loads, sc1 stores; 16384
 * tiles make 64 MiB), with each workgroup's start and end stamped from the
 * constant clock, to see where the time of back-to-back launches goes. */
constexpr unsigned kStampGrid = 16384, kStampBlock = 256;
__global__ void __launch_bounds__(256) k_copy_stamped(char *dst, const char *src, size_t nbytes,
                                                      unsigned long long *stamps) {
    const unsigned long long t0 = wall_clock64();
    nvshmemi::copy_tiled<16, 1, nvshmemi::NT_LOAD | nvshmemi::ST_SC1>(dst, src, nbytes, 1, blockIdx.x, gridDim.x,
                                                                      threadIdx.x, blockDim.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        stamps[2 * (size_t)blockIdx.x] = t0;
        stamps[2 * (size_t)blockIdx.x + 1] = wall_clock64();
    }
}

/* Workgroup-partitioned PEs post puts with nvshmemx_putmem_nbi: workgroup w
 * belongs to the PE nvshmemx_my_pe_d() names, and the `wgs` workgroups of a
 * PE share the requests it issued (issuer[i] == its PE; a null issuer array
 * gives every request to the caller's PE).  A request with kind[i] != 0 is
 * posted instead as the get that moves the same bytes: by the target's PE,
 * from the source in PE issuer[i]'s heap (local mode). */
__global__ void k_post_nbi(const nvshmemx_put_request_t *reqs, const int *issuer, const int *kind, int n, int wgs) {
    const int pe = nvshmemx_my_pe_d();
    const int sub = (int)(blockIdx.x % (unsigned)wgs);
    for (int i = sub * (int)blockDim.x + (int)threadIdx.x; i < n; i += wgs * (int)blockDim.x) {
        const nvshmemx_put_request_t r = reqs[i];
        if (kind && kind[i]) { /* the same bytes fetched by the target's PE from the issuer's source */
            if (r.pe == pe) nvshmemx_getmem_nbi(nvshmemx_ptr_d(r.target, r.pe), r.source, r.nbytes, issuer[i]);
        } else if (!issuer || issuer[i] == pe) {
            nvshmemx_putmem_nbi(r.target, r.source, r.nbytes, r.pe);
        }
    }
}

/* one typed non-blocking put (get: a get from `pe`) posted by thread 0 of
 * workgroup 0 (PE 0 in local mode): dtype 0 int, 1 float, 2 double */
__global__ void k_post_nbi_typed(int dtype, int get, void *target, const void *source, int len, int pe) {
    if (threadIdx.x || blockIdx.x) return;
    if (get) {
        if (dtype == 0) nvshmemx_int_get_nbi((int *)target, (const int *)source, len, pe);
        else if (dtype == 1) nvshmemx_float_get_nbi((float *)target, (const float *)source, len, pe);
        else nvshmemx_double_get_nbi((double *)target, (const double *)source, len, pe);
    } else {
        if (dtype == 0) nvshmemx_int_put_nbi((int *)target, (const int *)source, len, pe);
        else if (dtype == 1) nvshmemx_float_put_nbi((float *)target, (const float *)source, len, pe);
        else nvshmemx_double_put_nbi((double *)target, (const double *)source, len, pe);
    }
}

} // namespace

/* --------------------------------------------------------- C entry points -- */

/* Counts bytes whose target differs from the source over all requests
 * (synchronous). */
extern "C" unsigned long long nvshmemh_verify_requests(const nvshmemx_put_request_t *reqs, int n) {
    std::vector<verify_item> items((size_t)n);
    for (int i = 0; i < n; ++i)
        items[(size_t)i] = verify_item{(const unsigned char *)reqs[i].source,
                                       (const unsigned char *)nvshmem_ptr(reqs[i].target, reqs[i].pe),
                                       (unsigned long long)reqs[i].nbytes};
    verify_item *d_items = nullptr;
    unsigned long long *d_bad = nullptr, bad = 0;
    NVSHMEMI_HIP(hipMalloc((void **)&d_items, std::max<size_t>(1, (size_t)n) * sizeof(verify_item)));
    NVSHMEMI_HIP(hipMalloc((void **)&d_bad, sizeof(bad)));
    NVSHMEMI_HIP(hipMemset(d_bad, 0, sizeof(bad)));
    if (n) NVSHMEMI_HIP(hipMemcpy(d_items, items.data(), (size_t)n * sizeof(verify_item), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_verify, dim3(2048), dim3(256), 0, 0, d_items, n, d_bad);
    NVSHMEMI_HIP(hipGetLastError());
    NVSHMEMI_HIP(hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost));
    NVSHMEMI_HIP(hipFree(d_items));
    NVSHMEMI_HIP(hipFree(d_bad));
    return bad;
}

extern "C" {

/* Read (write) `nbytes` (a multiple of 16, 16-byte aligned) as one streaming
 * kernel; the write stores zeros.  `sink` is one device word. */
void nvshmemh_stream_read(const void *p, size_t nbytes, unsigned int *sink, hipStream_t s) {
    if ((nbytes & 15) || ((uintptr_t)p & 15)) NVSHMEMI_FATAL("nvshmemh_stream_read: %zu bytes at %p", nbytes, p);
    const size_t n = nbytes / 16;
    if (!n) return;
    hipLaunchKernelGGL(k_stream_read, dim3((unsigned)((n + 511) / 512)), dim3(256), 0, s, (const nvshmemi::u32x4 *)p, n, sink);
    NVSHMEMI_HIP(hipGetLastError());
}
/* policy 0: non-temporal stores; 1: sc1 stores */
void nvshmemh_stream_write(void *p, size_t nbytes, int policy, hipStream_t s) {
    if ((nbytes & 15) || ((uintptr_t)p & 15)) NVSHMEMI_FATAL("nvshmemh_stream_write: %zu bytes at %p", nbytes, p);
    if (policy != 0 && policy != 1) NVSHMEMI_FATAL("nvshmemh_stream_write: store policy %d", policy);
    const size_t n = nbytes / 16;
    if (!n) return;
    const dim3 grid((unsigned)((n + 511) / 512));
    if (policy == 0) hipLaunchKernelGGL(k_stream_write, grid, dim3(256), 0, s, (nvshmemi::u32x4 *)p, n);
    else hipLaunchKernelGGL(k_stream_write_sc1, grid, dim3(256), 0, s, (nvshmemi::u32x4 *)p, n);
    NVSHMEMI_HIP(hipGetLastError());
}

/* `launches` back-to-back stamped copies of `nbytes` on stream s; launch k
 * copies src + o to dst + o, o = (k * nbytes) mod span (both buffers hold
 * `span` bytes, a multiple of nbytes).  Workgroup b of launch k writes its
 * start and end (constant-clock ticks) to stamps[(k * grid + b) * 2 + {0, 1}];
 * `stamp_words` is the capacity.  Returns the grid (workgroups per launch). */
int nvshmemh_copy_stamped(void *dst, const void *src, size_t nbytes, size_t span, int launches,
                          unsigned long long *stamps, size_t stamp_words, hipStream_t s) {
    if (!nbytes || span < nbytes || span % nbytes || launches < 1)
        NVSHMEMI_FATAL("nvshmemh_copy_stamped: %zu bytes over a %zu-byte span, %d launches", nbytes, span, launches);
    if ((size_t)launches * kStampGrid * 2 > stamp_words)
        NVSHMEMI_FATAL("nvshmemh_copy_stamped: %d launches need %zu stamp words, %zu given", launches,
                       (size_t)launches * kStampGrid * 2, stamp_words);
    for (int k = 0; k < launches; ++k) {
        const size_t o = ((size_t)k * nbytes) % span;
        hipLaunchKernelGGL(k_copy_stamped, dim3(kStampGrid), dim3(kStampBlock), 0, s, (char *)dst + o,
                           (const char *)src + o, nbytes, stamps + (size_t)k * kStampGrid * 2);
    }
    NVSHMEMI_HIP(hipGetLastError());
    return (int)kStampGrid;
}

void nvshmemh_fill_pcg32(void *dst, size_t nwords, unsigned long long seed, unsigned long long seq, hipStream_t s) {
    if (nwords == 0) return;
    uint64_t inc = 0;
    const uint64_t s0 = pcg_seeded_state(seed, seq, inc);
    /* one 64-step jump as an affine map: state' = m64 * state + a64 */
    const uint64_t m64 = pcg_advance(1, 64, 0) /* multiplier only */;
    const uint64_t a64 = pcg_advance(0, 64, inc);
    const size_t words_per_wave = 64 * (size_t)kFillWordsPerLane;
    const size_t waves = (nwords + words_per_wave - 1) / words_per_wave;
    const size_t blocks = (waves * 64 + 255) / 256;
    hipLaunchKernelGGL(k_fill_pcg32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)dst, nwords, s0, inc, m64, a64);
    NVSHMEMI_HIP(hipGetLastError());
}

/* host-side generator with the same definition (used by tests to pin the
 * device generator against the oracle) */
void nvshmemh_pcg32_host(uint32_t *out, size_t nwords, unsigned long long seed, unsigned long long seq) {
    uint64_t inc = 0;
    uint64_t st = pcg_seeded_state(seed, seq, inc);
    for (size_t i = 0; i < nwords; ++i) {
        out[i] = pcg_output(st);
        st = st * kPcgMult + inc;
    }
}

/* The uneven-put workload of the batch benchmark, drawn from one PCG32
 * stream (seed, stream 54), five draws per request in this order: size
 * (log-uniform on [2^min_log2, 2^max_log2], rounded down to 4 bytes), element
 * type (int32 / f32), source PE, destination PE (uniform among the others),
 * 256-byte aligned source offset within `src_region`.  Destinations are packed
 * in request order per destination PE, 256-byte aligned, so they never
 * overlap.  Returns 0, or -1 when a destination region would overflow. */
int nvshmemh_gen_requests(unsigned long long seed, int nreq, int npes, unsigned long long src_region,
                          unsigned long long dst_region, int min_log2, int max_log2, int *src_pe, int *dst_pe,
                          unsigned long long *src_off, unsigned long long *dst_off, unsigned long long *nbytes,
                          int *dtype) {
    if (npes < 2 || npes > 1024) return -2;
    uint64_t inc = 0;
    uint64_t st = pcg_seeded_state(seed, 54, inc);
    std::vector<unsigned long long> cursor((size_t)npes, 0ull);
    for (int i = 0; i < nreq; ++i) {
        uint32_t draw[5];
        for (uint32_t &d : draw) {
            d = pcg_output(st);
            st = st * kPcgMult + inc;
        }
        const double e = (double)min_log2 + (double)(max_log2 - min_log2) * ((double)draw[0] / 4294967296.0);
        unsigned long long size = ((unsigned long long)exp2(e)) & ~3ull;
        if (size < 4) size = 4;
        if (size > src_region) return -1;
        const int s = (int)(draw[2] % (uint32_t)npes);
        const int d = (s + 1 + (int)(draw[3] % (uint32_t)(npes - 1))) % npes;
        src_pe[i] = s;
        dst_pe[i] = d;
        dtype[i] = (int)(draw[1] & 1u);
        nbytes[i] = size;
        src_off[i] = (unsigned long long)(draw[4] % ((src_region - size) / 256 + 1)) * 256ull;
        dst_off[i] = cursor[(size_t)d];
        cursor[(size_t)d] = (cursor[(size_t)d] + size + 255) / 256 * 256;
        if (cursor[(size_t)d] > dst_region) return -1;
    }
    return 0;
}

/* dtype: 0 int, 1 float, 2 double; op: 0 put, 1 get */
void nvshmemh_thread_putget(int op, int dtype, void *target, void *source, int len, int chunk, int pe, hipStream_t s) {
    if (len <= 0) return;
    if (chunk <= 0) chunk = len;
    const int threads = (len + chunk - 1) / chunk;
    const int block = threads < 256 ? threads : 256;
    const int grid = (threads + block - 1) / block;
    if (dtype == 0)
        hipLaunchKernelGGL(k_thread_putget<int>, dim3(grid), dim3(block), 0, s, op, (int *)target, (int *)source, len, chunk, pe);
    else if (dtype == 1)
        hipLaunchKernelGGL(k_thread_putget<float>, dim3(grid), dim3(block), 0, s, op, (float *)target, (float *)source, len, chunk, pe);
    else
        hipLaunchKernelGGL(k_thread_putget<double>, dim3(grid), dim3(block), 0, s, op, (double *)target, (double *)source, len, chunk, pe);
    NVSHMEMI_HIP(hipGetLastError());
}

/* scope: 0 wave, 1 block; each wave / workgroup copies `slice` bytes */
void nvshmemh_coop_putget(int op, int scope, void *target, const void *source, size_t nbytes, size_t slice, int pe,
                          int block, hipStream_t s) {
    if (nbytes == 0) return;
    if (slice == 0) slice = nbytes;
    const size_t units = (nbytes + slice - 1) / slice;
    const size_t per_block = scope == 0 ? (size_t)block / 64 : 1;
    const size_t grid = (units + per_block - 1) / per_block;
    hipLaunchKernelGGL(k_coop_putget, dim3((unsigned)grid), dim3(block), 0, s, op, scope, (char *)target,
                       (const char *)source, nbytes, slice, pe);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_pg(int op, int dtype, void *target, const void *values, void *out, int n, int pe, hipStream_t s) {
    if (n <= 0) return;
    const int block = 256, grid = (n + block - 1) / block;
    if (dtype == 0)
        hipLaunchKernelGGL(k_pg<int>, dim3(grid), dim3(block), 0, s, op, (int *)target, (const int *)values, (int *)out, n, pe);
    else if (dtype == 1)
        hipLaunchKernelGGL(k_pg<float>, dim3(grid), dim3(block), 0, s, op, (float *)target, (const float *)values, (float *)out, n, pe);
    else
        hipLaunchKernelGGL(k_pg<double>, dim3(grid), dim3(block), 0, s, op, (double *)target, (const double *)values, (double *)out, n, pe);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_pingpong_process(int *flag_sym, int peer, int initiator, int iters, int *payload_sym, int payload_len,
                               long long *samples, hipStream_t s) {
    hipLaunchKernelGGL(k_pingpong_process, dim3(1), dim3(64), 0, s, flag_sym, nvshmem_my_pe(), peer, initiator, iters,
                       payload_sym, payload_len, samples);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_strided_device(int op, int dtype, void *target, const void *source, long long tst, long long sst,
                             size_t n, size_t tpitch, size_t spitch, size_t width, size_t height, int pe, int block,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_strided_device, dim3(1), dim3(block), 0, s, op, dtype, target, source, tst, sst, n, tpitch,
                       spitch, width, height, pe);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_barrier_latency(int iters, long long *samples, hipStream_t s) {
    if (nvshmemx_mode() == NVSHMEMX_MODE_LOCAL) {
        /* every arrival must be resident at once: one workgroup per local PE */
        void *args[] = {&iters, &samples};
        nvshmemx_collective_launch((const void *)k_barrier_latency, dim3(nvshmem_n_pes()), dim3(64), args, 0, s);
        return;
    }
    /* one arrival per process, a plain launch (cooperative launches are
     * exclusive to one process of the device at a time) */
    hipLaunchKernelGGL(k_barrier_latency, dim3(1), dim3(64), 0, s, iters, samples);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_pingpong_local(int *flag_sym, int pe_a, int pe_b, int iters, int *payload_sym, int payload_len,
                             long long *samples, hipStream_t s) {
    hipLaunchKernelGGL(k_pingpong_local, dim3(2), dim3(64), 0, s, flag_sym, pe_a, pe_b, iters, payload_sym,
                       payload_len, samples);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_order_local(int *data_sym, int *stage_sym, int *flag_sym, int *ack_sym, int n, int rounds, int pe_a,
                          int pe_b, unsigned long long *violations, hipStream_t s) {
    hipLaunchKernelGGL(k_order_local, dim3(2), dim3(256), 0, s, data_sym, stage_sym, flag_sym, ack_sym, n, rounds,
                       pe_a, pe_b, violations);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_wg_barrier_check(int *slots_sym, int rounds, unsigned long long *violations, int block, hipStream_t s) {
    const int npes = nvshmem_n_pes();
    const int grid = npes * nvshmemx_wgs_per_pe();
    void *args[] = {&slots_sym, &rounds, &violations};
    nvshmemx_collective_launch((const void *)k_wg_barrier_check, dim3(grid), dim3(block), args, 0, s);
}

void nvshmemh_wait_once(int *ivar, int cmp_value, int until, int *out, hipStream_t s) {
    hipLaunchKernelGGL(k_wait_once, dim3(1), dim3(64), 0, s, ivar, cmp_value, until, out);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_traced_ops(int *ints_sym, float *floats_sym, double *doubles_sym, int *flag_sym, int me, int pe,
                         int reps, int block, int grid, hipStream_t s) {
    /* `grid` workgroups each issue the sequence from PE me's buffers (in local
     * mode workgroups 0 .. wgs_per_pe - 1 are PE 0, so callers pass me = 0) */
    if (grid < 1) NVSHMEMI_FATAL("nvshmemh_traced_ops: grid %d", grid);
    hipLaunchKernelGGL(k_traced_ops, dim3((unsigned)grid), dim3(block), 0, s, ints_sym, floats_sym, doubles_sym,
                       flag_sym, me, pe, reps);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_traced_lanes(int *ints_sym, int pe, int lanes, int reps, int block, int grid, hipStream_t s) {
    if (grid < 1 || block < 1 || block > 1024 || lanes < 0 || lanes > block)
        NVSHMEMI_FATAL("nvshmemh_traced_lanes: grid %d block %d lanes %d", grid, block, lanes);
    hipLaunchKernelGGL(k_traced_lanes, dim3((unsigned)grid), dim3((unsigned)block), 0, s, ints_sym, pe, lanes, reps);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_publish_round(int *slots_sym, int r, hipStream_t s) {
    hipLaunchKernelGGL(k_publish_round, dim3(1), dim3(128), 0, s, slots_sym, nvshmem_my_pe(), nvshmem_n_pes(), r);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_check_round(const int *slots_local, int r, unsigned long long *violations, hipStream_t s) {
    hipLaunchKernelGGL(k_check_round, dim3(1), dim3(128), 0, s, slots_local, nvshmem_n_pes(), r, violations);
    NVSHMEMI_HIP(hipGetLastError());
}

/* Post `n` requests (device array) from the PEs that issued them: in local
 * mode issuer[i] (device array) names request i's PE and each PE posts from
 * its own workgroups; in process mode this PE posts every request (issuer
 * may be null).  256 threads a workgroup. */
void nvshmemh_post_nbi(const nvshmemx_put_request_t *reqs, const int *issuer, int n, hipStream_t s) {
    if (n <= 0) return;
    const bool local = nvshmemx_mode() == NVSHMEMI_MODE_LOCAL;
    if (local && !issuer) NVSHMEMI_FATAL("nvshmemh_post_nbi: local mode needs the issuing PE of every request");
    const int wgs = local ? nvshmemx_wgs_per_pe() : 4;
    const int grid = local ? nvshmem_n_pes() * wgs : wgs;
    hipLaunchKernelGGL(k_post_nbi, dim3((unsigned)grid), dim3(256), 0, s, reqs, local ? issuer : nullptr,
                       (const int *)nullptr, n, wgs);
    NVSHMEMI_HIP(hipGetLastError());
}

/* The same in local mode, request i posted as a get by its target's PE where
 * kind[i] != 0 (device array), as a put by issuer[i] otherwise. */
void nvshmemh_post_nbi_mixed(const nvshmemx_put_request_t *reqs, const int *issuer, const int *kind, int n,
                             hipStream_t s) {
    if (n <= 0) return;
    if (nvshmemx_mode() != NVSHMEMI_MODE_LOCAL || !issuer || !kind)
        NVSHMEMI_FATAL("nvshmemh_post_nbi_mixed: local mode, with the issuing PE and kind of every request");
    const int wgs = nvshmemx_wgs_per_pe();
    hipLaunchKernelGGL(k_post_nbi, dim3((unsigned)(nvshmem_n_pes() * wgs)), dim3(256), 0, s, reqs, issuer, kind, n,
                       wgs);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_post_nbi_typed(int dtype, int get, void *target, const void *source, int len, int pe, hipStream_t s) {
    hipLaunchKernelGGL(k_post_nbi_typed, dim3(1), dim3(64), 0, s, dtype, get, target, source, len, pe);
    NVSHMEMI_HIP(hipGetLastError());
}

} /* extern "C" */
