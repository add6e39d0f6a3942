/*
 * oracle.c -- CPU restatement of the reference semantics (see oracle.h).
 * TEST INFRASTRUCTURE ONLY; never linked into the product.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

/* ------------------------------------------------------------- PCG32 -- */
/* O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
 * Algorithms for Random Number Generation" (2014); minimal C reference
 * pcg32_srandom_r / pcg32_random_r.  Not part of the reference repository:
 * it generates the seeded workloads its configurations describe. */
#define PCG_MULT 6364136223846793005ULL

static uint32_t pcg_out(uint64_t old) {
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

void oracle_pcg32_fill(uint32_t *out, size_t n, uint64_t seed, uint64_t seq) {
    uint64_t inc = (seq << 1u) | 1u, st = 0;
    st = st * PCG_MULT + inc;
    st += seed;
    st = st * PCG_MULT + inc;
    for (size_t i = 0; i < n; ++i) {
        out[i] = pcg_out(st);
        st = st * PCG_MULT + inc;
    }
}

/* words offset .. offset+n-1 of the same stream, by O'Neill's jump-ahead
 * (log2(offset) squarings of the affine step) */
void oracle_pcg32_fill_at(uint32_t *out, size_t n, uint64_t seed, uint64_t seq, uint64_t offset) {
    uint64_t inc = (seq << 1u) | 1u, st = 0;
    st = st * PCG_MULT + inc;
    st += seed;
    st = st * PCG_MULT + inc;
    uint64_t acc_mult = 1, acc_plus = 0, cur_mult = PCG_MULT, cur_plus = inc;
    for (uint64_t d = offset; d; d >>= 1) {
        if (d & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
    }
    st = acc_mult * st + acc_plus;
    for (size_t i = 0; i < n; ++i) {
        out[i] = pcg_out(st);
        st = st * PCG_MULT + inc;
    }
}

void oracle_pcg32_uniform_pm1(float *out, size_t n, uint64_t seed, uint64_t seq) {
    uint32_t *w = (uint32_t *)out;
    oracle_pcg32_fill(w, n, seed, seq);
    for (size_t i = 0; i < n; ++i) {
        uint32_t u = w[i] >> 8; /* 24 bits */
        out[i] = (float)u * (1.0f / 8388608.0f) - 1.0f; /* u * 2^-23 - 1: exact */
    }
}

void oracle_pcg32_uniform01_f64(double *out, size_t n, uint64_t seed, uint64_t seq) {
    uint32_t *w = (uint32_t *)out; /* 2n words fit exactly in n doubles */
    oracle_pcg32_fill(w, 2 * n, seed, seq);
    for (size_t i = 0; i < n; ++i) {
        uint64_t hi = w[2 * i], lo = w[2 * i + 1];
        uint64_t u = ((hi << 32) | lo) >> 11; /* 53 bits */
        out[i] = (double)u * (1.0 / 9007199254740992.0);
    }
}

/* -------------------------------------------------------------- heap -- */
/* reference nvshmem.cu:775-802 (nvshmalloc), 805-819 (nvshmcleanup) */

void oracle_heap_init(oracle_heap_t *h, size_t heap_size, size_t align) {
    h->heap_size = heap_size;
    h->tail = 0;
    h->align = align ? align : 1;
    h->nregions = 0;
}

long long oracle_nvshmalloc(oracle_heap_t *h, size_t size) {
    size_t off = (h->tail + h->align - 1) / h->align * h->align;
    if (off > h->heap_size || size > h->heap_size - off) return -1; /* nvshmem.cu:780-783 */
    h->tail = off + size;                                          /* nvshmem.cu:785-786 */
    h->nregions++;                                                 /* nvshmem.cu:788-799 */
    return (long long)off;
}

void oracle_nvshmcleanup(oracle_heap_t *h) {
    h->tail = 0; /* nvshmem.cu:818 */
    h->nregions = 0;
}

/* ------------------------------------------------------- translation -- */
/* reference nvshmem.cu:882: heap_pointers[pe] + (local - my_heap).  With
 * mask = all ones and anchor = my heap base this is that formula; the local
 * (single-process) layout uses anchor = PE 0's base and mask = stride - 1. */
uint64_t oracle_translate(const uint64_t *heap_base, uint64_t anchor, uint64_t mask, uint64_t addr, int pe) {
    return heap_base[pe] + ((addr - anchor) & mask);
}

/* ----------------------------------------------------- data movement -- */

void oracle_put(uint8_t **heaps, size_t heap_bytes, size_t target_off, const void *source, size_t nbytes, int pe) {
    /* reference nvshmem_device.cu:494-500: remote[i] = source[i], i < len */
    if (target_off > heap_bytes || nbytes > heap_bytes - target_off) return;
    memcpy(heaps[pe] + target_off, source, nbytes);
}

void oracle_get(uint8_t **heaps, size_t heap_bytes, void *target, size_t source_off, size_t nbytes, int pe) {
    /* reference nvshmem_device.cu:521-527: target[i] = remote[i], i < len */
    if (source_off > heap_bytes || nbytes > heap_bytes - source_off) return;
    memcpy(target, heaps[pe] + source_off, nbytes);
}

void oracle_p32(uint8_t **heaps, size_t target_off, uint32_t bits, int pe) {
    memcpy(heaps[pe] + target_off, &bits, 4); /* nvshmem_device.cu:445-448 */
}
void oracle_p64(uint8_t **heaps, size_t target_off, uint64_t bits, int pe) {
    memcpy(heaps[pe] + target_off, &bits, 8);
}
uint32_t oracle_g32(uint8_t *const *heaps, size_t source_off, int pe) {
    uint32_t v;
    memcpy(&v, heaps[pe] + source_off, 4); /* nvshmem_device.cu:469-472 */
    return v;
}
uint64_t oracle_g64(uint8_t *const *heaps, size_t source_off, int pe) {
    uint64_t v;
    memcpy(&v, heaps[pe] + source_off, 8);
    return v;
}

/* ------------------------------------------------------------- waits -- */

int oracle_int_wait_done(int observed, int cmp_value) {
    return observed != cmp_value; /* nvshmem_device.cu:551: spin while == */
}

int oracle_int_wait_until_done(int observed, int cmp, int cmp_value) {
    (void)cmp;                     /* README:170: "ignored currently" */
    return observed == cmp_value;  /* nvshmem_device.cu:573: spin while != */
}

int oracle_cmp_holds(int v, int op, int ref) {
    switch (op) {
    case 0: return v == ref;
    case 1: return v != ref;
    case 2: return v > ref;
    case 3: return v >= ref;
    case 4: return v < ref;
    case 5: return v <= ref;
    default: return 0;
    }
}

/* ----------------------------------------------------------- barrier -- */

int oracle_ref_barrier_value(long long k) {
    /* nvshmem.cu:139 starts at 1; :877 advances mod 64*1024 */
    return (int)((1 + k) % 65536);
}
int oracle_ref_barrier_half(long long k) {
    return oracle_ref_barrier_value(k) % 2; /* nvshmem.cu:873 */
}

void oracle_gen_barrier_init(oracle_gen_barrier_t *b, int npes, uint32_t generation) {
    b->word = (generation & 0xFFFFu) << 16;
    b->npes = npes;
}
int oracle_gen_barrier_arrive(oracle_gen_barrier_t *b) {
    const uint32_t old = b->word;
    b->word = old + 1u; /* the fetch-add */
    const uint32_t g0 = old >> 16;
    if ((old & 0xFFFFu) == (uint32_t)(b->npes - 1)) b->word = ((g0 + 1u) & 0xFFFFu) << 16;
    return (int)g0;
}
int oracle_gen_barrier_released(const oracle_gen_barrier_t *b, int gen_seen) {
    return (int)(b->word >> 16) != gen_seen;
}

/* ------------------------------------------------------------- trace -- */

static const char *const kOps[8] = {"put_int", "put_float", "put_double", "get_int",
                                    "get_float", "get_double", "wait_int", "wait_until_int"}; /* nvshmem.cu:117 */

const char *oracle_trace_header(void) {
    /* nvshmem.cu:390 */
    return "# timestamp, gridid, blockid, tid, type, local_address, remote_address, size, remote_pe, timespent\n";
}
const char *oracle_trace_op_name(int type) { return (type >= 0 && type < 8) ? kOps[type] : "unknown"; }

int oracle_trace_row(char *buf, size_t cap, long long tstamp_ns, unsigned long long gridid, int blockid, int tid,
                     int type, uint64_t local_addr, uint64_t remote_addr, int size, int remote_pe,
                     long long tspent_ns) {
    /* nvshmem.cu:411-421 (block-level format) */
    return snprintf(buf, cap, "%lld, %llu, %d, %d, %s, %p, %p, %d, %d, %lld\n", tstamp_ns, gridid, blockid, tid,
                    oracle_trace_op_name(type), (void *)(uintptr_t)local_addr, (void *)(uintptr_t)remote_addr, size,
                    remote_pe, tspent_ns);
}

/* ---------------------------------------------------- batch workload -- */
/* The 4096-put workload: sizes log-uniform on [2^min_log2, 2^max_log2]
 * rounded down to a multiple of 4 bytes, int32/f32 50/50, (src, dst) PE pairs
 * uniform with src != dst, 256-byte aligned source offsets inside the source
 * region, destinations packed in request order per destination PE (so they
 * are disjoint and the result is order-independent).  One PCG32 stream
 * (seed, seq 54) drives every draw, in the order written below. */
int oracle_batch_generate(uint64_t seed, int nreq, int npes, uint64_t src_region, uint64_t dst_region, int min_log2,
                          int max_log2, int *src_pe, int *dst_pe, uint64_t *src_off, uint64_t *dst_off,
                          uint64_t *nbytes, int *dtype) {
    uint64_t inc = (54u << 1u) | 1u, st = 0;
    st = st * PCG_MULT + inc;
    st += seed;
    st = st * PCG_MULT + inc;
    uint64_t cursor[1024];
    if (npes < 2 || npes > 1024) return -2;
    for (int p = 0; p < npes; ++p) cursor[p] = 0;
    for (int i = 0; i < nreq; ++i) {
        uint32_t u[5];
        for (int k = 0; k < 5; ++k) {
            u[k] = pcg_out(st);
            st = st * PCG_MULT + inc;
        }
        const double e = (double)min_log2 + (double)(max_log2 - min_log2) * ((double)u[0] / 4294967296.0);
        uint64_t size = ((uint64_t)exp2(e)) & ~(uint64_t)3;
        if (size < 4) size = 4;
        const int s = (int)(u[2] % (uint32_t)npes);
        const int d = (int)((s + 1 + (int)(u[3] % (uint32_t)(npes - 1))) % npes);
        if (size > src_region) return -1;
        const uint64_t slots = (src_region - size) / 256 + 1;
        src_pe[i] = s;
        dst_pe[i] = d;
        dtype[i] = (int)(u[1] & 1u);
        nbytes[i] = size;
        src_off[i] = (u[4] % slots) * 256;
        dst_off[i] = cursor[d];
        cursor[d] = (cursor[d] + size + 255) / 256 * 256;
        if (cursor[d] > dst_region) return -1;
    }
    return 0;
}

/* applies the batch to host heap images in request order: heaps[pe] is PE
 * pe's image; sources at src_base + src_off, targets at dst_base + dst_off */
void oracle_batch_apply(uint8_t **heaps, int nreq, const int *src_pe, const int *dst_pe, const uint64_t *src_off,
                        const uint64_t *dst_off, const uint64_t *nbytes, uint64_t src_base, uint64_t dst_base) {
    for (int i = 0; i < nreq; ++i)
        memcpy(heaps[dst_pe[i]] + dst_base + dst_off[i], heaps[src_pe[i]] + src_base + src_off[i], nbytes[i]);
}

/* ---- strided and 2-D transfers (OpenSHMEM iput/iget semantics; the
 * reference has contiguous put/get only, nvshmem_device.cu:487-540) ---- */
void oracle_iput(uint8_t **heaps, size_t target_off, ptrdiff_t tst, const void *source, ptrdiff_t sst, size_t nelems,
                 size_t elem_bytes, int pe) {
    const uint8_t *src = (const uint8_t *)source;
    for (size_t i = 0; i < nelems; ++i)
        memcpy(heaps[pe] + target_off + (size_t)((ptrdiff_t)i * tst) * elem_bytes,
               src + (size_t)((ptrdiff_t)i * sst) * elem_bytes, elem_bytes);
}

void oracle_iget(uint8_t *const *heaps, void *target, ptrdiff_t tst, size_t source_off, ptrdiff_t sst, size_t nelems,
                 size_t elem_bytes, int pe) {
    uint8_t *dst = (uint8_t *)target;
    for (size_t i = 0; i < nelems; ++i)
        memcpy(dst + (size_t)((ptrdiff_t)i * tst) * elem_bytes,
               heaps[pe] + source_off + (size_t)((ptrdiff_t)i * sst) * elem_bytes, elem_bytes);
}

void oracle_put_2d(uint8_t **heaps, size_t target_off, size_t tpitch, const void *source, size_t spitch, size_t width,
                   size_t height, int pe) {
    for (size_t r = 0; r < height; ++r)
        memcpy(heaps[pe] + target_off + r * tpitch, (const uint8_t *)source + r * spitch, width);
}
