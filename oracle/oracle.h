/*
 * oracle.h -- CPU restatement of the reference's semantics for the
 * device-initiated OpenSHMEM subset.  TEST INFRASTRUCTURE ONLY: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark use it as the checker.
 * The product (df-nvshmem-prototype_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The reference ships no golden vectors (it has no tests at all); each
 * function's pinning check is listed in tests/test_oracle.py.
 */
#ifndef NVSHMEM_ORACLE_H
#define NVSHMEM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded data (workload generator; PCG32 minimal C reference) ---- */
void oracle_pcg32_fill(uint32_t *out, size_t n, uint64_t seed, uint64_t seq);
/* words offset .. offset+n-1 of the stream oracle_pcg32_fill starts at 0 */
void oracle_pcg32_fill_at(uint32_t *out, size_t n, uint64_t seed, uint64_t seq, uint64_t offset);
/* u32 -> (u >> 8) * 2^-24 * 2 - 1, exactly representable in f32, in [-1, 1) */
void oracle_pcg32_uniform_pm1(float *out, size_t n, uint64_t seed, uint64_t seq);
/* u32 pairs -> (u64 >> 11) * 2^-53, exactly representable in f64, in [0, 1) */
void oracle_pcg32_uniform01_f64(double *out, size_t n, uint64_t seed, uint64_t seq);

/* ---- symmetric heap (reference nvshmem.cu:775-819) ---- */
typedef struct oracle_heap {
    size_t heap_size; /* user bytes */
    size_t tail;      /* next free offset */
    size_t align;     /* 1 restates the reference exactly; 256 is this build */
    size_t nregions;
} oracle_heap_t;
void oracle_heap_init(oracle_heap_t *h, size_t heap_size, size_t align);
/* returns the offset of the new block, or -1 where the reference exits
 * (tail + size > heap, nvshmem.cu:780-783) */
long long oracle_nvshmalloc(oracle_heap_t *h, size_t size);
void oracle_nvshmcleanup(oracle_heap_t *h);

/* ---- translation (reference nvshmem.cu:880-883, nvshmem_device.cu:445-446) ---- */
uint64_t oracle_translate(const uint64_t *heap_base, uint64_t anchor, uint64_t mask, uint64_t addr, int pe);

/* ---- data movement on PE-indexed host heaps (nvshmem_device.cu:438-540) ----
 * heaps[pe] points at PE pe's heap image (heap_bytes each).  Offsets are
 * relative to the heap base.  Copies are bytewise: element type only sets the
 * length unit (reference put/get differ only in element width). */
void oracle_put(uint8_t **heaps, size_t heap_bytes, size_t target_off, const void *source, size_t nbytes, int pe);
void oracle_get(uint8_t **heaps, size_t heap_bytes, void *target, size_t source_off, size_t nbytes, int pe);
void oracle_p32(uint8_t **heaps, size_t target_off, uint32_t value_bits, int pe);
void oracle_p64(uint8_t **heaps, size_t target_off, uint64_t value_bits, int pe);
uint32_t oracle_g32(uint8_t *const *heaps, size_t source_off, int pe);
uint64_t oracle_g64(uint8_t *const *heaps, size_t source_off, int pe);

/* ---- strided and 2-D transfers (OpenSHMEM iput/iget; strides in elements) ---- */
void oracle_iput(uint8_t **heaps, size_t target_off, ptrdiff_t tst, const void *source, ptrdiff_t sst, size_t nelems,
                 size_t elem_bytes, int pe);
void oracle_iget(uint8_t *const *heaps, void *target, ptrdiff_t tst, size_t source_off, ptrdiff_t sst, size_t nelems,
                 size_t elem_bytes, int pe);
void oracle_put_2d(uint8_t **heaps, size_t target_off, size_t tpitch, const void *source, size_t spitch, size_t width,
                   size_t height, int pe);

/* ---- waits (reference nvshmem_device.cu:542-584, README:157-174) ----
 * 1 when the wait returns for the observed value */
int oracle_int_wait_done(int observed, int cmp_value);              /* returns once observed != cmp_value */
int oracle_int_wait_until_done(int observed, int cmp, int cmp_value); /* returns once observed == cmp_value; cmp ignored */
int oracle_cmp_holds(int observed, int op, int ref);                 /* OpenSHMEM comparison set (extension) */

/* ---- offload barrier (reference nvshmem.cu:869-878, nvshmem_kernels.cu:56-102) ---- */
/* value used by the k-th call (k = 0, 1, ...): starts at 1, wraps mod 65536 */
int oracle_ref_barrier_value(long long k);
/* sync-array half used by the k-th call: value % 2 */
int oracle_ref_barrier_half(long long k);

/* Count-and-generation barrier of this build, as a state machine over its one
 * 32-bit word (arrival count in the low 16 bits, generation in the high 16):
 * arrive() is the fetch-add, returns the generation it observed, and on the
 * last arrival stores the next generation with a zero count; released() is the
 * waiters' poll.  Used to check exhaustively over arrival orders that no PE is
 * released early. */
typedef struct oracle_gen_barrier {
    uint32_t word;
    int npes;
} oracle_gen_barrier_t;
void oracle_gen_barrier_init(oracle_gen_barrier_t *b, int npes, uint32_t generation);
int oracle_gen_barrier_arrive(oracle_gen_barrier_t *b);
int oracle_gen_barrier_released(const oracle_gen_barrier_t *b, int gen_seen);

/* ---- trace CSV (reference nvshmem.cu:117, 390, 411) ---- */
const char *oracle_trace_header(void);
const char *oracle_trace_op_name(int type);
/* formats one row exactly like the reference's block-level sprintf */
int oracle_trace_row(char *buf, size_t cap, long long tstamp_ns, unsigned long long gridid, int blockid, int tid,
                     int type, uint64_t local_addr, uint64_t remote_addr, int size, int remote_pe,
                     long long tspent_ns);

/* ---- the 4096-put batch workload (see oracle.c) ---- */
int oracle_batch_generate(uint64_t seed, int nreq, int npes, uint64_t src_region, uint64_t dst_region, int min_log2,
                          int max_log2, int *src_pe, int *dst_pe, uint64_t *src_off, uint64_t *dst_off,
                          uint64_t *nbytes, int *dtype);
void oracle_batch_apply(uint8_t **heaps, int nreq, const int *src_pe, const int *dst_pe, const uint64_t *src_off,
                        const uint64_t *dst_off, const uint64_t *nbytes, uint64_t src_base, uint64_t dst_base);

/* ---- Jacobi host replay (oracle_jacobi.c) ---- */
void oracle_jacobi_init(double *u, int G, uint64_t seed);
void oracle_jacobi_step(const double *u, double *v, int G);
int oracle_jacobi_run(double *u, double *v, int G, int steps);
int oracle_jacobi_run_mt(double *u, double *v, int G, int steps, int threads);

/* ---- cost model (oracle_mlp.c): the 4 -> 32 ReLU -> 32 ReLU -> C network's
 * forward pass and mean-squared-error gradient in double precision ---- */
int oracle_mlp_num_params(int c);
void oracle_mlp_forward(const double *P, int C, const double *X, int N, double *Y);
double oracle_mlp_loss_grad(const double *P, int C, const double *X, const double *T, int N, double *grad);

#ifdef __cplusplus
}
#endif

#endif
