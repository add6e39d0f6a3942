/*
 * nvshmemx.h -- host-side extensions of the MI355X-native OpenSHMEM subset.
 *
 * None of these has a counterpart in the reference's host header
 * (reference nvshmem.h:44-52); they are what the MI355X design adds around the
 * same semantics: host-launched grid-wide put/get (the bandwidth path), stream
 * selection, workgroup-partitioned PEs in one process, sticky device errors in
 * place of the reference's print-and-continue time-outs
 * (nvshmem_kernels.cu:65-68), and the batched-put engine with its autotuner
 * and learned cost model.  Plain C ABI: pointers, sizes and ints only.
 */
#ifndef NVSHMEM_AMD_NVSHMEMX_H
#define NVSHMEM_AMD_NVSHMEMX_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "nvshmem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sticky device error bits returned by nvshmemx_get_last_error() */
#define NVSHMEMX_ERR_WAIT_TIMEOUT 0x1    /* a device wait outlived NVSHMEM_WAIT_TIMEOUT_MS */
#define NVSHMEMX_ERR_BARRIER_TIMEOUT 0x2 /* an offload barrier outlived the same budget */
#define NVSHMEMX_ERR_BAD_PE 0x4          /* a device call named a PE outside [0, npes); it did nothing */
#define NVSHMEMX_ERR_TRACE_OVERFLOW 0x8
#define NVSHMEMX_ERR_TRACE_CONFLICT 0x10
#define NVSHMEMX_ERR_BAD_ADDRESS 0x20     /* a device call's symmetric range left the heap; it did nothing */

#define NVSHMEMX_MODE_PROCESS 0 /* one PE per process, peers' heaps IPC-mapped */
#define NVSHMEMX_MODE_LOCAL 1   /* all PEs in this process, workgroup-partitioned */

/* ---- layout and state ---- */

/* Start `npes` PEs inside this process: one heap allocation with a
 * power-of-two stride per PE; in kernels, workgroup b belongs to PE
 * b / wgs_per_pe (wgs_per_pe a power of two).  Same as nvstart_pes() with
 * NVSHMEM_LOCAL_PES / NVSHMEM_WGS_PER_PE set.  Returns 0. */
int nvshmemx_start_local_pes(int npes, int wgs_per_pe);
int nvshmemx_initialized(void);
int nvshmemx_mode(void);
int nvshmemx_wgs_per_pe(void);
size_t nvshmemx_heap_size(void);
/* How the symmetric heap was allocated (NVSHMEM_HEAP_KIND): coarse-grained
 * hipMalloc memory (the default) or fine-grained hipExtMallocWithFlags memory. */
#define NVSHMEMX_HEAP_COARSE 0
#define NVSHMEMX_HEAP_FINE 1
int nvshmemx_heap_kind(void);
const char *nvshmemx_bootstrap_name(void);
int nvshmemx_wallclock_khz(void);

/* Synchronise the device and return the sticky NVSHMEMX_ERR_* bits. */
int nvshmemx_get_last_error(void);
void nvshmemx_clear_errors(void);
/* Budget of every device wait and barrier before it records a time-out. */
void nvshmemx_set_wait_timeout_ms(unsigned long long ms);

/* Stream used by nvshmem_barrier_all_offload (default: the null stream). */
void nvshmemx_set_stream(hipStream_t stream);
hipStream_t nvshmemx_get_stream(void);

/* ---- host-launched bulk copies (grid-cooperative, 16-byte vectorised) ---- */

/* target is symmetric (resolved on `pe`), source is a local device pointer */
void nvshmemx_putmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t stream);
/* target is a local device pointer, source is symmetric (read on `pe`) */
void nvshmemx_getmem_on_stream(void *target, const void *source, size_t nbytes, int pe, hipStream_t stream);
void nvshmemx_int_put_on_stream(int *target, const int *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_float_put_on_stream(float *target, const float *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_double_put_on_stream(double *target, const double *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_int_get_on_stream(int *target, const int *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_float_get_on_stream(float *target, const float *source, size_t len, int pe, hipStream_t stream);
void nvshmemx_double_get_on_stream(double *target, const double *source, size_t len, int pe, hipStream_t stream);

/* Launch shape of a host-launched copy.  Negative fields keep the library's
 * choice for that field, and so do zero workgroups, block, vec_bytes, unroll
 * and streams (zero is a valid nontemporal policy and mode). */
typedef struct nvshmemx_copy_config {
    int workgroups;  /* grid size */
    int block;       /* threads per workgroup (64..1024, multiple of 64) */
    int vec_bytes;   /* widest access: 4, 8 or 16 */
    int unroll;      /* independent accesses in flight per lane: 1, 2, 4 or 8 */
    int nontemporal; /* bit 0: non-temporal loads, bit 1: non-temporal stores, bit 2: stores written through the L2 (sc1) instead */
    int mode;        /* 0: lanes of the whole grid interleave; 1: one contiguous slice per workgroup;
                        2: tiles of unroll * block accesses alternating between `streams` address runs
                           (workgroups 0: one tile per workgroup) */
    int streams;     /* mode 2: 1..64 address runs */
} nvshmemx_copy_config_t;

void nvshmemx_putmem_config_on_stream(void *target, const void *source, size_t nbytes, int pe,
                                      const nvshmemx_copy_config_t *cfg, hipStream_t stream);
/* The launch shape nvshmemx_{put,get}mem_on_stream use for `nbytes`. */
void nvshmemx_default_copy_config(size_t nbytes, nvshmemx_copy_config_t *out);

/* Host-launched strided and 2-D transfers (grid-wide).  iput: target[i * tst]
 * = source[i * sst] on `pe`, i < nelems; iget: the same read from `pe`
 * (strides in elements, >= 1).  2-D: `height` rows of `width` bytes, row r at
 * target + r * tpitch and source + r * spitch (pitches >= width); the
 * symmetric side is the target of a put and the source of a get. */
void nvshmemx_putmem_2d_on_stream(void *target, size_t tpitch, const void *source, size_t spitch, size_t width,
                                  size_t height, int pe, hipStream_t stream);
void nvshmemx_getmem_2d_on_stream(void *target, size_t tpitch, const void *source, size_t spitch, size_t width,
                                  size_t height, int pe, hipStream_t stream);
void nvshmemx_int_iput_on_stream(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems, int pe,
                                 hipStream_t stream);
void nvshmemx_float_iput_on_stream(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                   int pe, hipStream_t stream);
void nvshmemx_double_iput_on_stream(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                    int pe, hipStream_t stream);
void nvshmemx_int_iget_on_stream(int *target, const int *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems, int pe,
                                 hipStream_t stream);
void nvshmemx_float_iget_on_stream(float *target, const float *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                   int pe, hipStream_t stream);
void nvshmemx_double_iget_on_stream(double *target, const double *source, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,
                                    int pe, hipStream_t stream);

/* Stream-ordered device barrier on an explicit stream. */
void nvshmemx_barrier_all_on_stream(hipStream_t stream);

/* ---- batched puts: many puts of uneven size in one fused launch ---- */

typedef struct nvshmemx_put_request {
    void *target;       /* symmetric address, resolved on `pe` */
    const void *source; /* local device address */
    size_t nbytes;
    int pe;
    int elem_bytes;     /* element width of the payload (4 or 8) */
} nvshmemx_put_request_t;

/* How one request runs inside the fused launch. */
typedef struct nvshmemx_batch_choice {
    int pieces;    /* 1..256 parts; a part is one workgroup's work (mode 1) or four wavefronts' (mode 0) */
    int vec_bytes; /* widest access: 4, 8 or 16 */
    int mode;      /* 0: wavefront pieces, 1: workgroup pieces, 2: tiles of 512 accesses of the widest width
                      <= vec_bytes the alignment allows, one workgroup of 256 threads a tile, all of a
                      plan's tiles in source order in a launch of their own (pieces is ignored) */
} nvshmemx_batch_choice_t;

#define NVSHMEMX_BATCH_SHARED 0 /* every workgroup of the launch serves every request */
#define NVSHMEMX_BATCH_OWNER 1  /* local PEs: PE p's workgroups serve the requests whose source is in PE p's heap */
#define NVSHMEMX_BATCH_SOURCE_ORDER 2 /* as SHARED, but each workgroup takes a run of pieces in source-address order */

typedef struct nvshmemx_batch_launch {
    int workgroups;  /* fused grid size (0: one workgroup per piece, at most 8192) */
    int block;       /* threads per workgroup, multiple of 64 (0: 1024) */
    int nontemporal; /* bit 0: non-temporal loads, bit 1: non-temporal stores, bit 2: stores written through the
                        L2 (sc1) instead (NULL launch: non-temporal stores only) */
    int discipline;  /* NVSHMEMX_BATCH_SHARED or NVSHMEMX_BATCH_OWNER */
} nvshmemx_batch_launch_t;

typedef struct nvshmemx_batch_plan *nvshmemx_batch_plan_t;

/* Build a plan (device-resident piece queues) for `nreq` requests.  `choices`
 * may be NULL (library rule, nvshmemx_batch_default_choice), `launch` may be
 * NULL (grid sized to the work, at most 8192 workgroups of 1024 threads,
 * non-temporal stores, NVSHMEMX_BATCH_SOURCE_ORDER).  Returns 0. */
int nvshmemx_batch_plan_create(const nvshmemx_put_request_t *reqs, int nreq, const nvshmemx_batch_choice_t *choices,
                               const nvshmemx_batch_launch_t *launch, nvshmemx_batch_plan_t *plan);
/* One fused launch executing every request of the plan. */
void nvshmemx_batch_plan_launch(nvshmemx_batch_plan_t plan, hipStream_t stream);
/* Capture `steps` back-to-back launches of the plan into a hipGraph; replay
 * it with nvshmemx_batch_plan_graph_launch.  The graph holds kernel nodes
 * only.  Returns 0. */
int nvshmemx_batch_plan_capture(nvshmemx_batch_plan_t plan, int steps);
void nvshmemx_batch_plan_graph_launch(nvshmemx_batch_plan_t plan, hipStream_t stream);
/* Select the fastest of `nplans` plans built for the same exchange (for instance with
 * the cost model's choices, the autotuner table's and the untuned rule): one
 * timed warm-up launch each, then `reps` rounds that time every plan in turn
 * on `stream` over a burst of back-to-back launches (about 50 us, at most
 * 64 launches, between two HIP events); returns the index of the plan with
 * the smallest median time per launch and, if ms_out is not NULL, writes
 * every plan's median (ms per launch).  Each launch executes the plan's puts,
 * so select when the sources hold what the exchange sends.  No counterpart in
 * the reference. */
int nvshmemx_batch_plan_select(const nvshmemx_batch_plan_t *plans, int nplans, int reps, hipStream_t stream,
                              float *ms_out);
size_t nvshmemx_batch_plan_pieces(nvshmemx_batch_plan_t plan);
size_t nvshmemx_batch_plan_bytes(nvshmemx_batch_plan_t plan);
void nvshmemx_batch_plan_destroy(nvshmemx_batch_plan_t plan);
/* The library's untuned rule for one request. */
void nvshmemx_batch_default_choice(size_t nbytes, int elem_bytes, nvshmemx_batch_choice_t *out);
/* The per-call path: one grid put launch per request. */
void nvshmemx_put_requests_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t stream);

/* ---- batched puts planned on the device by the issuing PEs ---- */

/* Execute `nreq` puts (the request array in DEVICE memory) with one
 * persistent kernel on `stream`.  A request belongs to the PE that issued it:
 * in local mode the PE whose heap holds its source (a private source counts
 * for PE i mod npes), in process mode this PE.  Workgroup p < npes (local
 * mode; workgroup 0 in process mode) is PE p's leader and cuts that PE's
 * requests into chunks on the device (a PE with at most 64 requests and
 * 128 KiB of payload is copied by its leader instead); then every workgroup
 * copies chunks of all PEs' queues in PE order, claimed one at a time from
 * one counter.  Destinations must not overlap (puts are unordered, as in the
 * fused plans).  A request naming a bad PE or a target range outside the heap
 * is skipped and sets NVSHMEMX_ERR_BAD_PE / NVSHMEMX_ERR_BAD_ADDRESS.
 * The library allocates the kernel's scratch (a control block, a 256 KiB
 * chunk table and 8 bytes per request), growing it on demand; launches tag
 * their flags with a launch number the kernel counts itself, so nothing is
 * cleared between launches and a captured call replays without a memset.
 * The first call (or one with more requests than any before) allocates and
 * so cannot be captured.  Calls that overlap on one GPU share the scratch
 * and race.  Returns 0. */
int nvshmemx_batch_put_persistent_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t stream);
/* workgroups of the persistent kernel's grid (every one resident at once) */
int nvshmemx_batch_persistent_grid(void);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same kernel and results as
 * nvshmemx_batch_put_persistent_on_stream, with its scratch in a static
 * per-GPU array of the library (1.25 MiB, set up by nvstart_pes /
 * nvshmemx_start_local_pes) instead of allocated memory; the call itself
 * allocates nothing and only launches the kernel, so even a first call can be
 * captured.  At most nvshmemx_batch_persistent_max_requests_noalloc()
 * (65,536) requests: a larger list is refused with a message and -1, nothing
 * launched.  Every call on a GPU shares that GPU's array (each launch
 * leaves it ready for the next; nothing is cleared between calls); calls
 * that overlap on one GPU race.  Returns 0. */
int nvshmemx_batch_put_persistent_noalloc_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t stream);
int nvshmemx_batch_persistent_max_requests_noalloc(void);
/* No counterpart in the reference.  Move every put and get posted on the
 * device with nvshmemx_putmem_nbi / nvshmemx_getmem_nbi (nvshmemx_device.h)
 * since the last flush: one launch
 * of the persistent kernel on `stream`, in which each local PE's leader
 * workgroup plans its own PE's queue (or copies it itself when it is small)
 * and every workgroup copies chunks of all queues in PE order.  Stream-order
 * it after the kernels that posted and before anything that reads the
 * targets; kernels that post while a flush runs race with it.  The queues
 * and the flush's scratch are static per-GPU arrays of the library (3.25 MiB
 * together): a flush allocates nothing and can be captured in a hipGraph.
 * Returns 0. */
int nvshmemx_nbi_flush_on_stream(hipStream_t stream);
/* posts a PE's queue holds between flushes (65,536 / local PEs in local
 * mode, 65,536 in process mode); later posts are copied as they are made */
int nvshmemx_nbi_capacity(void);

/* ---- exhaustive autotuner of batch choices ---- */

/* The 54 candidates: pieces {1,2,..,256} x access {4,8,16} B x {wave, workgroup}. */
int nvshmemx_tune_candidate_count(void);
void nvshmemx_tune_candidate(int index, nvshmemx_batch_choice_t *out);
/* 0: source and target agree modulo 16; 1: modulo 8; 2: otherwise */
int nvshmemx_align_class(const void *source, const void *target);
/* the four normalised model inputs of one request */
void nvshmemx_tune_features(double log2_bytes, int elem_bytes, int concurrency, int align_class, float *features4);
/* Measure every candidate at every point (fused launches over private
 * scratch buffers, HIP-event timed); times_us[p * 54 + c].  Returns 0. */
int nvshmemx_tune_measure(int npts, const double *log2_bytes, const int *elem_bytes, const int *concurrency,
                          const int *align_class, int reps, int workgroups, int block, float *times_us);
/* Concurrency feature of a request inside a batch of `batch_bytes`: the
 * number of same-size requests carrying that many bytes (clamped to
 * [1, 32768]); `given` > 0 is returned unchanged.  The choose functions below
 * use it when their `concurrency` argument is <= 0. */
int nvshmemx_effective_concurrency(size_t nbytes, size_t batch_bytes, int given);
/* Per request: the measured-best candidate of the nearest measured point. */
int nvshmemx_tune_table_choose(const float *features, const float *times_us, int npts,
                               const nvshmemx_put_request_t *reqs, int nreq, int concurrency,
                               nvshmemx_batch_choice_t *out);

/* ---- learned cost model: fp32 MLP 4 -> 32 -> 32 -> C, trained on the GPU ---- */

typedef struct nvshmemx_costmodel *nvshmemx_costmodel_t;
int nvshmemx_costmodel_create(int ncand, unsigned long long seed, nvshmemx_costmodel_t *model);
int nvshmemx_costmodel_num_params(nvshmemx_costmodel_t model);
void nvshmemx_costmodel_get_params(nvshmemx_costmodel_t model, float *params);
void nvshmemx_costmodel_set_params(nvshmemx_costmodel_t model, const float *params);
/* gradient of the mean squared error of log-times over (X: N x 4, T: N x C) */
void nvshmemx_costmodel_gradient(nvshmemx_costmodel_t model, const float *X, const float *T, int N, float *grad);
/* `steps` full-batch Adam steps (graph-captured) */
void nvshmemx_costmodel_train(nvshmemx_costmodel_t model, const float *X, const float *T, int N, int steps, float lr);
/* predicted log(time in microseconds) per candidate: Y is N x C */
void nvshmemx_costmodel_predict(nvshmemx_costmodel_t model, const float *X, int N, float *Y);
/* per request: the candidate with the smallest predicted time */
int nvshmemx_costmodel_choose(nvshmemx_costmodel_t model, const nvshmemx_put_request_t *reqs, int nreq,
                              int concurrency, nvshmemx_batch_choice_t *out);
int nvshmemx_costmodel_save(nvshmemx_costmodel_t model, const char *path);
int nvshmemx_costmodel_load(nvshmemx_costmodel_t model, const char *path);
void nvshmemx_costmodel_destroy(nvshmemx_costmodel_t model);
int nvshmemx_costmodel_num_outputs(nvshmemx_costmodel_t model);

/* ---- batch-level policy: how a whole batch is cut into pieces ---- */

/* Policies: 0 the library's rule, 1 the per-request model's choices,
 * 2..5 every request cut into wavefront pieces of about 4/8/16/32 KiB,
 * 6..10 into workgroup pieces of about 64/128/256/512 KiB and 1 MiB,
 * 11 one workgroup piece per request, 12 every request as tiles (choice
 * mode 2). */
int nvshmemx_batch_policy_count(void);
/* the four normalised batch-level inputs of the policy model: log2 of the
 * batch's bytes, of its request count, mean and largest log2 request size */
void nvshmemx_batch_features(const nvshmemx_put_request_t *reqs, int nreq, float *features);
/* per-request choices of `policy` (policy 1 needs the per-request model) */
int nvshmemx_batch_policy_choices(const nvshmemx_put_request_t *reqs, int nreq, int policy,
                                  nvshmemx_costmodel_t request_model, nvshmemx_batch_choice_t *out);
/* Exhaustive batch-level judge: every policy's plan timed on the device
 * (nvshmemx_batch_plan_select with `reps` rounds; the puts run, so the
 * sources must hold the batch's data).  ms_out: one time per policy.
 * Returns the fastest policy.  Without a per-request model policy 1 repeats
 * the rule. */
int nvshmemx_tune_measure_batch(const nvshmemx_put_request_t *reqs, int nreq, nvshmemx_costmodel_t request_model,
                                int reps, float *ms_out);
/* The policy model's choice (a cost model created with
 * nvshmemx_batch_policy_count() outputs, trained on batch features against
 * log-times of the policies): the policy with the smallest predicted time
 * plans the batch.  policy_out (optional) receives it. */
int nvshmemx_costmodel_choose_batch(nvshmemx_costmodel_t batch_model, nvshmemx_costmodel_t request_model,
                                    const nvshmemx_put_request_t *reqs, int nreq, nvshmemx_batch_choice_t *out,
                                    int *policy_out);

/* Tracing (library built with NVSHMEM_TRACING): drain the ring now; report
 * whether records are being kept.  Both are no-ops / 0 otherwise. */
void nvshmemx_trace_flush(void);
int nvshmemx_trace_enabled(void);

/* Launch a kernel whose workgroups must all be resident at once (kernels that
 * call nvshmemx_barrier_all_wg); exits with an error when the grid does not
 * fit the device. */
void nvshmemx_collective_launch(const void *func, dim3 grid, dim3 block, void **args, size_t shmem,
                                hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NVSHMEM_AMD_NVSHMEMX_H */
