/*
 * autotune.hip -- exhaustive autotuner for the per-request choices of the
 * fused batched put (no counterpart in the reference).
 *
 * Candidate set: pieces {1,2,4,...,256} x widest access {4,8,16} B x
 * {wavefront, workgroup} pieces = 54 candidates (a documented subset of
 * workgroups 1..256; powers of two cover the range at 2x resolution).
 *
 * A feature point is (log2 bytes, element width, concurrent requests,
 * alignment class).  Measuring a point: K = concurrency requests of that size
 * (capped so they fit a 256 MiB scratch pool) are planned with one candidate
 * applied to all of them and launched as one fused kernel; the time per
 * launch (HIP events, `reps` launches after one warm-up) is that candidate's
 * time.  Every candidate of every point is measured: the exhaustive table.
 * The same measurements are the cost model's training samples.
 *
 * Alignment class of a request: 0 when target and source agree modulo 16,
 * 1 modulo 8, 2 otherwise (modulo 4: payloads are 4-byte elements).
 */
#include <algorithm>
#include <math.h>
#include <vector>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

using namespace nvshmemi;

namespace {
const int kPieces[9] = {1, 2, 4, 8, 16, 32, 64, 128, 256};
const int kVec[3] = {4, 8, 16};
constexpr int kNumCand = 54;
constexpr size_t kPool = (size_t)256 << 20;
constexpr size_t kMaxPieces = 65536;

void candidate_of(int i, nvshmemx_batch_choice_t *c) {
    c->pieces = kPieces[i % 9];
    c->vec_bytes = kVec[(i / 9) % 3];
    c->mode = i / 27;
}
} // namespace

extern "C" int nvshmemx_tune_candidate_count(void) { return kNumCand; }

extern "C" void nvshmemx_tune_candidate(int i, nvshmemx_batch_choice_t *out) {
    if (i < 0 || i >= kNumCand) NVSHMEMI_FATAL("candidate %d out of range [0, %d)", i, kNumCand);
    candidate_of(i, out);
}

extern "C" int nvshmemx_align_class(const void *src, const void *dst) {
    const uintptr_t rel = (uintptr_t)src ^ (uintptr_t)dst;
    if ((rel & 15) == 0) return 0;
    if ((rel & 7) == 0) return 1;
    return 2;
}

/* normalised model inputs for one request */
extern "C" void nvshmemx_tune_features(double log2_bytes, int elem_bytes, int concurrency, int align_class, float *f) {
    f[0] = (float)((log2_bytes - 14.0) / 8.0);
    f[1] = (float)((double)elem_bytes - 6.0) / 2.0f;
    f[2] = (float)((log2((double)std::max(1, concurrency)) - 6.0) / 6.0);
    f[3] = (float)align_class / 2.0f;
}

/* Measures every candidate at every point; times_us[p * 54 + c].  Needs no
 * PE state (scratch buffers are private to the call).  Returns 0. */
extern "C" int nvshmemx_tune_measure(int npts, const double *log2_bytes, const int *elem_bytes, const int *concurrency,
                                     const int *align_class, int reps, int workgroups, int block, float *times_us) {
    (void)elem_bytes;
    if (npts < 0 || reps < 1) NVSHMEMI_FATAL("nvshmemx_tune_measure: bad arguments");
    if (workgroups <= 0) workgroups = 1024;
    if (block <= 0) block = 256;
    char *src = nullptr, *dst = nullptr;
    NVSHMEMI_HIP(hipMalloc((void **)&src, kPool + 4096));
    NVSHMEMI_HIP(hipMalloc((void **)&dst, kPool + 4096));
    NVSHMEMI_HIP(hipMemset(src, 0x5A, kPool + 4096));
    NVSHMEMI_HIP(hipMemset(dst, 0, kPool + 4096));
    hipEvent_t e0, e1;
    NVSHMEMI_HIP(hipEventCreate(&e0));
    NVSHMEMI_HIP(hipEventCreate(&e1));
    nvshmemx_batch_plan *plan = plan_new();
    std::vector<resolved_req> reqs;
    for (int p = 0; p < npts; ++p) {
        size_t size = (size_t)floor(exp2(log2_bytes[p]));
        size = std::max<size_t>(4, size & ~(size_t)3);
        const size_t stride = (size + 32 + 255) / 256 * 256;
        /* the most pieces any candidate makes of one request bounds how many
         * requests a point replays, keeping every plan <= kMaxPieces pieces
         * (identical for all 54 candidates of the point) */
        const size_t max_parts = std::max<size_t>(1, std::min<size_t>(1024, size / 256));
        long long kcap = std::min<long long>((long long)(kPool / stride), (long long)(kMaxPieces / max_parts));
        const size_t k = (size_t)std::max<long long>(1, std::min<long long>(concurrency[p], std::max(1LL, kcap)));
        const size_t doff = align_class[p] == 0 ? 0 : (align_class[p] == 1 ? 8 : 4);
        reqs.resize(k);
        for (size_t i = 0; i < k; ++i) reqs[i] = resolved_req{src + i * stride, dst + i * stride + doff, size, 0};
        for (int c = 0; c < kNumCand; ++c) {
            nvshmemx_batch_choice_t ch;
            candidate_of(c, &ch);
            plan_fill(plan, reqs, nullptr, &ch, 1, workgroups, block, NT_BOTH, false);
            plan_launch(plan, nullptr);
            NVSHMEMI_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r) plan_launch(plan, nullptr);
            NVSHMEMI_HIP(hipEventRecord(e1, nullptr));
            NVSHMEMI_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NVSHMEMI_HIP(hipEventElapsedTime(&ms, e0, e1));
            times_us[(size_t)p * kNumCand + c] = ms * 1000.0f / (float)reps;
        }
    }
    nvshmemx_batch_plan_destroy(plan);
    NVSHMEMI_HIP(hipEventDestroy(e0));
    NVSHMEMI_HIP(hipEventDestroy(e1));
    NVSHMEMI_HIP(hipFree(src));
    NVSHMEMI_HIP(hipFree(dst));
    return 0;
}

/* Table choice: each request takes the measured-best candidate of the
 * nearest point (Euclidean distance in the normalised feature space).
 * `features` are npts x 4 (nvshmemx_tune_features), `times_us` npts x 54. */
extern "C" int nvshmemx_tune_table_choose(const float *features, const float *times_us, int npts,
                                          const nvshmemx_put_request_t *reqs, int nreq, int concurrency,
                                          nvshmemx_batch_choice_t *out) {
    if (npts < 1) NVSHMEMI_FATAL("nvshmemx_tune_table_choose: empty table");
    std::vector<int> best((size_t)npts);
    for (int p = 0; p < npts; ++p) {
        const float *t = times_us + (size_t)p * kNumCand;
        best[(size_t)p] = (int)(std::min_element(t, t + kNumCand) - t);
    }
    for (int i = 0; i < nreq; ++i) {
        float f[4];
        const void *dst = g.initialized ? nvshmem_ptr(reqs[i].target, reqs[i].pe) : reqs[i].target;
        nvshmemx_tune_features(log2((double)std::max<size_t>(1, reqs[i].nbytes)), reqs[i].elem_bytes, concurrency,
                               nvshmemx_align_class(reqs[i].source, dst), f);
        int arg = 0;
        float dbest = 1e30f;
        for (int p = 0; p < npts; ++p) {
            const float *q = features + (size_t)p * 4;
            float d = 0;
            for (int j = 0; j < 4; ++j) d += (f[j] - q[j]) * (f[j] - q[j]);
            if (d < dbest) dbest = d, arg = p;
        }
        candidate_of(best[(size_t)arg], &out[i]);
    }
    return 0;
}
