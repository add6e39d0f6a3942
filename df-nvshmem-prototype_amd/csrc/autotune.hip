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
 * (capped so they fit a 1 GiB scratch pool) are planned with one candidate
 * applied to all of them and launched as one fused kernel; the time per
 * launch (HIP events, `reps` launches after one warm-up) is that candidate's
 * time.  A candidate that would make more than 65536 pieces runs on fewer
 * requests and its time is scaled up to K.  Pieces are dealt to workers in a
 * fixed scattered order: inside a real mixed batch a request's pieces run
 * beside unrelated pieces, never as one address-ordered sweep, and small
 * pieces measured as a sweep look up to ~6% faster than they are there.  Every candidate of every point is measured: the exhaustive table.
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
constexpr size_t kPool = (size_t)1 << 30;
constexpr size_t kMaxPieces = 65536;
constexpr int kMaxConcurrency = 32768;

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

/* Concurrency a request of a mixed batch is planned with: the number of
 * requests of its own size that would carry the batch's total bytes, i.e.
 * the identical-request scenario the microbenchmarks measure, clamped to
 * [1, kMaxConcurrency].  `given` > 0 overrides it for every request. */
extern "C" int nvshmemx_effective_concurrency(size_t nbytes, size_t batch_bytes, int given) {
    if (given > 0) return given;
    const double k = (double)batch_bytes / (double)std::max<size_t>(1, nbytes);
    return (int)std::max(1.0, std::min((double)kMaxConcurrency, floor(k + 0.5)));
}

namespace nvshmemi {
size_t resident_threads() {
    static size_t n = 0;
    if (!n) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxThreadsPerMultiProcessor, dev) == hipSuccess &&
            cus > 0 && per_cu > 0)
            n = (size_t)cus * (size_t)per_cu;
        else
            n = (size_t)256 * 2048; /* MI355X: 256 compute units x 2048 threads */
    }
    return n;
}

/* Largest piece the schedule of a batch can absorb, as bytes per thread of
 * the piece's workers.  Any list schedule ends within (mean load + largest
 * piece) (Graham's bound), so pieces held to 1/32 of a thread's fair share of
 * the batch keep the tail within about 3% of the mean.  The identical-request
 * microbenchmarks cannot see this (all their pieces are equal, so any piece
 * size balances), hence the cap is applied when choosing.  The cap is 0 (none)
 * below 16 bytes per thread: such batches do not fill the device and the
 * measured concurrency feature covers them. */
size_t balance_cap(size_t batch_bytes) {
    const size_t cap = batch_bytes / resident_threads() / 32;
    return cap >= 16 ? cap : 0;
}

/* Contiguous bytes of one piece of candidate `c` on a request of `nbytes`. */
static size_t piece_bytes(size_t nbytes, int c) {
    nvshmemx_batch_choice_t ch;
    candidate_of(c, &ch);
    const size_t parts = std::min<size_t>((size_t)ch.pieces * (ch.mode ? 1 : 4), std::max<size_t>(1, nbytes / kPieceQuantum));
    return ((nbytes + parts - 1) / parts + kPieceQuantum - 1) / kPieceQuantum * kPieceQuantum;
}

/* Bytes per thread of one piece: a workgroup piece is shared by the default
 * batch workgroup's threads, a wavefront piece by 64 lanes. */
static size_t piece_cost(size_t nbytes, int c) {
    return piece_bytes(nbytes, c) / (size_t)(c / 27 ? kBatchBlock : 64);
}

/* Constrained choice over the 54 candidates' costs (times, or log-times when
 * `log_cost`).  Only candidates whose pieces fit under `cap` compete (when
 * none fits: the candidate with the smallest pieces).  Among the competitors
 * whose cost is within kTieTolerance of the cheapest, the one with the
 * longest contiguous pieces, then the widest accesses, wins: measured costs
 * carry a few per cent of noise, and longer pieces and wider accesses mean
 * fewer descriptors, fewer instructions and fewer scattered accesses in the
 * fused launch (the one-standard-error rule of model selection). */
constexpr double kTieTolerance = 0.05;
int pick_candidate(const float *cost, size_t nbytes, size_t cap, bool log_cost) {
    int best = -1, finest = 0;
    size_t pc[kNumCand];
    for (int c = 0; c < kNumCand; ++c) {
        pc[c] = piece_cost(nbytes, c);
        if ((!cap || pc[c] <= cap) && (best < 0 || cost[c] < cost[best])) best = c;
        if (pc[c] < pc[finest] || (pc[c] == pc[finest] && cost[c] < cost[finest])) finest = c;
    }
    if (best < 0) return finest;
    const double limit = log_cost ? (double)cost[best] + log1p(kTieTolerance) : (double)cost[best] * (1.0 + kTieTolerance);
    int pick = best;
    size_t pb_pick = piece_bytes(nbytes, pick);
    for (int c = 0; c < kNumCand; ++c) {
        if ((cap && pc[c] > cap) || (double)cost[c] > limit) continue;
        const size_t pb = piece_bytes(nbytes, c);
        const int vc = kVec[(c / 9) % 3], vp = kVec[(pick / 9) % 3];
        if (pb > pb_pick || (pb == pb_pick && (vc > vp || (vc == vp && cost[c] < cost[pick])))) pick = c, pb_pick = pb;
    }
    return pick;
}
} // namespace nvshmemi

static size_t batch_bytes_of(const nvshmemx_put_request_t *reqs, int n) {
    size_t t = 0;
    for (int i = 0; i < n; ++i) t += reqs[i].nbytes;
    return t;
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
    if (workgroups < 0) workgroups = 0; /* 0: sized to the work, as the batch plans' default */
    if (block <= 0) block = kBatchBlock;
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
        /* k_full same-size requests (as many of the nominal concurrency as
         * the pool holds) are what a point stands for; a candidate that cuts
         * a request into many parts replays fewer of them, so that no plan
         * exceeds kMaxPieces pieces, and its time is scaled to k_full */
        const size_t k_full = (size_t)std::max<long long>(
            1, std::min<long long>(concurrency[p], (long long)(kPool / stride)));
        const size_t doff = align_class[p] == 0 ? 0 : (align_class[p] == 1 ? 8 : 4);
        for (int c = 0; c < kNumCand; ++c) {
            nvshmemx_batch_choice_t ch;
            candidate_of(c, &ch);
            const size_t parts = std::min<size_t>((size_t)ch.pieces * (ch.mode ? 1 : 4), std::max<size_t>(1, size / kPieceQuantum));
            const size_t k = std::max<size_t>(1, std::min(k_full, kMaxPieces / parts));
            reqs.resize(k);
            for (size_t i = 0; i < k; ++i) reqs[i] = resolved_req{src + i * stride, dst + i * stride + doff, size, 0};
            plan_fill(plan, reqs, nullptr, &ch, 1, workgroups, block, kBatchNontemporal, ASSIGN_SCATTERED);
            plan_launch(plan, nullptr);
            NVSHMEMI_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r) plan_launch(plan, nullptr);
            NVSHMEMI_HIP(hipEventRecord(e1, nullptr));
            NVSHMEMI_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NVSHMEMI_HIP(hipEventElapsedTime(&ms, e0, e1));
            times_us[(size_t)p * kNumCand + c] = ms * 1000.0f / (float)reps * (float)k_full / (float)k;
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
 * nearest point (Euclidean distance in the normalised feature space) among
 * those whose pieces fit the batch's balance cap.
 * `features` are npts x 4 (nvshmemx_tune_features), `times_us` npts x 54. */
extern "C" int nvshmemx_tune_table_choose(const float *features, const float *times_us, int npts,
                                          const nvshmemx_put_request_t *reqs, int nreq, int concurrency,
                                          nvshmemx_batch_choice_t *out) {
    if (npts < 1) NVSHMEMI_FATAL("nvshmemx_tune_table_choose: empty table");
    const size_t total = batch_bytes_of(reqs, nreq);
    const size_t cap = balance_cap(total);
    for (int i = 0; i < nreq; ++i) {
        float f[4];
        const void *dst = g.initialized ? nvshmem_ptr(reqs[i].target, reqs[i].pe) : reqs[i].target;
        nvshmemx_tune_features(log2((double)std::max<size_t>(1, reqs[i].nbytes)), reqs[i].elem_bytes,
                               nvshmemx_effective_concurrency(reqs[i].nbytes, total, concurrency),
                               nvshmemx_align_class(reqs[i].source, dst), f);
        int arg = 0;
        float dbest = 1e30f;
        for (int p = 0; p < npts; ++p) {
            const float *q = features + (size_t)p * 4;
            float d = 0;
            for (int j = 0; j < 4; ++j) d += (f[j] - q[j]) * (f[j] - q[j]);
            if (d < dbest) dbest = d, arg = p;
        }
        candidate_of(pick_candidate(times_us + (size_t)arg * kNumCand, reqs[i].nbytes, cap, false), &out[i]);
    }
    return 0;
}
