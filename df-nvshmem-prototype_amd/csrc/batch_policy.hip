/*
 * batch_policy.hip -- batch-level choice of how a batch of puts is cut into
 * pieces (no counterpart in the reference).
 *
 * The per-request cost model (costmodel.hip) predicts, for one request, the
 * time of every candidate split from the request's own features and an
 * "effective concurrency" -- the identical-request scenario its
 * microbenchmark samples were measured in.  What a request's best split is
 * also depends on the rest of the batch: a short launch of 128 mixed puts is
 * bound by how evenly its pieces fill the device, a long one of 4096 by how
 * the pieces walk memory.  Measured on the device (64 PEs, held-out batches
 * of the uneven-put workload), among the piece policies cutting every request
 * into 64 KiB workgroup pieces was the best plan of 128-put and 4096-put
 * batches, and the per-request network's own choices the best of 16-put
 * batches; every request cut into tiles (one short-lived workgroup a tile,
 * policy 12) beat them all at 4096 puts and matched the best at 16 and 128.
 *
 * So the batch is planned by one of a family of policies:
 *   0       the library's rule (per-request default_choice);
 *   1       the per-request network's choices;
 *   2 .. 5  every request cut into wavefront pieces of about 4, 8, 16, 32 KiB;
 *   6 .. 10 into workgroup pieces of about 64, 128, 256, 512 KiB and 1 MiB;
 *   11      one workgroup piece per request;
 *   12      every request cut into tiles of 512 accesses, one workgroup of
 *           256 threads a tile, all tiles in source order (choice mode 2);
 * (16-byte accesses where the alignment allows; at most 256 pieces a
 * request; pieces of at least 4 KiB, as every plan).  Timing every policy's
 * plan on the device is the exhaustive batch-level judge.  A second network
 * of the same shape (4 -> 32 -> 32 ReLU -> 13, the same HIP kernels) reads
 * four batch-level inputs -- log2 of the batch's bytes, of its request count,
 * the mean and the largest log2 request size -- and predicts the log-time of
 * every policy; it is trained on batches the library generates and times
 * itself, and picks the policy with the smallest prediction.
 */
#include <algorithm>
#include <math.h>
#include <vector>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_host.h"
#include "batch_plan.h"

using namespace nvshmemi;

namespace {
struct policy {
    int kind;       /* 0 rule, 1 per-request network, 2 uniform pieces, 3 tiles */
    size_t gran;    /* uniform: bytes per piece (0: one piece per request) */
    int mode;       /* uniform: 0 wavefront pieces, 1 workgroup pieces; tiles: 2 */
};
const policy kPolicies[] = {
    {0, 0, 1},         {1, 0, 1},         {2, 4u << 10, 0},   {2, 8u << 10, 0},
    {2, 16u << 10, 0}, {2, 32u << 10, 0}, {2, 64u << 10, 1},  {2, 128u << 10, 1},
    {2, 256u << 10, 1}, {2, 512u << 10, 1}, {2, 1u << 20, 1}, {2, 0, 1},
    {3, 0, 2},
};
constexpr int kNumPolicies = (int)(sizeof(kPolicies) / sizeof(kPolicies[0]));

void check_policy(int p, const char *fn) {
    if (p < 0 || p >= kNumPolicies) NVSHMEMI_FATAL("%s: policy %d out of range [0, %d)", fn, p, kNumPolicies);
}
} // namespace

extern "C" int nvshmemx_batch_policy_count(void) { return kNumPolicies; }

/* normalised batch-level inputs of the policy network */
extern "C" void nvshmemx_batch_features(const nvshmemx_put_request_t *reqs, int nreq, float *f) {
    double total = 0.0, sum_log = 0.0, max_log = 0.0;
    for (int i = 0; i < nreq; ++i) {
        const double b = (double)std::max<size_t>(1, reqs[i].nbytes);
        total += b;
        const double l = log2(b);
        sum_log += l;
        max_log = std::max(max_log, l);
    }
    const int n = std::max(1, nreq);
    f[0] = (float)((log2(std::max(1.0, total)) - 24.0) / 8.0);
    f[1] = (float)((log2((double)n) - 7.0) / 4.0);
    f[2] = (float)((sum_log / n - 14.0) / 8.0);
    f[3] = (float)((max_log - 14.0) / 8.0);
}

/* the per-request choices policy `p` makes for a batch; the per-request
 * network's policy needs `request_model` */
extern "C" int nvshmemx_batch_policy_choices(const nvshmemx_put_request_t *reqs, int nreq, int p,
                                             nvshmemx_costmodel_t request_model, nvshmemx_batch_choice_t *out) {
    check_policy(p, "nvshmemx_batch_policy_choices");
    if (nreq < 1) return 0;
    const policy &pol = kPolicies[p];
    if (pol.kind == 1) {
        if (!request_model) NVSHMEMI_FATAL("nvshmemx_batch_policy_choices: policy 1 needs the per-request model");
        return nvshmemx_costmodel_choose(request_model, reqs, nreq, 0, out);
    }
    for (int i = 0; i < nreq; ++i) {
        if (pol.kind == 0) {
            out[i] = default_choice(reqs[i].nbytes);
            continue;
        }
        if (pol.kind == 3) {
            out[i].pieces = 1;
            out[i].vec_bytes = 16;
            out[i].mode = 2;
            continue;
        }
        size_t pieces = 1;
        if (pol.gran) {
            const size_t per = pol.gran * (pol.mode ? 1 : 4); /* a wavefront choice's pieces are cut in four */
            pieces = (reqs[i].nbytes + per - 1) / per;
        }
        out[i].pieces = (int)std::max<size_t>(1, std::min<size_t>(pieces, 256));
        out[i].vec_bytes = 16;
        out[i].mode = pol.mode;
    }
    return 0;
}

/* The exhaustive batch-level judge: one plan per policy, all timed on the
 * device by nvshmemx_batch_plan_select (a warm-up launch each, then `reps`
 * rounds of bursts of back-to-back launches; median per launch).  The puts
 * run, so the sources must hold the batch's data.  ms_out receives every
 * policy's time; returns the fastest policy. */
extern "C" int nvshmemx_tune_measure_batch(const nvshmemx_put_request_t *reqs, int nreq,
                                           nvshmemx_costmodel_t request_model, int reps, float *ms_out) {
    require_init("nvshmemx_tune_measure_batch");
    if (nreq < 1 || !reqs || reps < 1) NVSHMEMI_FATAL("nvshmemx_tune_measure_batch: bad arguments");
    std::vector<nvshmemx_batch_choice_t> ch((size_t)nreq);
    std::vector<nvshmemx_batch_plan_t> plans((size_t)kNumPolicies, nullptr);
    for (int p = 0; p < kNumPolicies; ++p) {
        if (kPolicies[p].kind == 1 && !request_model) { /* no per-request model: its slot repeats the rule */
            nvshmemx_batch_policy_choices(reqs, nreq, 0, nullptr, ch.data());
        } else {
            nvshmemx_batch_policy_choices(reqs, nreq, p, request_model, ch.data());
        }
        nvshmemx_batch_plan_create(reqs, nreq, ch.data(), nullptr, &plans[(size_t)p]);
    }
    const int best = nvshmemx_batch_plan_select(plans.data(), kNumPolicies, reps, nullptr, ms_out);
    for (nvshmemx_batch_plan_t pl : plans) nvshmemx_batch_plan_destroy(pl);
    return best;
}

/* The policy network's choice: predicted log-time of every policy from the
 * batch's features, smallest wins; the batch is then planned by that policy.
 * policy_out (optional) receives the policy. */
extern "C" int nvshmemx_costmodel_choose_batch(nvshmemx_costmodel_t batch_model, nvshmemx_costmodel_t request_model,
                                               const nvshmemx_put_request_t *reqs, int nreq,
                                               nvshmemx_batch_choice_t *out, int *policy_out) {
    if (!batch_model || nvshmemx_costmodel_num_outputs(batch_model) != kNumPolicies)
        NVSHMEMI_FATAL("nvshmemx_costmodel_choose_batch: the batch model must predict all %d policies", kNumPolicies);
    float f[4];
    nvshmemx_batch_features(reqs, nreq, f);
    std::vector<float> y((size_t)kNumPolicies);
    nvshmemx_costmodel_predict(batch_model, f, 1, y.data());
    int best = 0;
    for (int p = 1; p < kNumPolicies; ++p)
        if (y[(size_t)p] < y[(size_t)best]) best = p;
    if (kPolicies[best].kind == 1 && !request_model) best = 0;
    if (policy_out) *policy_out = best;
    return nvshmemx_batch_policy_choices(reqs, nreq, best, request_model, out);
}
