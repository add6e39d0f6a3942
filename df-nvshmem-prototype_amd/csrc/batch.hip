/*
 * batch.hip -- fused execution of many puts of uneven size in one launch.
 *
 * No counterpart in the reference, whose puts are one call per thread
 * (nvshmem_device.cu:487-512).  A batch of put requests is turned into a plan:
 *
 *  - each request is split into pieces by its choice (pieces x vector width x
 *    wavefront/workgroup mode); piece boundaries fall on 4 KiB offsets of the
 *    request, so every piece keeps the request's relative alignment;
 *  - the pieces are dealt to workers on the host (batch_plan.h): workgroup
 *    pieces to workgroups (all the workgroup's wavefronts share a piece) by
 *    the plan's discipline, wavefront pieces to single wavefronts longest
 *    first, each to the least-loaded one (LPT list scheduling).  A piece costs
 *    its bytes plus a fixed per-piece charge standing for the first access's
 *    latency.  Per-worker lists are laid out contiguously with offset tables;
 *  - one kernel runs every list: workgroup b copies its workgroup pieces with
 *    all its threads, then each of its wavefronts copies its own wavefront
 *    pieces.  No atomics and no per-launch state, so a plan replays from a
 *    captured hipGraph as is.
 *
 * A plan is static: the assignment is computed once and reused every step,
 * which is what the batched-exchange pattern needs.  (A first version handed
 * pieces out with a shared atomic ticket; on MI355X the ticket counter became
 * the bottleneck as the grid grew, which the host-side assignment removes.)
 *
 * Worker sets: SHARED = one set of workers serving every request, pieces
 * assigned by LPT; SOURCE_ORDER (the default) = one set, each workgroup taking
 * a run of pieces in source-address order (workgroups dispatched together read
 * neighbouring sources); OWNER = one set per PE in local mode, serving only
 * the requests whose source is in that PE's heap (workgroup-partitioned PEs,
 * each executing the puts it initiated).
 */
#include <algorithm>
#include <cmath>
#include <vector>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"
#include "batch_plan.h"

namespace nvshmemi {

template <int NT>
__device__ __forceinline__ void copy_piece(const piece_t &p, unsigned lane, unsigned n) {
    switch (p.vec) {
    case 16: copy<16, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    case 8: copy<8, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    default: copy<4, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    }
}

template <int NT>
__global__ void __launch_bounds__(1024) k_batch(const piece_t *__restrict__ pieces,
                                                const unsigned int *__restrict__ wg_off,
                                                const unsigned int *__restrict__ wave_off) {
    const unsigned int b = blockIdx.x;
    const unsigned int wg_end = wg_off[b + 1];
    for (unsigned int j = wg_off[b]; j < wg_end; ++j) copy_piece<NT>(pieces[j], threadIdx.x, blockDim.x);
    const unsigned int w = b * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const unsigned int wave_end = wave_off[w + 1];
    for (unsigned int j = wave_off[w]; j < wave_end; ++j) copy_piece<NT>(pieces[j], threadIdx.x & 63u, 64);
}

/* One tile per workgroup (choice mode 2): lane l copies accesses l, l +
 * blockDim, ... of the tile (two for a whole tile of kTileAccesses with
 * kTileBlock threads), each load stored before the next is issued; width-1
 * tiles (unaligned heads and tails) go byte by byte. */
template <typename V, int NT>
__device__ __forceinline__ void tile_copy(const piece_t &p) {
    const size_t step = (size_t)blockDim.x * sizeof(V);
    for (size_t o = (size_t)threadIdx.x * sizeof(V); o < p.nbytes; o += step)
        vstore_p<NT>((V *)(p.dst + o), vload<(NT & NT_LOAD) != 0>((const V *)(p.src + o)));
}
template <int NT>
__global__ void __launch_bounds__(kTileBlock) k_batch_tiles(const piece_t *__restrict__ tiles, size_t ntiles) {
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) { /* one pass unless the grid was capped */
        const piece_t p = tiles[t];
        switch (p.vec) {
        case 16: tile_copy<u32x4, NT>(p); break;
        case 8: tile_copy<u32x2, NT>(p); break;
        case 4: tile_copy<unsigned int, NT>(p); break;
        case 2: tile_copy<unsigned short, NT>(p); break;
        default:
            for (size_t i = threadIdx.x; i < p.nbytes; i += blockDim.x) p.dst[i] = p.src[i];
            break;
        }
    }
}

} // namespace nvshmemi

using namespace nvshmemi;

struct nvshmemx_batch_plan {
    int nsets = 1;
    int wgs_per_set = 1;
    int block = 256;
    int nontemporal = NT_BOTH;
    size_t npieces = 0;     /* pieces of the lists and tiles */
    size_t tile0 = 0;       /* first tile in d_pieces; pieces before it belong to the lists */
    size_t ntiles = 0;
    size_t total_bytes = 0;
    size_t piece_capacity = 0;
    size_t wg_capacity = 0;
    size_t wave_capacity = 0;
    piece_t *d_pieces = nullptr;
    unsigned int *d_wg_off = nullptr;
    unsigned int *d_wave_off = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    plan_layout layout; /* host arrays, reused across rebuilds */
};

namespace nvshmemi {

static void ensure_capacity(void **d, size_t &cap, size_t need, size_t elem) {
    if (need <= cap) return;
    if (*d) NVSHMEMI_HIP(hipFree(*d));
    cap = std::max<size_t>(need, 1024);
    NVSHMEMI_HIP(hipMalloc(d, cap * elem));
}

/* (Re)build plan `p` from resolved requests; device buffers are reused when
 * they are large enough, so the autotuner can rebuild per candidate without
 * allocating.  wgs_per_set <= 0 sizes the grid to the pieces (at most the
 * default grid).  ASSIGN_LPT balances uneven pieces; ASSIGN_ROUND_ROBIN deals
 * pieces in request order (same-size pieces); ASSIGN_SCATTERED deals them in
 * a fixed pseudo-random order, so that pieces running at the same time are not
 * address-adjacent -- as in a mixed batch, whose LPT order interleaves the
 * pieces of unrelated requests. */
void plan_fill(nvshmemx_batch_plan *p, const std::vector<resolved_req> &reqs, const nvshmemx_batch_choice_t *choices,
               const nvshmemx_batch_choice_t *one_choice, int nsets, int wgs_per_set, int block, int nontemporal,
               int assign) {
    plan_layout &L = p->layout;
    build_layout(reqs, choices, one_choice, nsets, wgs_per_set, block, assign, L);
    wgs_per_set = L.wgs_per_set;
    const size_t G = (size_t)nsets * wgs_per_set, W = G * (size_t)(block / 64);
    const std::vector<piece_t> &all = L.pieces;
    const std::vector<unsigned int> &wo = L.wg_off, &vo = L.wave_off;
    const size_t total = L.total_bytes;
    if (all.size() >= 0xFFFFFFF0u) NVSHMEMI_FATAL("batch of %zu pieces is too large", all.size());
    ensure_capacity((void **)&p->d_pieces, p->piece_capacity, all.size(), sizeof(piece_t));
    ensure_capacity((void **)&p->d_wg_off, p->wg_capacity, G + 1, sizeof(unsigned int));
    ensure_capacity((void **)&p->d_wave_off, p->wave_capacity, W + 1, sizeof(unsigned int));
    if (!all.empty())
        NVSHMEMI_HIP(hipMemcpy(p->d_pieces, all.data(), all.size() * sizeof(piece_t), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemcpy(p->d_wg_off, wo.data(), (G + 1) * sizeof(unsigned int), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemcpy(p->d_wave_off, vo.data(), (W + 1) * sizeof(unsigned int), hipMemcpyHostToDevice));
    p->nsets = nsets;
    p->wgs_per_set = wgs_per_set;
    p->block = block;
    p->nontemporal = nontemporal & 7;
    p->npieces = all.size();
    p->tile0 = L.tile0;
    p->ntiles = L.ntiles;
    p->total_bytes = total;
}

nvshmemx_batch_plan *plan_new() { return new nvshmemx_batch_plan(); }

/* one workgroup per tile, the grid's threads within 32 bits */
template <int NT> static void launch_tiles(const nvshmemx_batch_plan *p, hipStream_t s) {
    const size_t grid = std::min<size_t>(p->ntiles, 0xFFFFFFFFu / kTileBlock);
    hipLaunchKernelGGL(k_batch_tiles<NT>, dim3((unsigned)grid), dim3(kTileBlock), 0, s, p->d_pieces + p->tile0,
                       p->ntiles);
}

void plan_launch(nvshmemx_batch_plan *p, hipStream_t s) {
    if (p->ntiles) { /* the tiles' launch; a plan of tiles only is this one launch */
        switch (p->nontemporal) {
        case NT_BOTH: launch_tiles<NT_BOTH>(p, s); break;
        case NT_LOAD: launch_tiles<NT_LOAD>(p, s); break;
        case NT_STORE: launch_tiles<NT_STORE>(p, s); break;
        case ST_SC1:
        case ST_SC1 | NT_STORE: launch_tiles<ST_SC1>(p, s); break;
        case ST_SC1 | NT_LOAD:
        case ST_SC1 | NT_BOTH: launch_tiles<ST_SC1 | NT_LOAD>(p, s); break;
        default: launch_tiles<NT_NONE>(p, s); break;
        }
        NVSHMEMI_HIP(hipGetLastError());
        if (p->tile0 == 0) return;
    }
    const unsigned grid = (unsigned)(p->nsets * p->wgs_per_set);
    switch (p->nontemporal) {
    case NT_BOTH:
        hipLaunchKernelGGL(k_batch<NT_BOTH>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off, p->d_wave_off);
        break;
    case NT_LOAD:
        hipLaunchKernelGGL(k_batch<NT_LOAD>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off, p->d_wave_off);
        break;
    case NT_STORE:
        hipLaunchKernelGGL(k_batch<NT_STORE>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off,
                           p->d_wave_off);
        break;
    case ST_SC1:
    case ST_SC1 | NT_STORE: /* write-through takes precedence over non-temporal stores */
        hipLaunchKernelGGL(k_batch<ST_SC1>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off, p->d_wave_off);
        break;
    case ST_SC1 | NT_LOAD:
    case ST_SC1 | NT_BOTH:
        hipLaunchKernelGGL(k_batch<ST_SC1 | NT_LOAD>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off,
                           p->d_wave_off);
        break;
    default:
        hipLaunchKernelGGL(k_batch<NT_NONE>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_wg_off, p->d_wave_off);
    }
    NVSHMEMI_HIP(hipGetLastError());
}

} // namespace nvshmemi

extern "C" void nvshmemx_batch_default_choice(size_t nbytes, int elem_bytes, nvshmemx_batch_choice_t *out) {
    (void)elem_bytes;
    *out = default_choice(nbytes);
}

static int source_pe(const void *src) {
    if (g.mode != NVSHMEMI_MODE_LOCAL) return g.mype;
    const char *s = (const char *)src;
    if (s < g.heap_base[0] || s >= g.heap_base[0] + g.heap_stride * (size_t)g.npes) return 0;
    return (int)((size_t)(s - g.heap_base[0]) / g.heap_stride);
}

extern "C" int nvshmemx_batch_plan_create(const nvshmemx_put_request_t *reqs, int nreq,
                                          const nvshmemx_batch_choice_t *choices,
                                          const nvshmemx_batch_launch_t *launch, nvshmemx_batch_plan_t *out) {
    require_init("nvshmemx_batch_plan_create");
    if (nreq < 0 || (nreq > 0 && !reqs) || !out) NVSHMEMI_FATAL("nvshmemx_batch_plan_create: bad arguments");
    nvshmemx_batch_launch_t L = {0, kBatchBlock, kBatchNontemporal, NVSHMEMX_BATCH_SOURCE_ORDER};
    if (launch) {
        if (launch->workgroups > 0) L.workgroups = launch->workgroups;
        if (launch->block > 0) L.block = launch->block;
        L.nontemporal = launch->nontemporal;
        L.discipline = launch->discipline;
    }
    if (L.block < 64 || L.block > 1024 || L.block % 64)
        NVSHMEMI_FATAL("batch block size %d must be a multiple of 64 in [64, 1024]", L.block);
    if (L.discipline < NVSHMEMX_BATCH_SHARED || L.discipline > NVSHMEMX_BATCH_SOURCE_ORDER)
        NVSHMEMI_FATAL("unknown batch discipline %d", L.discipline);
    const bool owner = L.discipline == NVSHMEMX_BATCH_OWNER;
    if (owner && g.mode != NVSHMEMI_MODE_LOCAL)
        NVSHMEMI_FATAL("the owner discipline needs local (workgroup-partitioned) PEs");
    const int nsets = owner ? g.npes : 1;
    std::vector<resolved_req> rr((size_t)nreq);
    for (int i = 0; i < nreq; ++i) {
        const nvshmemx_put_request_t &r = reqs[i];
        if (r.pe < 0 || r.pe >= g.npes) NVSHMEMI_FATAL("request %d: PE %d out of range [0, %d)", i, r.pe, g.npes);
        check_symmetric(r.target, r.nbytes, "nvshmemx_batch_plan_create");
        rr[(size_t)i] = resolved_req{(const char *)r.source, (char *)nvshmem_ptr(r.target, r.pe), r.nbytes,
                                     owner ? source_pe(r.source) : 0};
    }
    const int wgs_per_set = owner ? (L.workgroups > 0 ? std::max(1, L.workgroups / nsets) : g.wgs_per_pe)
                                  : (L.workgroups > 0 ? L.workgroups : 0 /* sized to the work */);
    nvshmemx_batch_plan *p = new nvshmemx_batch_plan();
    plan_fill(p, rr, choices, nullptr, nsets, wgs_per_set, L.block, L.nontemporal,
              L.discipline == NVSHMEMX_BATCH_SOURCE_ORDER ? ASSIGN_SOURCE_ORDER : ASSIGN_LPT);
    *out = p;
    return 0;
}

extern "C" void nvshmemx_batch_plan_launch(nvshmemx_batch_plan_t p, hipStream_t s) {
    if (!p) NVSHMEMI_FATAL("nvshmemx_batch_plan_launch: null plan");
    plan_launch(p, s);
}

extern "C" int nvshmemx_batch_plan_capture(nvshmemx_batch_plan_t p, int steps) {
    if (!p || steps < 1) NVSHMEMI_FATAL("nvshmemx_batch_plan_capture: bad arguments");
    if (p->exec) NVSHMEMI_HIP(hipGraphExecDestroy(p->exec));
    if (p->graph) NVSHMEMI_HIP(hipGraphDestroy(p->graph));
    if (!p->capture_stream) NVSHMEMI_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
    NVSHMEMI_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) plan_launch(p, p->capture_stream);
    NVSHMEMI_HIP(hipStreamEndCapture(p->capture_stream, &p->graph));
    NVSHMEMI_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return 0;
}

extern "C" void nvshmemx_batch_plan_graph_launch(nvshmemx_batch_plan_t p, hipStream_t s) {
    if (!p || !p->exec) NVSHMEMI_FATAL("nvshmemx_batch_plan_graph_launch: plan has no captured graph");
    NVSHMEMI_HIP(hipGraphLaunch(p->exec, s));
}

/* Select the fastest of plans built for the same exchange: each plan is
 * launched once, timed, to warm up; then `reps` rounds time every plan in
 * turn over a burst of back-to-back launches between two HIP events on `s`,
 * the burst long enough (about 50 us, at most 64 launches) that a batch of a
 * few microseconds is not lost in the events' resolution and launch gaps;
 * the plan with the smallest median time per launch wins.  Every launch
 * executes the plan's puts (idempotent: the same sources to the same
 * targets), so the caller selects when the sources hold what the exchange
 * sends.  ms_out (optional, nplans entries) receives each plan's median time
 * per launch. */
extern "C" int nvshmemx_batch_plan_select(const nvshmemx_batch_plan_t *plans, int nplans, int reps, hipStream_t s,
                                         float *ms_out) {
    require_init("nvshmemx_batch_plan_select");
    if (!plans || nplans < 1 || reps < 1) NVSHMEMI_FATAL("nvshmemx_batch_plan_select: bad arguments");
    for (int i = 0; i < nplans; ++i)
        if (!plans[i]) NVSHMEMI_FATAL("nvshmemx_batch_plan_select: plan %d is null", i);
    constexpr float kBurstMs = 0.05f;
    constexpr int kMaxBurst = 64;
    const size_t nev = (size_t)nplans * (size_t)(reps + 1); /* the warm-up round, then `reps` rounds */
    std::vector<hipEvent_t> ev(2 * nev);
    for (hipEvent_t &e : ev) NVSHMEMI_HIP(hipEventCreate(&e));
    auto timed_round = [&](int r, int burst) {
        for (int i = 0; i < nplans; ++i) {
            const size_t k = (size_t)r * (size_t)nplans + (size_t)i;
            NVSHMEMI_HIP(hipEventRecord(ev[2 * k], s));
            for (int b = 0; b < burst; ++b) plan_launch(plans[i], s);
            NVSHMEMI_HIP(hipEventRecord(ev[2 * k + 1], s));
        }
    };
    auto elapsed = [&](int r, int i) {
        const size_t k = (size_t)r * (size_t)nplans + (size_t)i;
        float ms = 0.0f;
        NVSHMEMI_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        return ms;
    };
    timed_round(0, 1);
    NVSHMEMI_HIP(hipEventSynchronize(ev[2 * (size_t)nplans - 1]));
    float fastest = elapsed(0, 0);
    for (int i = 1; i < nplans; ++i) fastest = std::min(fastest, elapsed(0, i));
    const int burst = fastest > 0.0f ? (int)std::min<float>((float)kMaxBurst, std::ceil(kBurstMs / fastest)) : kMaxBurst;
    for (int r = 1; r <= reps; ++r) timed_round(r, std::max(1, burst));
    NVSHMEMI_HIP(hipEventSynchronize(ev.back()));
    int best = 0;
    float best_ms = 0.0f;
    std::vector<float> t((size_t)reps);
    for (int i = 0; i < nplans; ++i) {
        for (int r = 1; r <= reps; ++r) t[(size_t)r - 1] = elapsed(r, i) / (float)std::max(1, burst);
        std::sort(t.begin(), t.end());
        const float med = t[(size_t)reps / 2];
        if (ms_out) ms_out[i] = med;
        if (i == 0 || med < best_ms) best = i, best_ms = med;
    }
    for (hipEvent_t &e : ev) NVSHMEMI_HIP(hipEventDestroy(e));
    return best;
}

extern "C" size_t nvshmemx_batch_plan_pieces(nvshmemx_batch_plan_t p) { return p ? p->npieces : 0; }
extern "C" size_t nvshmemx_batch_plan_bytes(nvshmemx_batch_plan_t p) { return p ? p->total_bytes : 0; }

extern "C" void nvshmemx_batch_plan_destroy(nvshmemx_batch_plan_t p) {
    if (!p) return;
    NVSHMEMI_HIP(hipDeviceSynchronize());
    if (p->exec) NVSHMEMI_HIP(hipGraphExecDestroy(p->exec));
    if (p->graph) NVSHMEMI_HIP(hipGraphDestroy(p->graph));
    if (p->capture_stream) NVSHMEMI_HIP(hipStreamDestroy(p->capture_stream));
    if (p->d_pieces) NVSHMEMI_HIP(hipFree(p->d_pieces));
    if (p->d_wg_off) NVSHMEMI_HIP(hipFree(p->d_wg_off));
    if (p->d_wave_off) NVSHMEMI_HIP(hipFree(p->d_wave_off));
    delete p;
}

/* The plain path the batch is measured against: one grid put launch per
 * request, each with the size-based default launch shape. */
extern "C" void nvshmemx_put_requests_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t s) {
    for (int i = 0; i < nreq; ++i)
        nvshmemx_putmem_on_stream(reqs[i].target, reqs[i].source, reqs[i].nbytes, reqs[i].pe, s);
}
