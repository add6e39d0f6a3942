/*
 * batch.hip -- fused execution of many puts of uneven size in one launch.
 *
 * No counterpart in the reference, whose puts are one call per thread
 * (nvshmem_device.cu:487-512).  A batch of put requests is turned into a plan:
 *
 *  - each request is split into pieces by its choice (pieces x vector width x
 *    wavefront/workgroup mode); piece boundaries fall on 256-byte offsets of the
 *    request, so every piece keeps the request's relative alignment;
 *  - pieces go into device queues (one workgroup queue and one wavefront queue
 *    per queue set), each sorted by size, largest first, so the dynamic
 *    schedule below approximates longest-processing-time-first;
 *  - one persistent kernel drains the queues: a workgroup takes whole
 *    workgroup pieces (one atomic ticket per piece, broadcast through LDS),
 *    then each wavefront takes wavefront pieces independently;
 *  - the ticket counters reset themselves: the last workgroup of a queue set
 *    to finish zeroes them, so a plan replays from a captured hipGraph with no
 *    memset node.
 *
 * Queue sets: SHARED = one set served by every workgroup; OWNER = one set per
 * PE in local mode, served only by that PE's workgroups (workgroup-partitioned
 * PEs, each executing the puts it initiated).
 */
#include <algorithm>
#include <vector>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

namespace nvshmemi {

struct piece_t {
    const char *src;
    char *dst;
    unsigned long long nbytes;
    unsigned int vec;
    unsigned int pad;
};

struct queue_t {
    unsigned int wg_begin, wg_end, wave_begin, wave_end;
};


constexpr int kCtrStride = 16; /* counters of one queue set: 64 bytes apart */

template <int NT>
__device__ __forceinline__ void copy_piece(const piece_t &p, unsigned lane, unsigned n) {
    switch (p.vec) {
    case 16: copy<16, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    case 8: copy<8, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    default: copy<4, 4, NT>(p.dst, p.src, p.nbytes, lane, n); break;
    }
}

template <int NT>
__global__ void __launch_bounds__(1024) k_batch(const piece_t *pieces, const queue_t *queues, unsigned int *ctr,
                                                int wgs_per_set) {
    const int set = (int)blockIdx.x / wgs_per_set;
    const queue_t q = queues[set];
    unsigned int *c = ctr + set * kCtrStride;
    __shared__ unsigned int s_ticket;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(&c[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const unsigned int j = q.wg_begin + s_ticket;
        __syncthreads();
        if (j >= q.wg_end) break;
        copy_piece<NT>(pieces[j], threadIdx.x, blockDim.x);
    }
    const unsigned int lane = threadIdx.x & 63u;
    for (;;) {
        unsigned int t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(&c[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = __builtin_amdgcn_readfirstlane(t);
        const unsigned int j = q.wave_begin + t;
        if (j >= q.wave_end) break;
        copy_piece<NT>(pieces[j], lane, 64);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int done = __hip_atomic_fetch_add(&c[2], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done == (unsigned int)wgs_per_set - 1) {
            __hip_atomic_store(&c[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c[2], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

/* one request's pieces under a choice; boundaries on 256-byte offsets */
static void split_request(const char *src, char *dst, size_t nbytes, const nvshmemx_batch_choice_t &ch,
                          std::vector<piece_t> &out) {
    if (nbytes == 0) return;
    const size_t quantum = 256;
    size_t want = (size_t)std::max(1, std::min(ch.pieces, 256));
    if (ch.mode == 0) want *= 4; /* a workgroup's worth of wavefronts per piece */
    const size_t max_parts = std::max<size_t>(1, nbytes / quantum);
    const size_t parts = std::min(want, max_parts);
    size_t per = (nbytes + parts - 1) / parts;
    per = (per + quantum - 1) / quantum * quantum;
    const unsigned vec = ch.vec_bytes >= 16 ? 16 : (ch.vec_bytes >= 8 ? 8 : 4);
    for (size_t off = 0; off < nbytes; off += per) {
        const size_t n = std::min(per, nbytes - off);
        out.push_back(piece_t{src + off, dst + off, (unsigned long long)n, vec, 0});
    }
}

} // namespace nvshmemi

using namespace nvshmemi;

struct nvshmemx_batch_plan {
    int nsets = 1;
    int wgs_per_set = 1;
    int block = 256;
    int nontemporal = NT_BOTH;
    size_t npieces = 0;
    size_t total_bytes = 0;
    size_t piece_capacity = 0;
    int set_capacity = 0;
    piece_t *d_pieces = nullptr;
    queue_t *d_queues = nullptr;
    unsigned int *d_ctr = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    std::vector<piece_t> host_pieces; /* staging, reused across rebuilds */
};

namespace nvshmemi {

/* (Re)build plan `p` from resolved requests; device buffers are reused when
 * they are large enough, so the autotuner can rebuild per candidate without
 * allocating. */
void plan_fill(nvshmemx_batch_plan *p, const std::vector<resolved_req> &reqs, const nvshmemx_batch_choice_t *choices,
               const nvshmemx_batch_choice_t *one_choice, int nsets, int wgs_per_set, int block, int nontemporal,
               bool sort_pieces) {
    std::vector<std::vector<piece_t>> wg(nsets), wave(nsets);
    size_t total = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const resolved_req &r = reqs[i];
        nvshmemx_batch_choice_t ch;
        if (one_choice) ch = *one_choice;
        else if (choices) ch = choices[i];
        else nvshmemx_batch_default_choice(r.nbytes, 4, &ch);
        split_request(r.src, r.dst, r.nbytes, ch, ch.mode ? wg[r.set] : wave[r.set]);
        total += r.nbytes;
    }
    std::vector<piece_t> &all = p->host_pieces;
    all.clear();
    std::vector<queue_t> queues(nsets);
    auto by_size = [](const piece_t &a, const piece_t &b) { return a.nbytes > b.nbytes; };
    for (int s = 0; s < nsets; ++s) {
        if (sort_pieces) {
            std::stable_sort(wg[s].begin(), wg[s].end(), by_size);
            std::stable_sort(wave[s].begin(), wave[s].end(), by_size);
        }
        queues[s].wg_begin = (unsigned)all.size();
        all.insert(all.end(), wg[s].begin(), wg[s].end());
        queues[s].wg_end = queues[s].wave_begin = (unsigned)all.size();
        all.insert(all.end(), wave[s].begin(), wave[s].end());
        queues[s].wave_end = (unsigned)all.size();
    }
    if (all.size() >= 0xFFFFFFF0u) NVSHMEMI_FATAL("batch of %zu pieces is too large", all.size());
    if (all.size() > p->piece_capacity) {
        if (p->d_pieces) NVSHMEMI_HIP(hipFree(p->d_pieces));
        p->piece_capacity = std::max<size_t>(all.size(), 1024);
        NVSHMEMI_HIP(hipMalloc((void **)&p->d_pieces, p->piece_capacity * sizeof(piece_t)));
    }
    if (nsets > p->set_capacity) {
        if (p->d_queues) NVSHMEMI_HIP(hipFree(p->d_queues));
        if (p->d_ctr) NVSHMEMI_HIP(hipFree(p->d_ctr));
        p->set_capacity = nsets;
        NVSHMEMI_HIP(hipMalloc((void **)&p->d_queues, nsets * sizeof(queue_t)));
        NVSHMEMI_HIP(hipMalloc((void **)&p->d_ctr, (size_t)nsets * kCtrStride * sizeof(unsigned int)));
        NVSHMEMI_HIP(hipMemset(p->d_ctr, 0, (size_t)nsets * kCtrStride * sizeof(unsigned int)));
    }
    if (!all.empty())
        NVSHMEMI_HIP(hipMemcpy(p->d_pieces, all.data(), all.size() * sizeof(piece_t), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemcpy(p->d_queues, queues.data(), nsets * sizeof(queue_t), hipMemcpyHostToDevice));
    p->nsets = nsets;
    p->wgs_per_set = wgs_per_set;
    p->block = block;
    p->nontemporal = nontemporal & 3;
    p->npieces = all.size();
    p->total_bytes = total;
}

} // namespace nvshmemi

extern "C" void nvshmemx_batch_default_choice(size_t nbytes, int elem_bytes, nvshmemx_batch_choice_t *out) {
    /* the rule used when no tuned choice is given: ~256 KiB per workgroup
     * piece, wavefront pieces below 64 KiB, widest access the element allows */
    nvshmemx_batch_choice_t c;
    c.mode = nbytes >= (64u << 10) ? 1 : 0;
    const size_t per = c.mode ? (256u << 10) : (16u << 10);
    size_t p = (nbytes + per - 1) / per;
    c.pieces = (int)std::max<size_t>(1, std::min<size_t>(p, 256));
    c.vec_bytes = 16;
    (void)elem_bytes;
    *out = c;
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
    nvshmemx_batch_launch_t L = {0, 256, NT_BOTH, NVSHMEMX_BATCH_SHARED};
    if (launch) {
        if (launch->workgroups > 0) L.workgroups = launch->workgroups;
        if (launch->block > 0) L.block = launch->block;
        L.nontemporal = launch->nontemporal;
        L.discipline = launch->discipline;
    }
    if (L.block < 64 || L.block > 1024 || L.block % 64)
        NVSHMEMI_FATAL("batch block size %d must be a multiple of 64 in [64, 1024]", L.block);
    const bool owner = L.discipline == NVSHMEMX_BATCH_OWNER;
    if (owner && g.mode != NVSHMEMI_MODE_LOCAL)
        NVSHMEMI_FATAL("the owner discipline needs local (workgroup-partitioned) PEs");
    const int nsets = owner ? g.npes : 1;
    std::vector<resolved_req> rr((size_t)nreq);
    for (int i = 0; i < nreq; ++i) {
        const nvshmemx_put_request_t &r = reqs[i];
        if (r.pe < 0 || r.pe >= g.npes) NVSHMEMI_FATAL("request %d: PE %d out of range [0, %d)", i, r.pe, g.npes);
        rr[(size_t)i] = resolved_req{(const char *)r.source, (char *)nvshmem_ptr(r.target, r.pe), r.nbytes,
                                     owner ? source_pe(r.source) : 0};
    }
    const int wgs_per_set = owner ? (L.workgroups > 0 ? std::max(1, L.workgroups / nsets) : g.wgs_per_pe)
                                  : (L.workgroups > 0 ? L.workgroups : 1024);
    nvshmemx_batch_plan *p = new nvshmemx_batch_plan();
    plan_fill(p, rr, choices, nullptr, nsets, wgs_per_set, L.block, L.nontemporal, true);
    *out = p;
    return 0;
}

nvshmemx_batch_plan *nvshmemi::plan_new() { return new nvshmemx_batch_plan(); }

static void launch_plan(nvshmemx_batch_plan_t p, hipStream_t s) { nvshmemi::plan_launch(p, s); }

void nvshmemi::plan_launch(nvshmemx_batch_plan *p, hipStream_t s) {
    const unsigned grid = (unsigned)(p->nsets * p->wgs_per_set);
    switch (p->nontemporal) {
    case NT_BOTH:
        hipLaunchKernelGGL(k_batch<NT_BOTH>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_queues, p->d_ctr,
                           p->wgs_per_set);
        break;
    case NT_LOAD:
        hipLaunchKernelGGL(k_batch<NT_LOAD>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_queues, p->d_ctr,
                           p->wgs_per_set);
        break;
    case NT_STORE:
        hipLaunchKernelGGL(k_batch<NT_STORE>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_queues, p->d_ctr,
                           p->wgs_per_set);
        break;
    default:
        hipLaunchKernelGGL(k_batch<NT_NONE>, dim3(grid), dim3(p->block), 0, s, p->d_pieces, p->d_queues, p->d_ctr,
                           p->wgs_per_set);
    }
    NVSHMEMI_HIP(hipGetLastError());
}

extern "C" void nvshmemx_batch_plan_launch(nvshmemx_batch_plan_t p, hipStream_t s) {
    if (!p) NVSHMEMI_FATAL("nvshmemx_batch_plan_launch: null plan");
    launch_plan(p, s);
}

extern "C" int nvshmemx_batch_plan_capture(nvshmemx_batch_plan_t p, int steps) {
    if (!p || steps < 1) NVSHMEMI_FATAL("nvshmemx_batch_plan_capture: bad arguments");
    if (p->exec) NVSHMEMI_HIP(hipGraphExecDestroy(p->exec));
    if (p->graph) NVSHMEMI_HIP(hipGraphDestroy(p->graph));
    if (!p->capture_stream) NVSHMEMI_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
    NVSHMEMI_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) launch_plan(p, p->capture_stream);
    NVSHMEMI_HIP(hipStreamEndCapture(p->capture_stream, &p->graph));
    NVSHMEMI_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return 0;
}

extern "C" void nvshmemx_batch_plan_graph_launch(nvshmemx_batch_plan_t p, hipStream_t s) {
    if (!p || !p->exec) NVSHMEMI_FATAL("nvshmemx_batch_plan_graph_launch: plan has no captured graph");
    NVSHMEMI_HIP(hipGraphLaunch(p->exec, s));
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
    if (p->d_queues) NVSHMEMI_HIP(hipFree(p->d_queues));
    if (p->d_ctr) NVSHMEMI_HIP(hipFree(p->d_ctr));
    delete p;
}

/* The plain path the batch is measured against: one grid put launch per
 * request, each with the size-based default launch shape. */
extern "C" void nvshmemx_put_requests_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t s) {
    for (int i = 0; i < nreq; ++i)
        nvshmemx_putmem_on_stream(reqs[i].target, reqs[i].source, reqs[i].nbytes, reqs[i].pe, s);
}
