/*
 * persistent.hip -- a batch of puts planned on the device by the PEs that
 * issued them and executed by one persistent kernel.
 *
 * No counterpart in the reference, whose puts are one call per thread
 * (nvshmem_device.cu:487-512).  The fused plans of batch.hip are cut into
 * pieces and dealt to workers on the host; here the device does that itself,
 * every launch, from the request list as it is:
 *
 *  - the grid holds only as many workgroups as the device runs at once, so
 *    every workgroup is resident for the whole launch.  Workgroups 0 .. P-1
 *    (P = the local PEs, or 1 in process mode) -- dispatched first -- are
 *    the PEs' leaders;
 *  - each leader reads the whole request list and takes the requests its PE
 *    issued (those whose source lies in its heap; a private source counts for
 *    PE i mod P): it lays them end to end in a virtual byte stream, each
 *    padded to a 4 KiB granule, writes its entries (request index, start in
 *    the stream) and, for every chunk of the stream, the entry the chunk
 *    starts in; then publishes its queue with a release flag.  Leaders agree
 *    on where each PE's entries and chunks go by counting every PE's requests
 *    in the same pass, so no leader waits for another;
 *  - a PE with at most 64 requests and 128 KiB of payload gets no queue: its
 *    leader publishes an empty one and copies the requests itself, so a
 *    small batch is moved by the PEs that issued it, with no chunk to claim;
 *  - once every queue is published, the workgroups copy 128 KiB chunks of
 *    the queues laid end to end in PE order: workgroup b takes chunk b, then
 *    claims more with one atomic fetch-add on one counter each (the next
 *    claim is issued before the current chunk is copied), so the workgroups
 *    running at once copy neighbouring chunks of the same few PEs' queues.
 *    Inside a chunk, a request segment of 16 KiB or more is copied by the
 *    whole workgroup, shorter ones by one wavefront each, in turn, with
 *    16-byte accesses where the alignment allows, eight loads in flight per
 *    lane and non-temporal stores;
 *  - nothing is cleared between launches: each workgroup takes a launch
 *    ticket, flags carry the launch's number, and the claim counters
 *    alternate between launches, each launch zeroing the next one's; so the
 *    next launch (or graph replay) starts clean with no memset and no
 *    workgroup waits for the others to finish.
 *
 * Contention and coherence: each PE's ready flag and queue record sit in a
 * 4 KiB slot of their own.  Readers never take an agent-scope acquire (which
 * invalidates the XCD's L2 under the streams the other workgroups are
 * running): the leader publishes with one release, and everything it
 * published is read with coherent (agent-scope relaxed) loads after the flag
 * was seen to carry this launch's number.
 *
 * Launch shape, from measurements of the 4096-put step of 64 PEs (1.57 GB
 * payload, ms per step, before the leader and chunk changes below): one
 * workgroup of 1024 threads per CU, chunks of 128 KiB, eight loads in flight
 * per lane, 0.528; chunks of 64, 256 or 512 KiB
 * 0.584, 0.540, 0.560; sixteen loads in flight 0.562; two workgroups of 512
 * threads per CU 0.591, three of 512 threads with four loads in flight 0.659.
 * Fewer, wider workgroups keep fewer address streams open at once.  Each
 * workgroup first draining its own PE's queue and then stealing from the
 * longest (the PE-owned order, 64 source windows at once) took 0.636 at two
 * workgroups of 512 threads per CU against 0.591 for the PE-ordered claims;
 * every wavefront working alone on 32 KiB chunks 1.07 (each of ~8000
 * concurrent streams waited ~12 us per access: the DRAM pages thrash); claim
 * counters sharing one cache line and an acquire per poll, 1.49.  Stamped
 * from the constant clock, the leaders took 6-8 us for 16 requests and 15-18
 * us for 4096 while thread 0 summed every PE's chunks with a 64-bit division
 * each; wavefront 0 now does that with shuffles (4 and 13-17 us), and a chunk
 * reads its entries 64 at a time instead of one dependent pair of loads per
 * entry: the step went from 0.528 to 0.512, a batch of 16 requests of 64 B
 * from 23.5 to 17.1 us.  Launch tickets instead of a last-workgroup
 * clean-up, a first chunk taken without an atomic, and leaders copying
 * small PEs' requests themselves then took the step to 0.491 and 16
 * requests of 64 B to 10.4 us (128 such requests 10.6, 1024 12.1, 4096
 * 38.9).  The host-planned fused launch of batch.hip (source-ordered pieces,
 * no device-side planning) takes 0.509 for the same step and about 3.5 us
 * for a tiny batch.
 *
 * Scratch (per launch, device memory): the control slots, the chunk table
 * (65,536 entries; the chunk grows from 128 KiB so that a batch of any size
 * fits) and one 8-byte entry per request.
 * nvshmemx_batch_put_persistent_on_stream allocates it (grown on demand);
 * nvshmemx_batch_put_persistent_noalloc_on_stream uses a static array of this
 * file instead (up to 65,536 requests, 1.25 MiB per GPU, held until the
 * process ends), for callers that cannot allocate.  Calls that overlap on one
 * GPU share the scratch and race.
 */
#include <algorithm>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

namespace nvshmemi {

#ifndef NVSHMEMI_PCHUNK_UNITS
#define NVSHMEMI_PCHUNK_UNITS 32
#endif
#ifndef NVSHMEMI_PUNROLL
#define NVSHMEMI_PUNROLL 8
#endif
constexpr int kPBlock = 1024;
constexpr int kPWaves = kPBlock / 64;
constexpr unsigned kUnit = 4096;                           /* granule of the virtual stream, bytes */
constexpr unsigned kChunkUnitsMin = NVSHMEMI_PCHUNK_UNITS; /* chunk unless the table needs larger */
constexpr size_t kWaveSegMax = 16384; /* shorter request segments go to one wavefront */
constexpr int kPUnroll = NVSHMEMI_PUNROLL;                 /* loads in flight per lane */
constexpr unsigned kTableChunks = 65536;                   /* chunk-table entries */
constexpr unsigned kLeadCache = 4096;                      /* requests a leader keeps in LDS between passes */
constexpr int kMaxReqsNoalloc = 65536;                     /* requests per call without allocation */
constexpr int kPPerCU = 1;                                                /* workgroups per CU (see the launch shape above) */
constexpr unsigned kSelfCount = 64;      /* a PE with at most this many requests ... */
constexpr size_t kSelfBytes = 128 << 10; /* ... and payload bytes is copied by its own leader */
static_assert(kSelfCount <= 64, "a leader reads its own requests one per lane");
constexpr int kMaxDevices = 64;

struct pinfo {            /* PE p's queue, published by its leader */
    unsigned entry0;      /* first entry */
    unsigned nown;        /* entries */
    unsigned chunk0;      /* first chunk-table slot */
    unsigned nchunks;
    unsigned chunk_units; /* chunk size in granules, the same for every PE */
    unsigned pad[3];
};
struct pentry {
    unsigned req;    /* request index */
    unsigned vstart; /* start in the PE's virtual stream, granules */
};

/* scratch layout: per-PE 4 KiB slots (ready flag, queue record), the
 * launch ticket counter and the two claim counters, the chunk table, the
 * entries */
constexpr size_t kSlot = 4096;
constexpr size_t kReadyOff = 128, kInfoOff = 256;
constexpr size_t kTicketOff = kSlot * NVSHMEMI_MAX_PES;
constexpr size_t kClaimOff = kTicketOff + 128; /* claim counter of even launches; odd ones 128 B further */
constexpr size_t kCtlBytes = kTicketOff + 384; /* zeroed when the scratch is set up */
constexpr size_t kTableOff = kCtlBytes;
constexpr size_t kEntryOff = kTableOff + sizeof(unsigned) * kTableChunks;
constexpr size_t scratch_bytes(size_t nreq) { return kEntryOff + sizeof(pentry) * (nreq ? nreq : 1); }

/* the no-allocation entry point's scratch: each GPU holds its own copy.  In
 * host code this name is only a stand-in: kernels get the address that
 * hipGetSymbolAddress returned for the current device. */
static __device__ __align__(256) unsigned char scratch[scratch_bytes(kMaxReqsNoalloc)];

/* the posted puts (nvshmemx_putmem_nbi) and the scratch of their flush, kept
 * apart from `scratch` so that a flush and a no-allocation batch do not share
 * control slots; like `scratch`, host code reaches them only through
 * hipGetSymbolAddress */
__device__ nvshmemx_put_request_t nbi_queue[kNbiSlots];
__device__ unsigned nbi_count[NVSHMEMI_MAX_PES];
static __device__ __align__(256) unsigned char nbi_scratch[scratch_bytes(kNbiSlots)];

struct pscr {
    unsigned char *base;
    __device__ __forceinline__ unsigned *ready(int q) const { return (unsigned *)(base + q * kSlot + kReadyOff); }
    __device__ __forceinline__ pinfo *info(int q) const { return (pinfo *)(base + q * kSlot + kInfoOff); }
    __device__ __forceinline__ unsigned long long *ticket() const { return (unsigned long long *)(base + kTicketOff); }
    __device__ __forceinline__ unsigned *claim(unsigned parity) const { return (unsigned *)(base + kClaimOff + 128 * parity); }
    __device__ __forceinline__ unsigned *table() const { return (unsigned *)(base + kTableOff); }
    __device__ __forceinline__ pentry *entries() const { return (pentry *)(base + kEntryOff); }
};

/* ------------------------------------------------------------- device -- */

__device__ __forceinline__ unsigned coherent_u32(const unsigned *p) { return load_relaxed(p); }
__device__ __forceinline__ pinfo coherent_info(const pinfo *p) {
    pinfo r;
    r.entry0 = load_relaxed(&p->entry0);
    r.nown = load_relaxed(&p->nown);
    r.chunk0 = load_relaxed(&p->chunk0);
    r.nchunks = load_relaxed(&p->nchunks);
    r.chunk_units = load_relaxed(&p->chunk_units);
    return r;
}
__device__ __forceinline__ pentry coherent_entry(const pentry *p) {
    const unsigned long long v = load_relaxed((const long long *)p);
    return pentry{(unsigned)v, (unsigned)(v >> 32)};
}

/* poll a ready flag with relaxed loads (back to back, then the sleep
 * ladder) until it holds this launch's tag or the wait budget runs out; no
 * acquire: what the flag publishes is read with coherent loads.  Returns
 * whether the tag was seen. */
__device__ __forceinline__ bool wait_tag(const unsigned *flag, unsigned tag) {
    bool seen = load_relaxed(flag) == tag;
    for (int i = 0; i < NVSHMEMI_SPIN_POLLS && !seen; ++i) seen = load_relaxed(flag) == tag;
    if (!seen) {
        const uint64_t budget = nvshmemi_dstate.wait_timeout_ticks;
        const uint64_t t0 = wall_clock64();
        sleep_ladder pace;
        while (!(seen = load_relaxed(flag) == tag)) {
            pace.pause();
            if ((uint64_t)(wall_clock64() - t0) > budget) {
                set_error(NVSHMEMI_ERR_WAIT_TIMEOUT);
                break;
            }
        }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST); /* no later load moves above the poll */
    return seen;
}

__device__ __forceinline__ int req_issuer(const nvshmemx_put_request_t &r, unsigned i, int P) {
    if (P == 1) return 0;
    const nvshmemi_device_state_t &s = nvshmemi_dstate;
    const uint64_t stride = s.offset_mask + 1;
    const uint64_t off = (uint64_t)((const char *)r.source - s.heap_base[0]);
    if (off < stride * (uint64_t)s.npes) return (int)(off / stride);
    return (int)(i % (unsigned)P);
}

/* granules a request occupies in its PE's stream; 0 (and an error bit, set
 * by leader 0) for a request naming a bad PE or a symmetric range outside
 * the heap (the target of a put, the source of a posted get), which is then
 * skipped */
__device__ __forceinline__ unsigned req_units(const nvshmemx_put_request_t &r, bool report, bool get = false) {
    const nvshmemi_device_state_t &s = nvshmemi_dstate;
    if ((unsigned)r.pe >= (unsigned)s.npes) {
        if (report) set_error(NVSHMEMI_ERR_BAD_PE);
        return 0;
    }
    const uint64_t rel = (uint64_t)((const char *)(get ? r.source : r.target) - s.anchor);
    const uint64_t off = rel & s.offset_mask;
    if (rel >= s.anchor_span || off >= s.heap_bytes || r.nbytes > s.heap_bytes - off) {
        if (report) set_error(NVSHMEMI_ERR_BAD_ADDRESS);
        return 0;
    }
    return (unsigned)((r.nbytes + kUnit - 1) / kUnit);
}

/* exclusive block scan of two counters over kPBlock threads; ta/tb get the
 * block totals.  Uses s_w[2][kPWaves]; every thread must call it. */
__device__ __forceinline__ void block_scan2(unsigned a, unsigned b, unsigned &xa, unsigned &xb, unsigned &ta,
                                            unsigned &tb, unsigned (*s_w)[kPWaves]) {
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    unsigned ia = a, ib = b;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned na = __shfl_up(ia, d, 64), nb = __shfl_up(ib, d, 64);
        if (lane >= (unsigned)d) ia += na, ib += nb;
    }
    if (lane == 63) s_w[0][w] = ia, s_w[1][w] = ib;
    __syncthreads();
    unsigned oa = 0, ob = 0;
    ta = 0, tb = 0;
#pragma unroll
    for (int k = 0; k < kPWaves; ++k) {
        if ((unsigned)k < w) oa += s_w[0][k], ob += s_w[1][k];
        ta += s_w[0][k], tb += s_w[1][k];
    }
    xa = oa + ia - a;
    xb = ob + ib - b;
    __syncthreads();
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t v, unsigned j) {
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, j), hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

/* a PE whose requests are few and small enough is served by its own leader:
 * it gets no queue, so the leaders and workers must all decide alike, from
 * the per-PE counts every leader takes */
__device__ __forceinline__ bool leader_served(unsigned cnt, unsigned long long bytes) {
    return cnt <= kSelfCount && bytes <= kSelfBytes;
}

/* A leader copies its PE's own requests (cnt <= kSelfCount of them, indices
 * in own[], all in LDS), one per lane of every wavefront: those of 16 KiB or
 * more by the whole workgroup, in turn; the shorter ones at once, one
 * wavefront each when there are at most 16 of them, else 16 lanes each (at
 * most 64 of them).  Every thread of the workgroup calls. */
template <bool POSTED>
__device__ __forceinline__ void copy_own(const nvshmemx_put_request_t *__restrict__ reqs, unsigned cnt, const unsigned *own) {
    const unsigned lane = threadIdx.x & 63u;
    uint64_t tgt = 0, src = 0, nb = 0;
    unsigned pe = 0, get = 0;
    if (lane < cnt) {
        const nvshmemx_put_request_t r = reqs[own[lane]];
        tgt = (uint64_t)r.target, src = (uint64_t)r.source, nb = (uint64_t)r.nbytes, pe = (unsigned)r.pe;
        get = POSTED && r.elem_bytes == kNbiGet;
    }
    for (unsigned j = 0; j < cnt; ++j) {
        const size_t n = (size_t)lane_u64(nb, j);
        if (n < kWaveSegMax) continue;
        const int jpe = (int)__builtin_amdgcn_readlane(pe, j);
        const bool jget = POSTED && __builtin_amdgcn_readlane(get, j);
        char *dst = jget ? (char *)lane_u64(tgt, j) : translate((const void *)lane_u64(tgt, j), jpe);
        const char *srcp = jget ? translate((const void *)lane_u64(src, j), jpe) : (const char *)lane_u64(src, j);
        copy<16, kPUnroll, NT_STORE>(dst, srcp, n, threadIdx.x, kPBlock);
    }
    uint64_t smask = __ballot(lane < cnt && nb < kWaveSegMax);
    const unsigned nsmall = (unsigned)__builtin_popcountll(smask);
    const unsigned group = nsmall <= (unsigned)kPWaves ? 64u : 16u; /* lanes per short request */
    const unsigned rank = threadIdx.x / group;
    /* the rank-th short request sits in lane j; every lane takes part in the
     * shuffles (a shuffle reads nothing from a lane that is switched off) */
    for (unsigned k = 0; k < (rank < nsmall ? rank : 0u); ++k) smask &= smask - 1;
    const int j = smask ? __builtin_ctzll(smask) : 0;
    const uint64_t jt = ((uint64_t)(unsigned)__shfl((int)(tgt >> 32), j, 64) << 32) | (unsigned)__shfl((int)tgt, j, 64);
    const uint64_t js = ((uint64_t)(unsigned)__shfl((int)(src >> 32), j, 64) << 32) | (unsigned)__shfl((int)src, j, 64);
    const size_t n = (size_t)(unsigned)__shfl((int)nb, j, 64); /* below 16 KiB */
    const int jpe = __shfl((int)pe, j, 64);
    const bool jget = POSTED && __shfl((int)get, j, 64);
    char *dst = jget ? (char *)jt : translate((const void *)jt, jpe);
    const char *srcp = jget ? translate((const void *)js, jpe) : (const char *)js;
    if (rank < nsmall) copy<16, kPUnroll, NT_STORE>(dst, srcp, n, threadIdx.x % group, group);
}

/* leader of PE p: build and publish the PE's queue, or, for a PE the leader
 * serves itself, publish an empty queue and copy the PE's requests.
 * `ticket` (the workgroup's launch ticket) is read by thread 0 only, when it
 * publishes, so the ticket's atomic completes while the leader works. */
__device__ void lead(const nvshmemx_put_request_t *__restrict__ reqs, unsigned nreq, int P, int p, const pscr &S,
                     unsigned long long ticket) {
    /* issuer (high 7 bits) and granules (low 25 bits) of the first requests,
     * kept from the first pass so the second reads no request again */
    __shared__ unsigned s_req[kLeadCache];
    __shared__ unsigned s_cnt[NVSHMEMI_MAX_PES];
    __shared__ unsigned long long s_units[NVSHMEMI_MAX_PES], s_bytes[NVSHMEMI_MAX_PES];
    __shared__ unsigned s_w[2][kPWaves];
    __shared__ unsigned s_own[kSelfCount]; /* a self-served PE's requests, in request order */
    __shared__ pinfo s_me;
    __shared__ int s_self;
    unsigned *table = S.table();
    pentry *entries = S.entries();
    for (int q = threadIdx.x; q < P; q += kPBlock) s_cnt[q] = 0, s_units[q] = 0, s_bytes[q] = 0;
    __syncthreads();
    /* a PE's entries are its requests that move bytes: zero-byte and invalid
     * requests get none, so no chunk can reach them */
    for (unsigned i = threadIdx.x; i < nreq; i += kPBlock) {
        const nvshmemx_put_request_t r = reqs[i];
        const int q = req_issuer(r, i, P);
        const unsigned u = req_units(r, p == 0);
        if (u) {
            atomicAdd(&s_cnt[q], 1u);
            atomicAdd(&s_units[q], (unsigned long long)u);
            atomicAdd(&s_bytes[q], (unsigned long long)r.nbytes);
        }
        if (i < kLeadCache) s_req[i] = u < (1u << 25) ? ((unsigned)q << 25) | u : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (threadIdx.x < 64) { /* wavefront 0, two PEs per lane (P <= 128); self-served PEs count for nothing */
        const unsigned lane = threadIdx.x, qa = lane, qb = lane + 64;
        const bool ina = qa < (unsigned)P && !leader_served(s_cnt[qa], s_bytes[qa]);
        const bool inb = qb < (unsigned)P && !leader_served(s_cnt[qb], s_bytes[qb]);
        const unsigned long long ua = ina ? s_units[qa] : 0ull, ub = inb ? s_units[qb] : 0ull;
        unsigned long long total = ua + ub;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) total += __shfl_xor(total, d, 64);
        /* sum over PEs of ceil(units / cu) <= total / cu + P <= table size */
        const unsigned long long need = (total + (kTableChunks - P) - 1) / (kTableChunks - P);
        const unsigned cu = (unsigned)(need > kChunkUnitsMin ? need : kChunkUnitsMin);
        /* entries and chunks of the PEs before p */
        unsigned e0 = (qa < (unsigned)p && ina ? s_cnt[qa] : 0u) + (qb < (unsigned)p && inb ? s_cnt[qb] : 0u);
        unsigned c0 = (qa < (unsigned)p ? (unsigned)((ua + cu - 1) / cu) : 0u) +
                      (qb < (unsigned)p ? (unsigned)((ub + cu - 1) / cu) : 0u);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) e0 += __shfl_xor(e0, d, 64), c0 += __shfl_xor(c0, d, 64);
        if (lane == 0) {
            const bool self = leader_served(s_cnt[p], s_bytes[p]);
            s_self = self;
            s_me.entry0 = e0;
            s_me.nown = self ? 0u : s_cnt[p];
            s_me.chunk0 = c0;
            s_me.nchunks = self ? 0u : (unsigned)((s_units[p] + cu - 1) / cu);
            s_me.chunk_units = cu;
        }
    }
    __syncthreads();
    const bool self = s_self;
    const unsigned cu = s_me.chunk_units, e0 = s_me.entry0, c0 = s_me.chunk0;
    unsigned rc = 0, ru = 0; /* own entries and granules of earlier tiles */
    for (unsigned base = 0; base < nreq; base += kPBlock) {
        const unsigned i = base + threadIdx.x;
        unsigned u = 0;
        if (i < nreq) {
            const unsigned c = i < kLeadCache ? s_req[i] : 0xFFFFFFFFu;
            if (c != 0xFFFFFFFFu) {
                if ((int)(c >> 25) == p) u = c & ((1u << 25) - 1);
            } else {
                const nvshmemx_put_request_t r = reqs[i];
                if (req_issuer(r, i, P) == p) u = req_units(r, false);
            }
        }
        const bool own = u != 0;
        unsigned xo, xu, to, tu;
        block_scan2(own ? 1u : 0u, u, xo, xu, to, tu, s_w);
        if (own) {
            if (self) {
                s_own[rc + xo] = i;
            } else {
                const unsigned e = e0 + rc + xo, vs = ru + xu;
                entries[e] = pentry{i, vs};
                /* every chunk that starts inside [vs, vs + u) starts in entry e */
                for (unsigned k = (vs + cu - 1) / cu; k * cu < vs + u; ++k) table[c0 + k] = e;
            }
        }
        rc += to;
        ru += tu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *S.info(p) = s_me;
        store_release(S.ready(p), (unsigned)(ticket / gridDim.x) + 1u); /* writes back this XCD's L2 first */
    }
    if (self) copy_own<false>(reqs, s_cnt[p], s_own);
}

/* Leader of PE p in a flush of posted puts: its queue is slots
 * p * share .. of nbi_queue, min(posted, share) of them.  It lays them out as
 * lead() does, but in regions of its own (entries from p * share, chunk-table
 * slots from p * (kTableChunks / P), its chunk size chosen so its chunks fit
 * there), so it needs no other PE's counts; workers number the chunks of all
 * queues from the published counts.  It zeroes the PE's post counter for the
 * posts after this flush. */
__device__ void lead_posted(int P, int p, const pscr &S, unsigned long long ticket) {
    __shared__ unsigned s_w[2][kPWaves];
    __shared__ unsigned s_own[kSelfCount];
    __shared__ unsigned s_n, s_cnt;
    __shared__ unsigned long long s_units, s_bytes;
    __shared__ pinfo s_me;
    __shared__ int s_self;
    const unsigned share = kNbiSlots / (unsigned)P, tab = kTableChunks / (unsigned)P;
    const unsigned base = (unsigned)p * share;
    const nvshmemx_put_request_t *__restrict__ q = nbi_queue + base;
    if (threadIdx.x == 0) {
        const unsigned posted = load_relaxed(&nbi_count[p]);
        s_n = posted < share ? posted : share;
        s_cnt = 0, s_units = 0, s_bytes = 0;
        store_relaxed(&nbi_count[p], 0u);
    }
    __syncthreads();
    const unsigned n = s_n;
    for (unsigned i = threadIdx.x; i < n; i += kPBlock) {
        const nvshmemx_put_request_t r = q[i];
        const unsigned u = req_units(r, true, r.elem_bytes == kNbiGet);
        if (u) {
            atomicAdd(&s_cnt, 1u);
            atomicAdd(&s_units, (unsigned long long)u);
            atomicAdd(&s_bytes, (unsigned long long)r.nbytes);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const bool self = leader_served(s_cnt, s_bytes);
        const unsigned long long need = (s_units + tab - 1) / tab; /* so that the chunks fit the PE's slots */
        const unsigned cu = (unsigned)(need > kChunkUnitsMin ? need : kChunkUnitsMin);
        s_self = self;
        s_me.entry0 = base;
        s_me.nown = self ? 0u : s_cnt;
        s_me.chunk0 = (unsigned)p * tab;
        s_me.nchunks = self ? 0u : (unsigned)((s_units + cu - 1) / cu);
        s_me.chunk_units = cu;
    }
    __syncthreads();
    const bool self = s_self;
    const unsigned cu = s_me.chunk_units, c0 = s_me.chunk0;
    unsigned *table = S.table();
    pentry *entries = S.entries();
    unsigned rc = 0, ru = 0; /* entries and granules of earlier blocks */
    for (unsigned b = 0; b < n; b += kPBlock) {
        const unsigned i = b + threadIdx.x;
        const unsigned u = i < n ? req_units(q[i], false, q[i].elem_bytes == kNbiGet) : 0u;
        unsigned xo, xu, to, tu;
        block_scan2(u ? 1u : 0u, u, xo, xu, to, tu, s_w);
        if (u) {
            if (self) {
                s_own[rc + xo] = base + i;
            } else {
                const unsigned e = base + rc + xo, vs = ru + xu;
                entries[e] = pentry{base + i, vs};
                /* every chunk that starts inside [vs, vs + u) starts in entry e */
                for (unsigned k = (vs + cu - 1) / cu; k * cu < vs + u; ++k) table[c0 + k] = e;
            }
        }
        rc += to;
        ru += tu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *S.info(p) = s_me;
        store_release(S.ready(p), (unsigned)(ticket / gridDim.x) + 1u); /* writes back this XCD's L2 first */
    }
    if (self) copy_own<true>(nbi_queue, s_cnt, s_own);
}

/* a workgroup copies chunk k of a queue: a request segment of kWaveSegMax
 * bytes or more with all its threads, shorter ones one wavefront each, in
 * turn.  Entries are read 64 at a time, one per lane (the entry, then its
 * request), and then taken in order from the lanes, so a chunk of many small
 * requests waits for two memory round trips per 64 entries, not per entry. */
template <bool POSTED>
__device__ __forceinline__ void run_chunk(const nvshmemx_put_request_t *__restrict__ reqs, const pinfo &inf, unsigned k,
                                          const pscr &S) {
    const uint64_t c_lo = (uint64_t)k * inf.chunk_units * kUnit, c_hi = c_lo + (uint64_t)inf.chunk_units * kUnit;
    const unsigned e_end = inf.entry0 + inf.nown;
    const unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    unsigned small = 0;
    for (unsigned e = coherent_u32(S.table() + inf.chunk0 + k); e < e_end; e += 64) {
        uint64_t vs = ~0ull, tgt = 0, src = 0, nb = 0;
        unsigned pe = 0, get = 0;
        if (e + lane < e_end) {
            const pentry ent = coherent_entry(S.entries() + e + lane);
            vs = (uint64_t)ent.vstart * kUnit;
            if (vs < c_hi) {
                const nvshmemx_put_request_t r = reqs[ent.req];
                tgt = (uint64_t)r.target, src = (uint64_t)r.source, nb = (uint64_t)r.nbytes, pe = (unsigned)r.pe;
                get = POSTED && r.elem_bytes == kNbiGet;
            }
        }
        /* entries start in increasing order, so the lanes in the chunk are a prefix */
        const unsigned cnt = (unsigned)__builtin_popcountll(__ballot(vs < c_hi));
        for (unsigned j = 0; j < cnt; ++j) {
            const uint64_t jvs = lane_u64(vs, j), jnb = lane_u64(nb, j);
            const uint64_t lo = (c_lo > jvs ? c_lo : jvs) - jvs;
            const uint64_t end = jvs + jnb;
            const uint64_t hi = (c_hi < end ? c_hi : end) - jvs;
            if (hi <= lo) continue; /* the chunk starts in the request's padding */
            const int jpe = (int)__builtin_amdgcn_readlane(pe, j);
            const bool jget = POSTED && __builtin_amdgcn_readlane(get, j); /* a posted get: the source is remote */
            char *dst = (jget ? (char *)lane_u64(tgt, j) : translate((const void *)lane_u64(tgt, j), jpe)) + lo;
            const char *srcp = (jget ? translate((const void *)lane_u64(src, j), jpe)
                                     : (const char *)lane_u64(src, j)) + lo;
            const size_t n = (size_t)(hi - lo);
            if (n >= kWaveSegMax) {
                copy<16, kPUnroll, NT_STORE>(dst, srcp, n, threadIdx.x, kPBlock);
            } else {
                if (small % kPWaves == wave) copy<16, kPUnroll, NT_STORE>(dst, srcp, n, lane, 64);
                ++small;
            }
        }
        if (cnt < 64) break;
    }
}

/* every workgroup drains every queue in PE order: chunks are numbered over
 * the queues laid end to end (the chunk table already holds them in PE
 * order); workgroup b takes chunk b first, without an atomic, and claims the
 * rest from this launch's counter, so the workgroups running at once copy
 * neighbouring chunks of the same few PEs' queues.  `ticket` (read by thread
 * 0 only) is the workgroup's launch ticket. */
template <bool POSTED>
__device__ __forceinline__ void drain(const nvshmemx_put_request_t *__restrict__ reqs, const pscr &S, int P,
                                      unsigned long long ticket) {
    __shared__ pinfo s_inf[NVSHMEMI_MAX_PES];
    __shared__ unsigned s_g0[NVSHMEMI_MAX_PES]; /* each queue's first chunk in the launch's numbering */
    __shared__ unsigned s_total, s_k[2], s_tag, s_parity;
    __shared__ int s_qk[2], s_missing;
    if (threadIdx.x == 0) {
        const unsigned long long epoch = ticket / gridDim.x;
        s_tag = (unsigned)epoch + 1u;
        s_parity = (unsigned)epoch & 1u;
        s_total = 0, s_missing = 0;
        /* the next launch's claim counter: no workgroup of this launch uses it */
        if (blockIdx.x == 0) store_relaxed(S.claim(s_parity ^ 1u), 0u);
    }
    __syncthreads();
    const unsigned tag = s_tag;
    unsigned *claim = S.claim(s_parity);
    if ((int)threadIdx.x < P) {
        const int q = (int)threadIdx.x;
        pinfo inf;
        if (wait_tag(S.ready(q), tag)) {
            inf = coherent_info(S.info(q));
        } else { /* a queue that never appeared: its chunks are skipped */
            inf.entry0 = inf.nown = inf.chunk0 = inf.nchunks = inf.chunk_units = 0u;
            s_missing = 1;
        }
        s_inf[q] = inf;
        if (!POSTED && inf.nchunks) atomicMax(&s_total, inf.chunk0 + inf.nchunks);
    }
    __syncthreads();
    if constexpr (POSTED) {
        /* posted queues keep their chunks in regions of their own: number
         * them end to end in PE order (wavefront 0, two PEs a lane) */
        if (threadIdx.x < 64) {
            const unsigned lane = threadIdx.x, qa = 2 * lane, qb = 2 * lane + 1;
            const unsigned na = qa < (unsigned)P ? s_inf[qa].nchunks : 0u, nb = qb < (unsigned)P ? s_inf[qb].nchunks : 0u;
            unsigned incl = na + nb;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned o = __shfl_up(incl, d, 64);
                if (lane >= (unsigned)d) incl += o;
            }
            const unsigned excl = incl - na - nb;
            if (qa < (unsigned)P) s_g0[qa] = excl;
            if (qb < (unsigned)P) s_g0[qb] = excl + na;
            if (lane == 63) s_total = incl;
        }
    } else {
        if (threadIdx.x == 0 && s_missing) /* keep first chunks non-decreasing over a missing queue */
            for (int q = 1; q < P; ++q)
                if (!s_inf[q].chunk_units) s_inf[q].chunk0 = s_inf[q - 1].chunk0 + s_inf[q - 1].nchunks;
        __syncthreads();
        if ((int)threadIdx.x < P) s_g0[threadIdx.x] = s_inf[threadIdx.x].chunk0;
    }
    __syncthreads();
    const unsigned total = s_total;
    /* queue of chunk k: the last PE whose first chunk is at or before k (PEs
     * with no chunks never match: the range test below skips them) */
    auto owner = [&](unsigned k) -> int {
        int lo = 0, hi = P - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_g0[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    if (blockIdx.x >= total) return; /* no chunk for this workgroup: it claims nothing */
    if (threadIdx.x == 0) {
        s_k[0] = blockIdx.x;
        s_qk[0] = owner(blockIdx.x);
    }
    __syncthreads();
    unsigned it = 0;
    for (;;) {
        const unsigned k = __builtin_amdgcn_readfirstlane(s_k[it & 1]);
        const int q = __builtin_amdgcn_readfirstlane(s_qk[it & 1]);
        if (k >= total) break;
        if (threadIdx.x == 0) {
            const unsigned nk = gridDim.x + atomicAdd(claim, 1u);
            s_k[(it + 1) & 1] = nk;
            s_qk[(it + 1) & 1] = nk < total ? owner(nk) : -1;
        }
        const pinfo &inf = s_inf[q];
        if (k - s_g0[q] < inf.nchunks) run_chunk<POSTED>(reqs, inf, k - s_g0[q], S);
        __syncthreads();
        ++it;
    }
}

/* Launch ticket: every workgroup takes one, so launch E (counted from the
 * scratch's set-up, with the grid fixed in between) holds tickets
 * E * grid .. E * grid + grid - 1.  Ready flags carry E + 1 and are never
 * cleared; the claim counters alternate between launches and each launch
 * zeroes the next one's.  Nothing waits for the last workgroup, and a graph
 * replay needs no reset. */
__global__ void __launch_bounds__(kPBlock) k_batch_persistent(const nvshmemx_put_request_t *__restrict__ reqs,
                                                              unsigned nreq, int P, unsigned char *scr) {
    const pscr S{scr};
    unsigned long long ticket = 0;
    if (threadIdx.x == 0)
        ticket = __hip_atomic_fetch_add(S.ticket(), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int p = (int)(blockIdx.x % (unsigned)P);
    if ((int)blockIdx.x < P) lead(reqs, nreq, P, p, S, ticket);
    drain<false>(reqs, S, P, ticket);
}

/* The flush of posted puts: the same launch with each leader taking its
 * PE's own queue (lead_posted) instead of picking its requests out of a
 * shared list. */
__global__ void __launch_bounds__(kPBlock) k_flush_posted(int P, unsigned char *scr) {
    const pscr S{scr};
    unsigned long long ticket = 0;
    if (threadIdx.x == 0)
        ticket = __hip_atomic_fetch_add(S.ticket(), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int)blockIdx.x < P) lead_posted(P, (int)blockIdx.x, S, ticket);
    drain<true>(nbi_queue, S, P, ticket);
}

/* zeroes the control slots (set-up only; a fill kernel, not a memset) */
__global__ void k_persist_clear(unsigned *ctl, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        ctl[i] = 0u;
}

/* --------------------------------------------------------------- host -- */

struct PersistDevice {
    unsigned char *static_ptr = nullptr; /* this device's copy of `scratch` */
    unsigned char *posted_ptr = nullptr; /* ... of `nbi_scratch` */
    unsigned *count_ptr = nullptr;       /* ... of `nbi_count` */
    unsigned char *alloc_ptr = nullptr;  /* library-allocated scratch */
    size_t alloc_reqs = 0;
    int capacity = 0; /* workgroups of the persistent kernel resident at once */
};
static PersistDevice g_pdev[kMaxDevices];

static PersistDevice &pdev() {
    if (g.device < 0 || g.device >= kMaxDevices) NVSHMEMI_FATAL("device %d beyond the persistent-kernel table", g.device);
    return g_pdev[g.device];
}

static void clear_ctl(unsigned char *base) {
    hipLaunchKernelGGL(k_persist_clear, dim3(64), dim3(256), 0, 0, (unsigned *)base, kCtlBytes / sizeof(unsigned));
    NVSHMEMI_HIP(hipGetLastError());
}

/* Library set-up, at every start with the device current: the address of
 * this device's copy of `scratch`, and how many workgroups of the kernel the
 * device runs at once (at most kPPerCU per CU).  The control slots of both
 * scratches are cleared by a fill kernel here: launch tickets count from
 * zero with the grid of this start (it changes only with the number of local
 * PEs, i.e. between starts). */
void persistent_init() {
    PersistDevice &d = pdev();
    if (!d.static_ptr) {
        void *ptr = nullptr, *posted = nullptr, *count = nullptr;
        NVSHMEMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NVSHMEMI_HIP(hipGetSymbolAddress(&posted, HIP_SYMBOL(nbi_scratch)));
        NVSHMEMI_HIP(hipGetSymbolAddress(&count, HIP_SYMBOL(nbi_count)));
        if (!ptr || !posted || !count)
            NVSHMEMI_FATAL("hipGetSymbolAddress returned no address for the persistent kernel's scratch");
        d.static_ptr = (unsigned char *)ptr;
        d.posted_ptr = (unsigned char *)posted;
        d.count_ptr = (unsigned *)count;
    }
    clear_ctl(d.static_ptr);
    clear_ctl(d.posted_ptr);
    hipLaunchKernelGGL(k_persist_clear, dim3(1), dim3(256), 0, 0, d.count_ptr, (size_t)NVSHMEMI_MAX_PES);
    NVSHMEMI_HIP(hipGetLastError());
    if (d.alloc_ptr) clear_ctl(d.alloc_ptr);
    NVSHMEMI_HIP(hipDeviceSynchronize());
    if (!d.capacity) {
        int per_cu = 0, cus = 0;
        NVSHMEMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g.device));
        NVSHMEMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_batch_persistent, kPBlock, 0));
        d.capacity = std::max(1, std::min(per_cu, kPPerCU)) * std::max(1, cus);
    }
}

static int local_issuers() { return g.mode == NVSHMEMI_MODE_LOCAL ? g.npes : 1; }

static void persistent_launch(const nvshmemx_put_request_t *reqs, int nreq, unsigned char *scr, hipStream_t s) {
    const PersistDevice &d = pdev();
    const int P = local_issuers();
    if (d.capacity < P)
        NVSHMEMI_FATAL("the persistent batch kernel needs one resident workgroup per PE: %d PEs, %d resident", P,
                       d.capacity);
    const int grid = P * (d.capacity / P);
    hipLaunchKernelGGL(k_batch_persistent, dim3((unsigned)grid), dim3(kPBlock), 0, s, reqs, (unsigned)nreq, P, scr);
    NVSHMEMI_HIP(hipGetLastError());
}

static void check_args(const nvshmemx_put_request_t *reqs, int nreq, const char *fn) {
    require_init(fn);
    if (nreq < 0 || (nreq > 0 && !reqs)) NVSHMEMI_FATAL("%s: bad arguments (%d requests at %p)", fn, nreq, (const void *)reqs);
}

} // namespace nvshmemi

using namespace nvshmemi;

extern "C" int nvshmemx_batch_put_persistent_on_stream(const nvshmemx_put_request_t *reqs, int nreq, hipStream_t s) {
    check_args(reqs, nreq, "nvshmemx_batch_put_persistent_on_stream");
    if (nreq == 0) return 0;
    PersistDevice &d = pdev();
    if ((size_t)nreq > d.alloc_reqs) {
        NVSHMEMI_HIP(hipDeviceSynchronize()); /* the old scratch may be in use */
        if (d.alloc_ptr) NVSHMEMI_HIP(hipFree(d.alloc_ptr));
        const size_t want = std::max<size_t>((size_t)nreq, 4096);
        NVSHMEMI_HIP(hipMalloc((void **)&d.alloc_ptr, scratch_bytes(want)));
        d.alloc_reqs = want;
        clear_ctl(d.alloc_ptr);
        NVSHMEMI_HIP(hipDeviceSynchronize());
    }
    persistent_launch(reqs, nreq, d.alloc_ptr, s);
    return 0;
}

extern "C" int nvshmemx_batch_put_persistent_noalloc_on_stream(const nvshmemx_put_request_t *reqs, int nreq,
                                                                hipStream_t s) {
    check_args(reqs, nreq, "nvshmemx_batch_put_persistent_noalloc_on_stream");
    if (nreq > kMaxReqsNoalloc) {
        fprintf(stderr,
                "[nvshmem] [%d] nvshmemx_batch_put_persistent_noalloc_on_stream: %d requests exceed the %d its "
                "static scratch holds; nothing was launched\n",
                g.mype, nreq, kMaxReqsNoalloc);
        fflush(stderr);
        return -1;
    }
    if (nreq == 0) return 0;
    const PersistDevice &d = pdev();
    if (!d.static_ptr) NVSHMEMI_FATAL("nvshmemx_batch_put_persistent_noalloc_on_stream: the library's set-up did not run");
    persistent_launch(reqs, nreq, d.static_ptr, s);
    return 0;
}

extern "C" int nvshmemx_batch_persistent_max_requests_noalloc(void) { return kMaxReqsNoalloc; }

extern "C" int nvshmemx_nbi_flush_on_stream(hipStream_t s) {
    require_init("nvshmemx_nbi_flush_on_stream");
    const PersistDevice &d = pdev();
    if (!d.posted_ptr) NVSHMEMI_FATAL("nvshmemx_nbi_flush_on_stream: the library's set-up did not run");
    const int P = local_issuers();
    if (d.capacity < P)
        NVSHMEMI_FATAL("the persistent batch kernel needs one resident workgroup per PE: %d PEs, %d resident", P,
                       d.capacity);
    const int grid = P * (d.capacity / P);
    hipLaunchKernelGGL(k_flush_posted, dim3((unsigned)grid), dim3(kPBlock), 0, s, P, d.posted_ptr);
    NVSHMEMI_HIP(hipGetLastError());
    return 0;
}

extern "C" int nvshmemx_nbi_capacity(void) {
    require_init("nvshmemx_nbi_capacity");
    return (int)(kNbiSlots / (unsigned)local_issuers());
}

extern "C" int nvshmemx_batch_persistent_grid(void) {
    require_init("nvshmemx_batch_persistent_grid");
    const int P = local_issuers();
    return P * std::max(1, pdev().capacity / P);
}
