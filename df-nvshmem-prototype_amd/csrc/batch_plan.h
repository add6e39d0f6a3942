/*
 * batch_plan.h -- the host half of the batched-put planner: cutting requests
 * into pieces and dealing the pieces to workgroup and wavefront lists, laid
 * out flat with offset tables.  Host code only (no device code, no HIP
 * calls), so it is built into the library by batch.hip and, unchanged, into
 * the host-side checker of its invariants (tests/host/planner_check.cpp).
 */
#ifndef NVSHMEMI_BATCH_PLAN_H
#define NVSHMEMI_BATCH_PLAN_H

#include <algorithm>
#include <queue>
#include <stdlib.h>
#include <vector>

#include "nvshmemi_host.h"

namespace nvshmemi {

struct piece_t {
    const char *src;
    char *dst;
    unsigned long long nbytes;
    unsigned int vec;
    unsigned int pad;
};

/* per-piece charge in the LPT cost, in bytes: about one access latency of a
 * wavefront streaming 4 x 1 KiB (0 to 64 KiB measured the same) */
constexpr double kPieceCharge = 4096.0;

/* The library's untuned rule for one request: about 256 KiB per workgroup
 * piece from 64 KiB up, wavefront pieces of about 16 KiB below, 16-byte
 * accesses where the alignment allows. */
inline nvshmemx_batch_choice_t default_choice(size_t nbytes) {
    nvshmemx_batch_choice_t c;
    c.mode = nbytes >= (64u << 10) ? 1 : 0;
    const size_t per = c.mode ? (256u << 10) : (16u << 10);
    const size_t parts = (nbytes + per - 1) / per;
    c.pieces = (int)std::max<size_t>(1, std::min<size_t>(parts, 256));
    c.vec_bytes = 16;
    return c;
}

/* one request's pieces under a choice; boundaries on kPieceQuantum offsets */
inline void split_request(const char *src, char *dst, size_t nbytes, const nvshmemx_batch_choice_t &ch,
                          std::vector<piece_t> &out) {
    if (nbytes == 0) return;
    const size_t quantum = kPieceQuantum;
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

/* Tiles (choice mode 2): a request cut into tiles of kTileAccesses accesses
 * of the widest width <= maxw that the relative alignment of its source and
 * target allows, each tile copied by one workgroup of kTileBlock threads (two
 * accesses a lane, one after the other).  The bytes before the target's first
 * boundary of that width, and those after the last whole access, go as byte
 * tiles (width 1). */
constexpr unsigned kTileAccesses = 512;
constexpr int kTileBlock = 256;
inline void tile_request(const char *src, char *dst, size_t nbytes, int maxw, std::vector<piece_t> &out) {
    if (nbytes == 0) return;
    const uintptr_t rel = (uintptr_t)src ^ (uintptr_t)dst;
    size_t w = maxw >= 16 ? 16 : (maxw >= 8 ? 8 : 4);
    while (w > 1 && (rel & (w - 1))) w >>= 1;
    const size_t head = std::min(nbytes, (size_t)((w - ((uintptr_t)dst & (w - 1))) & (w - 1)));
    if (head) out.push_back(piece_t{src, dst, (unsigned long long)head, 1, 0});
    const size_t body = (nbytes - head) / w * w, step = (size_t)kTileAccesses * w;
    for (size_t off = 0; off < body; off += step)
        out.push_back(piece_t{src + head + off, dst + head + off, (unsigned long long)std::min(step, body - off),
                              (unsigned)w, 0});
    const size_t tail = nbytes - head - body;
    if (tail) out.push_back(piece_t{src + head + body, dst + head + body, (unsigned long long)tail, 1, 0});
}

/* LPT: pieces (any order) onto `workers` bins; returns per-worker lists.
 * `base_load[w]` is load a worker already carries; `share` divides a piece's
 * bytes among the lanes of a workgroup worker's wavefronts. */
inline std::vector<std::vector<piece_t>> lpt(std::vector<piece_t> pieces, int workers, size_t share,
                                            std::vector<double> &load) {
    std::vector<std::vector<piece_t>> lists((size_t)workers);
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const piece_t &a, const piece_t &b) { return a.nbytes > b.nbytes; });
    typedef std::pair<double, int> item;
    std::priority_queue<item, std::vector<item>, std::greater<item>> heap;
    for (int w = 0; w < workers; ++w) heap.push(item(load[(size_t)w], w));
    const double charge = kPieceCharge;
    for (const piece_t &p : pieces) {
        item top = heap.top();
        heap.pop();
        lists[(size_t)top.second].push_back(p);
        top.first += (double)(p.nbytes / share) + charge;
        load[(size_t)top.second] = top.first;
        heap.push(top);
    }
    return lists;
}

/* Source order: pieces sorted by source address and cut into `workers`
 * consecutive runs of (nearly) equal cost, run w to worker w.  The hardware
 * dispatches workgroups in index order, so the workgroups running at any
 * moment read neighbouring sources: puts whose sources overlap (or sit in the
 * same heap) then share cache lines instead of each streaming its own.  The
 * last run ends within one piece of the mean (runs close on the first piece
 * that reaches their share). */
inline std::vector<std::vector<piece_t>> source_runs(std::vector<piece_t> pieces, int workers, size_t share,
                                                    std::vector<double> &load) {
    std::vector<std::vector<piece_t>> lists((size_t)workers);
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const piece_t &a, const piece_t &b) { return a.src < b.src; });
    double total = 0;
    for (const piece_t &p : pieces) total += (double)(p.nbytes / share) + kPieceCharge;
    const double target = total / (double)workers;
    double acc = 0;
    size_t w = 0;
    for (const piece_t &p : pieces) {
        const double c = (double)(p.nbytes / share) + kPieceCharge;
        lists[w].push_back(p);
        load[w] += c;
        acc += c;
        if (acc >= target * (double)(w + 1) && w + 1 < (size_t)workers) ++w;
    }
    return lists;
}

/* Fisher-Yates with a fixed xorshift64 stream: a reproducible order */
inline void scatter(std::vector<piece_t> &v) {
    unsigned long long x = 0x9E3779B97F4A7C15ull;
    for (size_t i = v.size(); i > 1; --i) {
        x ^= x << 13, x ^= x >> 7, x ^= x << 17;
        std::swap(v[i - 1], v[(size_t)(x % i)]);
    }
}

/* A plan's host arrays: the pieces of every worker list back to back, the
 * offset of each workgroup's list (G + 1 entries) and of each wavefront's
 * list (G * block / 64 + 1 entries), G = nsets * wgs_per_set; then the tiles
 * of the requests cut into tiles, sorted by source address, from `tile0` on
 * (one workgroup each, whatever set their request is in). */
struct plan_layout {
    std::vector<piece_t> pieces;
    std::vector<unsigned int> wg_off, wave_off;
    int wgs_per_set = 0;
    size_t total_bytes = 0;
    size_t tile0 = 0, ntiles = 0;
};

inline void build_layout(const std::vector<resolved_req> &reqs, const nvshmemx_batch_choice_t *choices,
                         const nvshmemx_batch_choice_t *one_choice, int nsets, int wgs_per_set, int block, int assign,
                         plan_layout &L) {
    if (block < 64 || block > 1024 || block % 64) NVSHMEMI_FATAL("batch block size %d invalid", block);
    const int waves_per_wg = block / 64;
    std::vector<std::vector<piece_t>> wg(nsets), wave(nsets);
    std::vector<piece_t> tiles;
    size_t total = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const resolved_req &r = reqs[i];
        nvshmemx_batch_choice_t ch;
        if (one_choice) ch = *one_choice;
        else if (choices) ch = choices[i];
        else ch = default_choice(r.nbytes);
        if (ch.mode == 2) tile_request(r.src, r.dst, r.nbytes, ch.vec_bytes, tiles);
        else split_request(r.src, r.dst, r.nbytes, ch, ch.mode ? wg[r.set] : wave[r.set]);
        total += r.nbytes;
    }
    if (wgs_per_set <= 0) {
        /* grid sized to the work: one workgroup per workgroup piece or per
         * workgroup's worth of wavefront pieces, at most the default grid */
        size_t most = 1;
        for (int s = 0; s < nsets; ++s)
            most = std::max(most, wg[s].size() + (wave[s].size() + waves_per_wg - 1) / waves_per_wg);
        wgs_per_set = (int)std::min<size_t>(std::max(1, kBatchWorkgroups / nsets), most);
    }
    const size_t G = (size_t)nsets * wgs_per_set, W = G * waves_per_wg;
    const size_t wps = (size_t)wgs_per_set * waves_per_wg; /* wavefront workers per set */
    std::vector<piece_t> &all = L.pieces;
    all.clear();
    std::vector<unsigned int> &wo = L.wg_off, &vo = L.wave_off;
    wo.assign(G + 1, 0);
    vo.assign(W + 1, 0);
    if (assign == ASSIGN_LPT || assign == ASSIGN_SOURCE_ORDER) {
        std::vector<std::vector<piece_t>> wg_lists(G), wave_lists(W);
        for (int s = 0; s < nsets; ++s) {
            const size_t g0 = (size_t)s * wgs_per_set, w0 = g0 * waves_per_wg;
            std::vector<double> wg_load((size_t)wgs_per_set, 0.0);
            std::vector<std::vector<piece_t>> a = assign == ASSIGN_LPT
                                                      ? lpt(wg[s], wgs_per_set, (size_t)waves_per_wg, wg_load)
                                                      : source_runs(wg[s], wgs_per_set, (size_t)waves_per_wg, wg_load);
            /* a wavefront starts its own pieces after its workgroup's */
            std::vector<double> wave_load(wps);
            for (size_t w = 0; w < wps; ++w) wave_load[w] = wg_load[w / waves_per_wg];
            std::vector<std::vector<piece_t>> b = lpt(wave[s], (int)wps, 1, wave_load);
            for (size_t i = 0; i < a.size(); ++i) wg_lists[g0 + i] = std::move(a[i]);
            for (size_t i = 0; i < b.size(); ++i) wave_lists[w0 + i] = std::move(b[i]);
        }
        for (size_t i = 0; i < G; ++i) {
            wo[i] = (unsigned)all.size();
            all.insert(all.end(), wg_lists[i].begin(), wg_lists[i].end());
        }
        wo[G] = (unsigned)all.size();
        for (size_t i = 0; i < W; ++i) {
            vo[i] = (unsigned)all.size();
            all.insert(all.end(), wave_lists[i].begin(), wave_lists[i].end());
        }
        vo[W] = (unsigned)all.size();
    } else {
        /* worker i of a set takes pieces i, i + workers, i + 2 * workers, ...
         * of the set's list, written straight into the flat layout */
        size_t npieces = 0;
        for (int s = 0; s < nsets; ++s) {
            if (assign == ASSIGN_SCATTERED) scatter(wg[s]), scatter(wave[s]);
            npieces += wg[s].size() + wave[s].size();
        }
        all.reserve(npieces);
        for (int s = 0; s < nsets; ++s)
            for (size_t i = 0; i < (size_t)wgs_per_set; ++i) {
                wo[(size_t)s * wgs_per_set + i] = (unsigned)all.size();
                for (size_t j = i; j < wg[s].size(); j += (size_t)wgs_per_set) all.push_back(wg[s][j]);
            }
        wo[G] = (unsigned)all.size();
        for (int s = 0; s < nsets; ++s)
            for (size_t i = 0; i < wps; ++i) {
                vo[(size_t)s * wps + i] = (unsigned)all.size();
                for (size_t j = i; j < wave[s].size(); j += wps) all.push_back(wave[s][j]);
            }
        vo[W] = (unsigned)all.size();
    }
    /* tiles in source order: the workgroups running at any moment read
     * neighbouring sources (see source_runs) */
    std::stable_sort(tiles.begin(), tiles.end(), [](const piece_t &a, const piece_t &b) { return a.src < b.src; });
    L.tile0 = all.size();
    L.ntiles = tiles.size();
    all.insert(all.end(), tiles.begin(), tiles.end());
    L.wgs_per_set = wgs_per_set;
    L.total_bytes = total;
}

} // namespace nvshmemi

#endif /* NVSHMEMI_BATCH_PLAN_H */
