/*
 * planner_check.cpp -- invariants of the batched-put planner's host half
 * (csrc/batch_plan.h), checked over random request lists for every
 * assignment discipline, set count, grid size, block size and choice source.
 * Built with AddressSanitizer and UndefinedBehaviorSanitizer; no GPU.
 *
 * Checked for every plan:
 *  - offset tables: sizes G+1 and W+1, non-decreasing, contiguous, ending at
 *    the piece count (G = sets x workgroups per set, W = G x block / 64);
 *  - every request's bytes covered exactly once, each piece at the same
 *    offset in source and target, pieces starting on 4 KiB offsets of their
 *    request, access width from the choice;
 *  - workgroup pieces exactly the workgroup-mode requests', each in a list of
 *    its request's set; the grid sized to the work when asked;
 *  - longest-processing-time lists within Graham's bound (max load <= mean
 *    load + largest piece); source-order lists sorted and in order across the
 *    workgroups of a set;
 *  - tiles (choice mode 2) after the lists, sorted by source address, exactly
 *    the tiled requests' bytes: each at most 512 accesses of the widest width
 *    (<= the choice's) the relative alignment allows, source and target
 *    aligned to it, whole accesses only; byte tiles only before the first and
 *    after the last whole access of their request;
 *  - total bytes.
 * Planted mistakes (a dropped piece, a piece off its boundary, a dropped
 * tile, tiles out of source order, a tile narrower than its alignment
 * allows) must each fail.  Exit status 0 when every plan passes.
 */
#include <stdarg.h>
#include <stdio.h>
#include <unistd.h>

#include <map>
#include <random>

#include "batch_plan.h"

namespace nvshmemi {
void fatal(const char *file, int line, const char *fmt, ...) {
    fprintf(stderr, "[%s:%d] ", file, line);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    _exit(3);
}
} // namespace nvshmemi

using namespace nvshmemi;

static int failures = 0;
static bool report = true; /* off while the planted mistakes are checked */
#define CHECK(cond, ...)                                                            \
    do {                                                                            \
        if (!(cond)) {                                                              \
            if (report && failures < 20) {                                          \
                fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);     \
                fprintf(stderr, __VA_ARGS__);                                       \
                fprintf(stderr, "\n");                                              \
            }                                                                       \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

static unsigned vec_of(int v) { return v >= 16 ? 16u : (v >= 8 ? 8u : 4u); }

static void check_plan(const std::vector<resolved_req> &reqs, const std::vector<nvshmemx_batch_choice_t> &ch, int nsets,
                       int wgs_req, int block, int assign, const plan_layout &L, const char *tag) {
    const int waves = block / 64;
    const size_t G = (size_t)nsets * (size_t)L.wgs_per_set, W = G * (size_t)waves;
    CHECK(L.wgs_per_set >= 1, "%s", tag);
    CHECK(L.wg_off.size() == G + 1 && L.wave_off.size() == W + 1, "%s sizes", tag);
    if (L.wg_off.size() != G + 1 || L.wave_off.size() != W + 1) return;
    CHECK(L.wg_off[0] == 0 && L.wg_off[G] == L.wave_off[0] && L.wave_off[W] == L.tile0 &&
              L.tile0 + L.ntiles == L.pieces.size(), "%s ends", tag);
    for (size_t i = 0; i < G; ++i) CHECK(L.wg_off[i] <= L.wg_off[i + 1], "%s wg_off %zu", tag, i);
    for (size_t i = 0; i < W; ++i) CHECK(L.wave_off[i] <= L.wave_off[i + 1], "%s wave_off %zu", tag, i);
    /* each piece back to its request (sources are disjoint) */
    std::map<const char *, size_t> by_src;
    size_t total = 0;
    for (size_t r = 0; r < reqs.size(); ++r) {
        if (reqs[r].nbytes) by_src[reqs[r].src] = r;
        total += reqs[r].nbytes;
    }
    CHECK(L.total_bytes == total, "%s total %zu vs %zu", tag, L.total_bytes, total);
    std::vector<std::vector<std::pair<size_t, size_t>>> cover(reqs.size());
    std::vector<size_t> worker_of(L.pieces.size(), 0);
    for (size_t w = 0; w < G; ++w)
        for (unsigned j = L.wg_off[w]; j < L.wg_off[w + 1]; ++j) worker_of[j] = w;
    for (size_t w = 0; w < W; ++w)
        for (unsigned j = L.wave_off[w]; j < L.wave_off[w + 1]; ++j) worker_of[j] = w / (size_t)waves;
    for (size_t j = 0; j < L.pieces.size(); ++j) {
        const piece_t &p = L.pieces[j];
        auto it = by_src.upper_bound(p.src);
        CHECK(it != by_src.begin(), "%s piece %zu before every request", tag, j);
        if (it == by_src.begin()) continue;
        --it;
        const resolved_req &r = reqs[it->second];
        const size_t off = (size_t)(p.src - r.src);
        CHECK(off + p.nbytes <= r.nbytes, "%s piece %zu past its request", tag, j);
        CHECK((size_t)(p.dst - r.dst) == off, "%s piece %zu target offset", tag, j);
        if (j >= L.tile0) { /* a tile */
            CHECK(ch[it->second].mode == 2, "%s tile %zu of a request that is not tiled", tag, j);
            CHECK(j == L.tile0 || L.pieces[j - 1].src <= p.src, "%s tile %zu out of source order", tag, j);
            const uintptr_t rel = (uintptr_t)r.src ^ (uintptr_t)r.dst;
            unsigned w = vec_of(ch[it->second].vec_bytes);
            while (w > 1 && (rel & (w - 1))) w >>= 1;
            CHECK(p.nbytes > 0, "%s empty tile %zu", tag, j);
            if (p.vec > 1 || w == 1) { /* a body tile (all of them byte tiles when the widest width is 1) */
                CHECK(p.vec == w, "%s tile %zu width %u, widest allowed %u", tag, j, p.vec, w);
                CHECK((uintptr_t)p.src % w == 0 && (uintptr_t)p.dst % w == 0 && p.nbytes % w == 0 &&
                          p.nbytes <= (size_t)kTileAccesses * w, "%s tile %zu shape", tag, j);
            } else {
                CHECK(p.vec == 1 && p.nbytes < w, "%s byte tile %zu of %llu bytes, width %u", tag, j, p.nbytes, w);
                const size_t head = ((size_t)w - ((uintptr_t)r.dst & (w - 1))) & (w - 1);
                /* the head (up to the target's first width boundary) or the tail (after the last whole access) */
                const bool is_head = off == 0 && head > 0 && p.nbytes == std::min(head, r.nbytes);
                const bool is_tail = off + p.nbytes == r.nbytes && (r.nbytes - std::min(head, r.nbytes)) % w == p.nbytes;
                CHECK(is_head || is_tail, "%s byte tile %zu at %zu of %llu bytes is neither head nor tail", tag, j, off,
                      p.nbytes);
            }
            cover[it->second].push_back({off, (size_t)p.nbytes});
            continue;
        }
        CHECK(off % kPieceQuantum == 0, "%s piece %zu offset %zu", tag, j, off);
        CHECK(p.nbytes > 0, "%s empty piece %zu", tag, j);
        CHECK(p.vec == vec_of(ch[it->second].vec_bytes), "%s piece %zu width", tag, j);
        const bool wg_piece = j < L.wg_off[G];
        CHECK(wg_piece == (ch[it->second].mode != 0), "%s piece %zu mode", tag, j);
        CHECK(worker_of[j] / (size_t)L.wgs_per_set == (size_t)r.set, "%s piece %zu set", tag, j);
        cover[it->second].push_back({off, (size_t)p.nbytes});
    }
    for (size_t r = 0; r < reqs.size(); ++r) {
        auto &c = cover[r];
        std::sort(c.begin(), c.end());
        size_t at = 0;
        for (auto &seg : c) {
            CHECK(seg.first == at, "%s request %zu gap or overlap at %zu", tag, r, at);
            at = seg.first + seg.second;
        }
        CHECK(at == reqs[r].nbytes, "%s request %zu covered %zu of %zu", tag, r, at, reqs[r].nbytes);
    }
    if (wgs_req <= 0) {
        size_t most = 1;
        for (int s = 0; s < nsets; ++s) {
            size_t nwg = 0, nwave = 0;
            for (size_t j = 0; j < std::min(L.tile0, L.pieces.size()); ++j)
                if ((int)(worker_of[j] / (size_t)L.wgs_per_set) == s) (j < L.wg_off[G] ? nwg : nwave) += 1;
            most = std::max(most, nwg + (nwave + (size_t)waves - 1) / (size_t)waves);
        }
        const size_t want = std::min<size_t>((size_t)std::max(1, kBatchWorkgroups / nsets), most);
        CHECK((size_t)L.wgs_per_set == want, "%s auto grid %d vs %zu", tag, L.wgs_per_set, want);
    }
    /* discipline-specific properties of the workgroup lists, per set */
    for (int s = 0; s < nsets; ++s) {
        const size_t g0 = (size_t)s * (size_t)L.wgs_per_set, g1 = g0 + (size_t)L.wgs_per_set;
        if (assign == ASSIGN_LPT) {
            double sum = 0, mx = 0, big = 0;
            for (size_t w = g0; w < g1; ++w) {
                double load = 0;
                for (unsigned j = L.wg_off[w]; j < L.wg_off[w + 1]; ++j) {
                    const double c = (double)(L.pieces[j].nbytes / (size_t)waves) + kPieceCharge;
                    load += c;
                    big = std::max(big, c);
                }
                sum += load;
                mx = std::max(mx, load);
            }
            CHECK(mx <= sum / (double)L.wgs_per_set + big + 1e-6, "%s set %d LPT bound %.0f > %.0f + %.0f", tag, s, mx,
                  sum / L.wgs_per_set, big);
        }
        if (assign == ASSIGN_SOURCE_ORDER) {
            const char *last = nullptr;
            for (size_t w = g0; w < g1; ++w)
                for (unsigned j = L.wg_off[w]; j < L.wg_off[w + 1]; ++j) {
                    CHECK(!last || L.pieces[j].src >= last, "%s set %d source order at piece %u", tag, s, j);
                    last = L.pieces[j].src;
                }
        }
    }
}

int main(int argc, char **argv) {
    const int cases = argc > 1 ? atoi(argv[1]) : 60;
    std::mt19937_64 rng(20250101);
    /* addresses only (the planner never dereferences them), made from integers */
    auto addr = [](size_t a) { return (char *)(uintptr_t)a; };
    int plans = 0;
    for (int c = 0; c < cases; ++c) {
        const int nsets = (c % 3 == 0) ? 4 : 1;
        const int n = 1 + (int)(rng() % 300);
        std::vector<resolved_req> reqs((size_t)n);
        size_t src_at = 0, dst_at = 0;
        for (int i = 0; i < n; ++i) {
            const int kind = (int)(rng() % 8);
            size_t nb = kind == 0 ? 0 : (kind < 4 ? (size_t)(rng() % 70000) : (size_t)(rng() % (4u << 20)));
            if (c % 5 == 1) nb &= ~(size_t)3;   /* element-sized requests */
            const size_t sgap = (size_t)(rng() % 4096), dgap = (size_t)(rng() % 4096);
            reqs[(size_t)i] = resolved_req{addr((1ull << 32) + src_at + sgap), addr((1ull << 40) + dst_at + dgap), nb,
                                           (int)(rng() % (unsigned)nsets)};
            src_at += sgap + nb + 1;
            dst_at += dgap + nb + 1;
        }
        for (int source = 0; source < 3; ++source) {           /* rule, random per request, one for all */
            std::vector<nvshmemx_batch_choice_t> ch((size_t)n);
            nvshmemx_batch_choice_t one{};
            for (int i = 0; i < n; ++i) {
                if (source == 0) ch[(size_t)i] = default_choice(reqs[(size_t)i].nbytes);
                else {
                    ch[(size_t)i].pieces = 1 + (int)(rng() % 300);
                    ch[(size_t)i].vec_bytes = 4 << (rng() % 3);
                    ch[(size_t)i].mode = (int)(rng() % 3);
                }
            }
            if (source == 2) {
                one = ch[0];
                for (auto &x : ch) x = one;
            }
            for (int assign : {ASSIGN_LPT, ASSIGN_ROUND_ROBIN, ASSIGN_SCATTERED, ASSIGN_SOURCE_ORDER})
                for (int wgs : {0, 7, 64})
                    for (int block : {64, 256, 1024}) {
                        plan_layout L;
                        build_layout(reqs, source == 1 ? ch.data() : nullptr, source == 2 ? &one : nullptr, nsets, wgs,
                                     block, assign, L);
                        char tag[96];
                        snprintf(tag, sizeof tag, "case %d src %d assign %d wgs %d block %d", c, source, assign, wgs,
                                 block);
                        check_plan(reqs, ch, nsets, wgs, block, assign, L, tag);
                        ++plans;
                    }
        }
    }
    /* the checks are not vacuous: a plan with one piece dropped, and one with
     * a piece moved off its 4 KiB boundary, must both fail */
    const int real_failures = failures;
    {
        std::vector<resolved_req> reqs = {resolved_req{addr(1ull << 32), addr(1ull << 40), (size_t)1 << 20, 0}};
        std::vector<nvshmemx_batch_choice_t> ch = {default_choice((size_t)1 << 20)};
        plan_layout L;
        build_layout(reqs, nullptr, nullptr, 1, 0, 256, ASSIGN_LPT, L);
        report = false;
        plan_layout dropped = L;
        dropped.pieces.pop_back();
        for (auto &o : dropped.wave_off) o = std::min<unsigned>(o, (unsigned)dropped.pieces.size());
        for (auto &o : dropped.wg_off) o = std::min<unsigned>(o, (unsigned)dropped.pieces.size());
        dropped.tile0 = dropped.pieces.size(); /* the plan has no tiles */
        failures = 0;
        check_plan(reqs, ch, 1, 0, 256, ASSIGN_LPT, dropped, "planted: dropped piece");
        const bool caught_drop = failures > 0;
        plan_layout shifted = L;
        shifted.pieces.back().src += 16;
        shifted.pieces.back().dst += 16;
        shifted.pieces.back().nbytes -= 16;
        failures = 0;
        check_plan(reqs, ch, 1, 0, 256, ASSIGN_LPT, shifted, "planted: shifted piece");
        const bool caught_shift = failures > 0;
        /* a tile plan: one tile dropped, two tiles out of source order, a
         * tile narrower than the alignment allows */
        std::vector<resolved_req> treqs = {resolved_req{addr((1ull << 32) + 16), addr((1ull << 40) + 16), (size_t)1 << 20, 0},
                                           resolved_req{addr((1ull << 33) + 3), addr((1ull << 41) + 3), 70000, 0}};
        std::vector<nvshmemx_batch_choice_t> tch(2, nvshmemx_batch_choice_t{1, 16, 2});
        plan_layout T;
        build_layout(treqs, tch.data(), nullptr, 1, 0, 256, ASSIGN_SOURCE_ORDER, T);
        plan_layout tdrop = T;
        tdrop.pieces.pop_back();
        tdrop.ntiles -= 1;
        failures = 0;
        check_plan(treqs, tch, 1, 0, 256, ASSIGN_SOURCE_ORDER, tdrop, "planted: dropped tile");
        const bool caught_tdrop = failures > 0;
        plan_layout tswap = T;
        std::swap(tswap.pieces[tswap.tile0], tswap.pieces[tswap.tile0 + 1]);
        failures = 0;
        check_plan(treqs, tch, 1, 0, 256, ASSIGN_SOURCE_ORDER, tswap, "planted: tiles out of order");
        const bool caught_tswap = failures > 0;
        plan_layout tnarrow = T;
        tnarrow.pieces[tnarrow.tile0].vec = 8;
        failures = 0;
        check_plan(treqs, tch, 1, 0, 256, ASSIGN_SOURCE_ORDER, tnarrow, "planted: narrow tile");
        const bool caught_tnarrow = failures > 0;
        report = true;
        const bool all_caught = caught_drop && caught_shift && caught_tdrop && caught_tswap && caught_tnarrow;
        printf("planted mistakes caught: %s\n", all_caught ? "yes" : "NO");
        failures = real_failures + (all_caught ? 0 : 1);
    }
    printf("%s: %d plans, %d failed checks\n", failures ? "FAILED" : "ok", plans, failures);
    return failures ? 1 : 0;
}
