// The trace ring's host-side index math (csrc/trace_ring.h) driven on the
// CPU: records published by claim index into slot idx & (cap - 1) with
// seq = idx + 1.  Checks, for several capacities and starting counts
// (including counts that wrap the 32-bit index): an empty ring, records
// published in order, a gap (a later record published before an earlier
// one), copies that join the ring's end to its start in claim order, a full
// ring, and slots still holding the previous lap's records.  Prints "ok" or
// the first failure.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>

#include "trace_ring.h"

struct Rec {
    uint32_t seq;
    uint32_t payload; /* the claim index, so a copy's order can be checked */
};

static int fail(const char *what, unsigned cap, unsigned start, unsigned got, unsigned want) {
    printf("FAIL %s: cap %u start %u got %u want %u\n", what, cap, start, got, want);
    return 1;
}

static void publish(std::vector<Rec> &ring, unsigned idx) {
    const unsigned cap = (unsigned)ring.size();
    ring[idx & (cap - 1)].payload = idx;
    __atomic_store_n(&ring[idx & (cap - 1)].seq, idx + 1, __ATOMIC_RELEASE);
}

int main() {
    const unsigned caps[] = {1, 2, 8, 64, 1024};
    const unsigned starts[] = {0, 3, 1000, 0xFFFFFFF0u, 0xFFFFFFFFu};
    for (unsigned cap : caps) {
        for (unsigned start : starts) {
            std::vector<Rec> ring(cap);
            /* the slots hold the previous lap's records (or nothing at all) */
            for (unsigned k = 0; k < cap; ++k) {
                const unsigned old = start - cap + k; /* claim index of the slot's previous record */
                if (start >= cap || start == 0xFFFFFFF0u || start == 0xFFFFFFFFu) publish(ring, old);
            }
            if (nvshmemi::ring_published(ring.data(), cap, start) != 0)
                return fail("empty ring", cap, start, nvshmemi::ring_published(ring.data(), cap, start), 0);
            /* publish all but one record in the middle: the count stops at the gap */
            const unsigned gap = cap / 2;
            for (unsigned k = 0; k < cap; ++k)
                if (k != gap) publish(ring, start + k);
            unsigned n = nvshmemi::ring_published(ring.data(), cap, start);
            if (n != gap) return fail("gap", cap, start, n, gap);
            publish(ring, start + gap);
            n = nvshmemi::ring_published(ring.data(), cap, start);
            if (n != cap) return fail("full ring", cap, start, n, cap);
            /* a copy from every offset and length comes out in claim order */
            for (unsigned off = 0; off < cap; off += (cap > 8 ? cap / 8 : 1)) {
                const unsigned len = cap - off;
                std::vector<Rec> out(len);
                nvshmemi::ring_copy(out.data(), ring.data(), cap, start + off, len);
                for (unsigned k = 0; k < len; ++k)
                    if (out[k].payload != start + off + k || out[k].seq != start + off + k + 1)
                        return fail("copy order", cap, start + off, out[k].payload, start + off + k);
            }
            /* the host consumes half; the next lap's records then fill the freed slots */
            const unsigned half = (cap + 1) / 2, consumed = start + half;
            for (unsigned k = 0; k < half; ++k) publish(ring, start + cap + k);
            n = nvshmemi::ring_published(ring.data(), cap, consumed);
            if (n != cap) return fail("next lap", cap, consumed, n, cap);
            std::vector<Rec> out(cap);
            nvshmemi::ring_copy(out.data(), ring.data(), cap, consumed, cap);
            for (unsigned k = 0; k < cap; ++k)
                if (out[k].payload != consumed + k) return fail("next lap order", cap, consumed, out[k].payload, consumed + k);
        }
    }
    printf("ok\n");
    return 0;
}
