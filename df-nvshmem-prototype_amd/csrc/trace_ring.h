/*
 * trace_ring.h -- the host side of a trace ring's index math, as pure
 * functions over host memory (the drain in trace.hip uses them; the CPU
 * tests drive them with no device).
 *
 * The device claims records by increasing 32-bit index; record idx lives in
 * slot idx & (cap - 1) (cap a power of two) and is published when its seq
 * field reads idx + 1.  The host consumes records in claim order from its
 * consumed count, which wraps with the claim index.
 */
#ifndef NVSHMEMI_TRACE_RING_H
#define NVSHMEMI_TRACE_RING_H

#include <string.h>

namespace nvshmemi {

/* records published in claim order from `consumed` on: stops at the first
 * slot whose record is not the next one yet, and at a whole ring */
template <typename Rec>
inline unsigned int ring_published(const Rec *ring, unsigned int cap, unsigned int consumed) {
    unsigned int n = 0;
    while (n < cap) {
        const unsigned int idx = consumed + n;
        if (__atomic_load_n(&ring[idx & (cap - 1)].seq, __ATOMIC_ACQUIRE) != idx + 1) break;
        ++n;
    }
    return n;
}

/* copy n records from `consumed` on into out[0..n), in claim order, joining
 * the end of the ring to its start where they wrap */
template <typename Rec>
inline void ring_copy(Rec *out, const Rec *ring, unsigned int cap, unsigned int consumed, unsigned int n) {
    const unsigned int first = consumed & (cap - 1);
    const unsigned int n1 = (first + n <= cap) ? n : cap - first;
    memcpy(out, ring + first, (size_t)n1 * sizeof(Rec));
    if (n1 < n) memcpy(out + n1, ring, (size_t)(n - n1) * sizeof(Rec));
}

} // namespace nvshmemi

#endif /* NVSHMEMI_TRACE_RING_H */
