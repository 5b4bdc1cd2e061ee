#include "partition.h"

#include "geometry.h"

namespace ffg {

static uint32_t dbl_to_u32(double x) { return (uint32_t)(int64_t)x; }

std::vector<SumChunk> segPartition(uint32_t sumStart, uint32_t sumLimit, uint32_t threads)
{
    // Front pointer t*delta and back pointer (2T-t-1)*delta advance toward
    // each other; a pair launches only once both have moved (:723-760).
    const uint32_t numSum = sumLimit - sumStart + 1;
    const double delta = (double)numSum / (2 * threads);
    uint32_t prevFront = 0, prevBack = numSum;
    bool frontMoved = true, backMoved = true;
    std::vector<SumChunk> front, back;
    std::vector<bool> hasBack;
    for (uint32_t t = 0; t < threads; ++t) {
        const uint32_t a = dbl_to_u32(t * delta + .5);
        if (a > prevFront) {
            prevFront = a;
            frontMoved = true;
        }
        const uint32_t t2 = 2 * threads - t - 1;
        const uint32_t b = dbl_to_u32(t2 * delta + .5);
        if (b < prevBack) {
            prevBack = b;
            backMoved = true;
        }
        if (!(frontMoved && backMoved)) continue;
        frontMoved = backMoved = false;
        uint32_t aHi = dbl_to_u32((t + 1) * delta - .5);
        if (aHi < a) aHi = a;
        front.push_back({a + sumStart, aHi + sumStart});
        if (a < b) {
            uint32_t bHi = dbl_to_u32((t2 + 1) * delta - .5);
            if (bHi < b) bHi = b;
            back.push_back({b + sumStart, bHi + sumStart});
            hasBack.push_back(true);
        } else {
            back.push_back({0, 0});
            hasBack.push_back(false);
        }
    }
    // front chunks in launch order, then live back chunks in reverse (:764-791)
    std::vector<SumChunk> out(front);
    for (size_t i = back.size(); i-- > 0;)
        if (hasBack[i]) out.push_back(back[i]);
    return out;
}

std::vector<SumChunk> segRunChunks(uint32_t sumStart, uint32_t sumLimit, uint64_t maxPrimeMapValue)
{
    const uint64_t split = pmapSumLimit(maxPrimeMapValue);
    // sumStart above the split would underflow ff_seg's first call (its
    // numSum wraps); the range is then searched as one call.
    if (sumLimit <= split || sumStart > split) return segPartition(sumStart, sumLimit);
    std::vector<SumChunk> out = segPartition(sumStart, (uint32_t)split);
    std::vector<SumChunk> tail = segPartition((uint32_t)(split + 1) | 1u, sumLimit);
    out.insert(out.end(), tail.begin(), tail.end());
    return out;
}

bool chunkSpan(const std::vector<SumChunk> &chunks, uint32_t *lo, uint32_t *hi)
{
    bool any = false;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (const SumChunk &c : chunks) {
        const uint32_t first = c.lo | 1u;
        if (first > c.hi) continue;
        const uint32_t last = ((c.hi - first) & ~1u) + first;
        if (first < mn) mn = first;
        if (last > mx) mx = last;
        any = true;
    }
    if (any) {
        *lo = mn;
        *hi = mx;
    }
    return any;
}

}  // namespace ffg
