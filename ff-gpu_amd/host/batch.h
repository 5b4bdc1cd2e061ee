// Batched and served search: many (sumStart, sumLimit) ranges answered by
// one device search over the union of their sums.
//
// Each range's stdout is exactly what `ff_sieve <sumStart> <sumLimit>` (and
// ff_seg) prints for it: its own header lines, its solutions in ff_seg's
// chunk order numbered from 1, the timing lines.  F(s) does not depend on
// the range it is asked for, so overlapping ranges share one slot per sum.
#ifndef FFG_BATCH_H
#define FFG_BATCH_H

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "emission.h"
#include "ffx.h"
#include "geometry.h"
#include "partition.h"

namespace ffg {

struct RangeJob {
    uint32_t sumStart = 5, sumLimit = 5;  // after ff_seg's clamps
    SegGeometry geom;
    std::vector<SumChunk> chunks;
    uint32_t lo = 0, hi = 0;  // odd sums the chunks evaluate
    bool any = false;
};

// ff_seg's view of `<sumStart> <sumLimit>` (segmentedSieve.C:844-855).
RangeJob makeJob(uint32_t sumStart, uint32_t sumLimit, uint64_t primeMem);

// Sorted disjoint odd-sum intervals with consecutive slot offsets.
struct IntervalSet {
    std::vector<ffx_interval> iv;
    uint64_t totalSlots = 0;
    // Slot of odd sum s, or -1 when s is not covered.
    int64_t slotOf(uint32_t s) const;
    uint32_t lastSum() const { return iv.empty() ? 0 : iv.back().last; }
};

// Union of [lo, hi] odd-sum spans (each lo, hi odd, lo <= hi).
IntervalSet unionOf(std::vector<std::pair<uint32_t, uint32_t>> spans);

// Spans minus a covered set (both sorted, disjoint): what is still missing.
std::vector<std::pair<uint32_t, uint32_t>> subtract(const std::vector<std::pair<uint32_t, uint32_t>> &spans,
                                                    const std::vector<std::pair<uint32_t, uint32_t>> &covered);

// One range's stdout without the two timing lines: its header lines and
// its solution lines in chunk order, numbered from 1.  slotAt(s) returns the
// slot word of odd sum s (FFX_* encoding).  Returns the number of solution
// lines.
template <class SlotAt>
uint64_t emitBody(std::string &out, const RangeJob &job, SlotAt slotAt)
{
    for (const std::string &h : job.geom.headers) {
        out += h;
        out += '\n';
    }
    uint64_t n = 0;
    for (const SumChunk &c : job.chunks)
        for (uint64_t s = c.lo | 1u; s <= c.hi; s += 2) {
            const uint32_t v = slotAt((uint32_t)s);
            if (!v) continue;
            ffx_solution r;
            r.sum = (uint32_t)s;
            r.low = v & FFX_SLOT_LOW_MASK;
            r.high = r.sum - r.low;
            r.tags = v & ~FFX_SLOT_LOW_MASK;
            out += formatLine(++n, r);
        }
    return n;
}

// ff_seg's two timing lines (segmentedSieve.C:922).
void appendTimes(std::string &out, uint64_t primeUs, uint64_t searchUs);

// The timing lines with their digits replaced by N, as the reference's
// goldens store every stdout (scripts/verify.sh:37).
extern const char kNormalizedTimes[];

// SHA-256 (hex) of a range's stdout with normalized timing lines, from its
// body (emitBody's text).
std::string normalizedDigest(const std::string &body);

// Body + timing lines: the range's whole stdout.
template <class SlotAt>
uint64_t emitJob(std::string &out, const RangeJob &job, SlotAt slotAt, uint64_t primeUs, uint64_t searchUs)
{
    const uint64_t n = emitBody(out, job, slotAt);
    appendTimes(out, primeUs, searchUs);
    return n;
}

}  // namespace ffg

#endif
