#include "batch.h"

#include <algorithm>

#include "emission.h"
#include "sha256.h"

namespace ffg {

RangeJob makeJob(uint32_t sumStart, uint32_t sumLimit, uint64_t primeMem)
{
    RangeJob j;
    uint32_t a = sumStart | 1u, b = sumLimit;
    if (a < 5) a = 5;
    if (a > b) b = a;
    j.sumStart = a;
    j.sumLimit = b;
    j.geom = computeSegGeometry(a, b, primeMem);
    if (!j.geom.fatal) {
        j.chunks = segRunChunks(a, b, j.geom.maxPrimeMapValue);
        j.any = chunkSpan(j.chunks, &j.lo, &j.hi);
    }
    return j;
}

int64_t IntervalSet::slotOf(uint32_t s) const
{
    size_t lo = 0, hi = iv.size();
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (iv[mid].last < s) lo = mid + 1;
        else hi = mid;
    }
    if (lo == iv.size() || s < iv[lo].first || !(s & 1)) return -1;
    return (int64_t)(iv[lo].slotOffset + (s - iv[lo].first) / 2);
}

IntervalSet unionOf(std::vector<std::pair<uint32_t, uint32_t>> spans)
{
    std::sort(spans.begin(), spans.end());
    IntervalSet r;
    for (const auto &sp : spans) {
        if (!r.iv.empty() && sp.first <= (uint64_t)r.iv.back().last + 2) {
            if (sp.second > r.iv.back().last) r.iv.back().last = sp.second;
        } else {
            ffx_interval x;
            x.first = sp.first;
            x.last = sp.second;
            x.slotOffset = 0;
            r.iv.push_back(x);
        }
    }
    for (ffx_interval &x : r.iv) {
        x.slotOffset = r.totalSlots;
        r.totalSlots += (x.last - x.first) / 2 + 1;
    }
    return r;
}

std::vector<std::pair<uint32_t, uint32_t>> subtract(const std::vector<std::pair<uint32_t, uint32_t>> &spans,
                                                    const std::vector<std::pair<uint32_t, uint32_t>> &covered)
{
    std::vector<std::pair<uint32_t, uint32_t>> out;
    size_t c = 0;
    for (auto sp : spans) {
        uint64_t lo = sp.first;
        const uint64_t hi = sp.second;
        while (c < covered.size() && covered[c].second < lo) ++c;
        size_t k = c;
        while (lo <= hi) {
            if (k >= covered.size() || covered[k].first > hi) {
                out.push_back({(uint32_t)lo, (uint32_t)hi});
                break;
            }
            if (covered[k].first > lo) out.push_back({(uint32_t)lo, covered[k].first - 2});
            lo = (uint64_t)covered[k].second + 2;
            ++k;
        }
    }
    return out;
}

const char kNormalizedTimes[] = "Prime time: N \xce\xbcs\nFreudenthal time: N \xce\xbcs\n";

void appendTimes(std::string &out, uint64_t primeUs, uint64_t searchUs)
{
    char buf[96];
    std::snprintf(buf, sizeof buf, "Prime time: %llu \xce\xbcs\nFreudenthal time: %llu \xce\xbcs\n",
                  (unsigned long long)primeUs, (unsigned long long)searchUs);
    out += buf;
}

std::string normalizedDigest(const std::string &body)
{
    Sha256 h;
    h.update(body);
    h.update(kNormalizedTimes, sizeof kNormalizedTimes - 1);
    return h.hex();
}

}  // namespace ffg
