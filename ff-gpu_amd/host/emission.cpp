#include "emission.h"

#include <algorithm>

namespace ffg {

std::string formatLine(uint64_t ordinal, const ffx_solution &s)
{
    // low-term tag is followed by the padding that ends the low field,
    // high-term tag ends the line (OutputTags :663-677)
    char tl[24], th[16];
    if (s.tags & FFX_TAG_LOW_PRIME) {
        std::snprintf(tl, sizeof tl, " (prime), ");
    } else if (s.tags & FFX_TAG_LOW_POW2) {
        const unsigned k = (unsigned)__builtin_ctz(s.low);
        std::snprintf(tl, sizeof tl, " (2^%u),%s", k, k < 10 ? "   " : "  ");
    } else {
        std::snprintf(tl, sizeof tl, ",         ");
    }
    if (s.tags & FFX_TAG_HIGH_PRIME) std::snprintf(th, sizeof th, " (prime)");
    else if (s.tags & FFX_TAG_HIGH_POW2) std::snprintf(th, sizeof th, " (2^%u)", (unsigned)__builtin_ctz(s.high));
    else th[0] = 0;
    char buf[256];
    const int n = std::snprintf(buf, sizeof buf,
                                "%7llu) sum =%9u, product =%16llu,  low term =%9u%shigh term =%9u%s\n",
                                (unsigned long long)ordinal, s.sum, (unsigned long long)s.low * s.high, s.low, tl,
                                s.high, th);
    return std::string(buf, (size_t)n);
}

uint64_t emitChunks(FILE *out, const std::vector<SumChunk> &chunks, const ffx_solution *sorted, size_t n,
                    uint64_t *ordinal)
{
    uint64_t bytes = 0;
    std::string buf;
    buf.reserve(1 << 20);
    for (const SumChunk &c : chunks) {
        const uint32_t first = c.lo | 1u;
        if (first > c.hi) continue;
        const ffx_solution *it = std::lower_bound(sorted, sorted + n, first,
                                                  [](const ffx_solution &a, uint32_t v) { return a.sum < v; });
        for (; it != sorted + n && it->sum <= c.hi; ++it) {
            buf += formatLine(++*ordinal, *it);
            if (buf.size() > (1u << 20) - 256) {
                bytes += std::fwrite(buf.data(), 1, buf.size(), out);
                buf.clear();
            }
        }
    }
    bytes += std::fwrite(buf.data(), 1, buf.size(), out);
    return bytes;
}

void emitTimes(FILE *out, uint64_t primeUs, uint64_t searchUs)
{
    std::fprintf(out, "Prime time: %llu \xce\xbcs\nFreudenthal time: %llu \xce\xbcs\n", (unsigned long long)primeUs,
                 (unsigned long long)searchUs);
}

}  // namespace ffg
