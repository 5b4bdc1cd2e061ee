#include "budget.h"

#include <algorithm>
#include <cstdio>
#include <limits>

namespace ffg {

DeviceBudget computeBudget(uint64_t freeBytes, const Config &cfg, double fraction)
{
    DeviceBudget b;
    b.fraction = fraction;
    b.freeQueried = freeBytes;
    b.freeUsed = (uint64_t)(fraction * (double)freeBytes);
    b.cap = cfg.hasBudgetCap ? cfg.budgetCapBytes : std::numeric_limits<uint64_t>::max();
    b.budget = std::min(b.freeUsed, b.cap);
    b.scratch = cfg.hasScratch ? cfg.scratchBytes
                               : std::min((uint64_t)(0.15 * (double)b.budget), kAutoScratchCapBytes);
    b.headroom = kHeadroomBytes;
    uint64_t backing = 0;
    if (b.budget > b.scratch && b.budget - b.scratch > b.headroom) backing = b.budget - b.scratch - b.headroom;
    backing -= backing % 4096;
    b.backing = backing;
    b.slabCount = cfg.slabSizeBytes ? backing / cfg.slabSizeBytes : 0;
    return b;
}

DeviceBudget shrinkFractionForAlloc(uint64_t freeBytes, const Config &cfg, double fraction)
{
    const double floorF = 0.50;
    if (fraction <= floorF) return computeBudget(freeBytes, cfg, floorF);
    return computeBudget(freeBytes, cfg, 0.5 * (fraction + floorF));
}

std::string bytesToHuman(uint64_t bytes)
{
    char buf[64];
    if (bytes >= (1ull << 30)) std::snprintf(buf, sizeof buf, "%.2f GiB", (double)bytes / (double)(1ull << 30));
    else if (bytes >= (1ull << 20)) std::snprintf(buf, sizeof buf, "%.1f MiB", (double)bytes / (double)(1ull << 20));
    else if (bytes >= (1ull << 10)) std::snprintf(buf, sizeof buf, "%.1f KiB", (double)bytes / 1024.0);
    else std::snprintf(buf, sizeof buf, "%llu B", (unsigned long long)bytes);
    return buf;
}

uint64_t mapValueFor(uint64_t maxM, uint64_t cap)
{
    uint64_t M = maxM;
    if (cap && cap < M) M = cap;
    M = ((M + 1) & ~15ull) - 1;
    return M < 15 ? 15 : M;
}

uint64_t uncappedMapValue(uint32_t L)
{
    const uint64_t mgp = ((uint64_t)L + 1) >> 1;
    const uint64_t need = mgp * mgp / 4 + 2;
    return ((need + 15) & ~15ull) - 1;
}

uint64_t serveMapValue(uint64_t backing, uint64_t fixedBytes, uint32_t maxSum, uint64_t mapCap)
{
    if (backing < fixedBytes + (64ull << 20)) return 0;
    const uint64_t room = backing - fixedBytes;
    uint64_t M = mapValueFor(uncappedMapValue(maxSum), mapCap);
    if ((M >> 4) + 1 > room) M = mapValueFor(((room - 1) << 4) - 1, 0);
    return M;
}

}  // namespace ffg
