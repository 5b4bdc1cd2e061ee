#include "budget.h"

#include <algorithm>
#include <cmath>
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

// Sizes in the units of the reference's messages (source/budget.cpp:53-68):
// the largest binary unit the count reaches, two decimals for GiB, one for
// MiB and KiB, whole bytes below.
std::string bytesToHuman(uint64_t bytes)
{
    static const struct {
        int shift, decimals;
        const char *unit;
    } kUnits[] = {{30, 2, "GiB"}, {20, 1, "MiB"}, {10, 1, "KiB"}};
    char text[48];
    for (const auto &u : kUnits)
        if (bytes >> u.shift) {
            std::snprintf(text, sizeof text, "%.*f %s", u.decimals, std::ldexp((double)bytes, -u.shift), u.unit);
            return text;
        }
    return std::to_string(bytes) + " B";
}

TierPlan planTier(uint64_t backing, uint64_t hostTierBytes, uint64_t mapM, uint64_t slabBytes)
{
    TierPlan t;
    t.mapBytes = (mapM >> 4) + 1;
    t.capacity = hostTierBytes > std::numeric_limits<uint64_t>::max() - backing ? std::numeric_limits<uint64_t>::max()
                                                                                : backing + hostTierBytes;
    t.spill = t.mapBytes > backing;
    // three slab buffers when the backing holds them, else two of at most
    // half the backing; whole 4 KiB pages
    t.ring = slabBytes && backing / slabBytes >= 3 ? 3 : 2;
    t.ringSlabBytes = std::min(slabBytes ? slabBytes : backing, backing / (uint64_t)t.ring) & ~4095ull;
    t.gatePass = t.capacity >= t.mapBytes && (!t.spill || t.ringSlabBytes > 0);
    t.searchMapValue = !t.spill ? mapM : backing ? mapValueFor(backing * 16 - 1, 0) : 15;
    return t;
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
