// Per-device memory budget (source/budget.h:22-50, budget.cpp:15-51):
// budget = min(f * free, cap); scratch = fixed or min(0.15 * budget, 1 GiB);
// headroom 64 MiB; backing = budget - scratch - headroom floored to 4 KiB;
// slabs = backing / slab size.  The serving loop admits requests against
// the same backing figure.
#ifndef FFG_BUDGET_H
#define FFG_BUDGET_H

#include <cstdint>
#include <string>

#include "config.h"

namespace ffg {

constexpr uint64_t kHeadroomBytes = 64ull << 20;
constexpr uint64_t kAutoScratchCapBytes = 1ull << 30;

struct DeviceBudget {
    double fraction = 0.90;
    uint64_t freeQueried = 0, freeUsed = 0, cap = 0, budget = 0, scratch = 0, headroom = 0, backing = 0,
             slabCount = 0;
};

DeviceBudget computeBudget(uint64_t freeBytes, const Config &cfg, double fraction);
// One midpoint step toward the 0.50 floor after an allocation failure.
DeviceBudget shrinkFractionForAlloc(uint64_t freeBytes, const Config &cfg, double fraction);
std::string bytesToHuman(uint64_t bytes);

// How a leg's map of mapM values (mapBytes = mapM/16 + 1) fits one device
// with a host tier behind it.  The run-gate is the reference's aggregate one
// (source/main.cpp:296-334): backing + host tier >= mapBytes.  A map larger
// than the backing spills: --dump-map streams it through a ring of device
// slab buffers into host memory (ringSlabBytes * ring <= backing), and the
// search keeps the largest map the backing holds on the device, Miller-Rabin
// answering above it (the same answers).
struct TierPlan {
    uint64_t mapBytes = 0, capacity = 0;
    bool gatePass = false, spill = false;
    uint64_t ringSlabBytes = 0;
    int ring = 0;
    uint64_t searchMapValue = 0;
};
TierPlan planTier(uint64_t backing, uint64_t hostTierBytes, uint64_t mapM, uint64_t slabBytes);

// Device map value for a map that must answer values up to maxM: maxM, or
// the --map-cap bound when smaller (Miller-Rabin answers above the map), as
// a whole number of map bytes (16*bytes - 1), at least 15.
uint64_t mapValueFor(uint64_t maxM, uint64_t cap);

// The map value that keeps the probes of sums up to L in the bitmap as far
// as ff_seg's uncapped geometry does: ((L+1)>>1)^2/4 + 2, rounded up to a
// whole map byte (reference/segmentedSieve.C:874-879).
uint64_t uncappedMapValue(uint32_t L);

// Admission of the serving loop's resident map against the budget: the
// uncapped map for sums up to maxSum (bounded by mapCap), or the largest map
// the backing holds after fixedBytes of other buffers.  0 when the backing
// cannot hold the fixed buffers and 64 MiB of map.
uint64_t serveMapValue(uint64_t backing, uint64_t fixedBytes, uint32_t maxSum, uint64_t mapCap);

}  // namespace ffg

#endif
