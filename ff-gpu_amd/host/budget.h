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

}  // namespace ffg

#endif
