// Leg geometry and ff_seg's header lines.
//
// Two geometry chains exist in the reference and both are kept:
//  * computeLegGeometry: ff_sieve's chain (source/geometry.cpp:7-22), used for
//    budgets and the map size the device holds;
//  * SegGeometry: ff_seg's chain with the PRIMEMEM cap and the exact header
//    text (reference/segmentedSieve.C:874-879, :136-182), used for stdout.
#ifndef FFG_GEOMETRY_H
#define FFG_GEOMETRY_H

#include <cstdint>
#include <string>
#include <vector>

namespace ffg {

// ff_seg compile-time PRIMEMEM default (segmentedSieve.C:8-9).
constexpr uint64_t kSegPrimeMem = 85ull << 30;

struct LegGeometry {
    uint64_t sumStart = 5;
    uint64_t sumLimit = 2627;
    uint64_t maxGeneratedPrime = 0;  // (sumLimit+1)>>1
    uint64_t productLimit = 0;       // maxGeneratedPrime^2
    uint64_t primeLimit = 0;         // (productLimit>>2)+2
    uint64_t mapBytes = 0;           // ((primeLimit>1?primeLimit:2)+15)>>4
    uint64_t maxPrimeMapValue = 0;   // (mapBytes<<4)-1
};

// sumStart forced odd and >= 5, sumLimit >= sumStart (source/geometry.cpp:7-22).
LegGeometry computeLegGeometry(uint64_t sumStart, uint64_t sumLimit);

// Bytes of the reference CPU search's factor scratch
// (source/geometry.cpp:24-35, FreudenthalTools :408-413).
uint64_t searchWorkspaceBytes(uint64_t sumLimit);

struct SegGeometry {
    uint32_t sumStart = 5, sumLimit = 5;
    uint64_t productLimit = 0, primeLimit = 0;
    uint64_t numPrimesRequested = 0;
    uint64_t maxPrimeMapValue = 0;  // M after the cap
    uint64_t mapBytes = 0;          // bytes of the map ff_seg allocates
    uint64_t totalBytes = 0;
    uint64_t sqrtLimit = 0;
    bool capped = false;            // PRIMEMEM branch taken (3 or 4 header lines)
    bool fatal = false;             // ff_seg would exit(-1) after line 1
    std::vector<std::string> headers;
};

// ff_seg's geometry for the already-clamped (sumStart, sumLimit) pair.
SegGeometry computeSegGeometry(uint32_t sumStart, uint32_t sumLimit, uint64_t primeMem = kSegPrimeMem);

// ff_seg main's positional handling (segmentedSieve.C:844-855): returns the
// clamped (sumStart, sumLimit).
void segPositionals(const std::vector<std::string> &pos, uint32_t *sumStart, uint32_t *sumLimit);

// Floor square root, exact for every uint64_t.
uint64_t isqrt64(uint64_t n);

// The sum at which ff_seg splits its search in two (pmapSumLimit, :905).
uint64_t pmapSumLimit(uint64_t maxPrimeMapValue);

}  // namespace ffg

#endif
