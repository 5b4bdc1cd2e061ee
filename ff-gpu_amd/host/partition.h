// ff_seg's split of a sum range into thread chunks and the order in which it
// prints them (reference/segmentedSieve.C:706-798, :905-912).
//
// Per-sum output is a pure function of the sum, so the only way the thread
// split shows in stdout is which sums get printed, in which order: for very
// small ranges ff_seg evaluates sumLimit+1 or skips a sum, and overlapping
// chunks print twice.  Reproducing the split keeps ff_sieve's stdout
// identical to ff_seg's for every range, not only for large ones.
#ifndef FFG_PARTITION_H
#define FFG_PARTITION_H

#include <cstdint>
#include <vector>

namespace ffg {

// THREADS in ff_seg (segmentedSieve.C:11-13).
constexpr uint32_t kSegThreads = 31;

struct SumChunk {
    uint32_t lo, hi;  // sums lo|1, lo|1 + 2, ..., <= hi
};

// Chunks of one GenerateThreads(sumStart, sumLimit) call, in print order.
std::vector<SumChunk> segPartition(uint32_t sumStart, uint32_t sumLimit, uint32_t threads = kSegThreads);

// All chunks ff_seg prints for a run, in print order, including the second
// GenerateThreads call above pmapSumLimit (:905-912).
std::vector<SumChunk> segRunChunks(uint32_t sumStart, uint32_t sumLimit, uint64_t maxPrimeMapValue);

// Smallest and largest odd sum any chunk evaluates; false if none.
bool chunkSpan(const std::vector<SumChunk> &chunks, uint32_t *lo, uint32_t *hi);

}  // namespace ffg

#endif
