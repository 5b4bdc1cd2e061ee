// Seeded synthetic workloads: the batched sub-range set and the served
// request stream.  splitmix64 seeds a xoshiro256** generator, so a seed
// names the workload exactly on every host.
#ifndef FFG_WORKLOAD_H
#define FFG_WORKLOAD_H

#include <cstdint>
#include <vector>

namespace ffg {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();  // [0, 1) with 53 bits
};

struct SumRange {
    uint32_t sumStart, sumLimit;
};

struct Request {
    uint32_t sumStart, sumLimit;
    double arrivalS;  // offset from stream start
};

// count sub-ranges: start ~ U[5, startMax), width = floor(2^(wLo + (wHi-wLo)u))
// (log-uniform over [2^wLo, 2^wHi]), sumLimit = start + width - 1.
std::vector<SumRange> batchRanges(uint64_t seed, uint32_t count, uint32_t startMax, double wLo, double wHi);

// count requests: sumLimit ~ U[limLo, limHi], width log-uniform over
// [2^wLo, 2^wHi], sumStart = max(5, sumLimit - width + 1); Poisson arrivals
// at `rate` requests per second.
std::vector<Request> servedRequests(uint64_t seed, uint32_t count, uint32_t limLo, uint32_t limHi, double wLo,
                                    double wHi, double rate);

}  // namespace ffg

#endif
