// ff_sieve's batched (--batch) and served (--serve) modes.
#ifndef FFG_MODES_H
#define FFG_MODES_H

#include <string>
#include <vector>

#include "config.h"

namespace ffg {

struct BatchOptions {
    std::string rangesFile;      // "sumStart sumLimit" per line, or empty
    uint64_t seed = 20261203;    // seeded ranges when no file is given
    uint32_t count = 4096;
    uint32_t startMax = 1u << 22;
    double wLo = 8.0, wHi = 16.0;
    std::string outDir;          // one file per range when set
    uint32_t reps = 1;           // graph replays timed
    bool useGraph = true;
    std::string hashOut;         // per-range digests of the normalized stdout
    bool verify = false;         // check every range against one plain slot array
    std::string slotsOut;        // verify: the plain slot array (u32 per odd sum from 5)
};

struct ServeOptions {
    uint64_t seed = 20261204;
    uint32_t count = 20000;
    uint32_t limLo = 1u << 18, limHi = 1u << 23;
    double wLo = 6.0, wHi = 18.0;
    double rate = 2000.0;        // requests per second (Poisson)
    uint64_t batchSums = 65536;   // odd sums per persistent launch (the oldest queued first)
    uint32_t streams = 2;        // batches in flight
    uint32_t plainSample = 0;    // also time this many requests, spread evenly over the stream, as
                                 // sequential plain launches
    bool warm = false;           // sieve the whole map before the first arrival (else grown on demand)
    std::string outDir;
    std::string jsonOut;         // summary line
    std::string hashOut;         // per-request digests of the normalized stdout
    bool verify = false;         // check every request against one plain slot array
    std::string slotsOut;        // verify: the plain slot array (u32 per odd sum from 5)
};

struct AutotuneOptions {
    uint32_t sumLimit = 1048575;  // workload: odd sums [5, sumLimit] on ff_seg's map for it
    uint32_t reps = 2;            // timed launches per candidate (the minimum counts)
    std::string out = "ffx_tune.json";
};

int runBatch(const Config &cfg, const BatchOptions &opt);
int runAutotune(const Config &cfg, const AutotuneOptions &opt);
int runServe(const Config &cfg, const ServeOptions &opt);

}  // namespace ffg

#endif
