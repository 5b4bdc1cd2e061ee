// ff_hostcheck: host-only entry points of the ff_sieve host library for the
// unit tests (no GPU): geometry, headers, partition, budget, config parsing,
// line formatting and workload generation.  One subcommand per call; output
// is "key value" lines or one record per line.
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "batch.h"
#include "budget.h"
#include "config.h"
#include "emission.h"
#include "geometry.h"
#include "partition.h"
#include "sha256.h"
#include "workload.h"

using namespace ffg;

static uint64_t u64(const char *s) { return std::strtoull(s, nullptr, 10); }

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: ff_hostcheck <cmd> ...\n");
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "leg" && argc >= 4) {
        const LegGeometry g = computeLegGeometry(u64(argv[2]), u64(argv[3]));
        std::printf("sumStart %" PRIu64 "\nsumLimit %" PRIu64 "\nmaxGeneratedPrime %" PRIu64 "\nproductLimit %" PRIu64
                    "\nprimeLimit %" PRIu64 "\nmapBytes %" PRIu64 "\nmaxPrimeMapValue %" PRIu64 "\n",
                    g.sumStart, g.sumLimit, g.maxGeneratedPrime, g.productLimit, g.primeLimit, g.mapBytes,
                    g.maxPrimeMapValue);
        return 0;
    }
    if (cmd == "seg" && argc >= 3) {
        std::vector<std::string> pos(argv + 2, argv + argc);
        uint64_t primeMem = kSegPrimeMem;
        if (!pos.empty() && pos.back().rfind("primemem=", 0) == 0) {
            primeMem = u64(pos.back().c_str() + 9);
            pos.pop_back();
        }
        uint32_t a, b;
        segPositionals(pos, &a, &b);
        const SegGeometry g = computeSegGeometry(a, b, primeMem);
        std::printf("sumStart %u\nsumLimit %u\nfatal %d\ncapped %d\nM %" PRIu64 "\nR %" PRIu64 "\nmapBytes %" PRIu64
                    "\nsqrtLimit %" PRIu64 "\npmapSumLimit %" PRIu64 "\n",
                    a, b, (int)g.fatal, (int)g.capped, g.maxPrimeMapValue, g.numPrimesRequested, g.mapBytes,
                    g.sqrtLimit, g.fatal ? 0 : pmapSumLimit(g.maxPrimeMapValue));
        for (const std::string &h : g.headers) std::printf("header %s\n", h.c_str());
        return 0;
    }
    if (cmd == "chunks" && argc >= 4) {
        const uint32_t a = (uint32_t)u64(argv[2]), b = (uint32_t)u64(argv[3]);
        const std::vector<SumChunk> c =
            argc >= 5 ? segRunChunks(a, b, u64(argv[4])) : segPartition(a, b);
        for (const SumChunk &x : c) std::printf("%u %u\n", x.lo, x.hi);
        return 0;
    }
    if (cmd == "budget" && argc >= 4) {
        Config cfg;
        for (int i = 4; i < argc; ++i) {
            if (!std::strncmp(argv[i], "cap=", 4)) {
                cfg.hasBudgetCap = true;
                cfg.budgetCapBytes = u64(argv[i] + 4);
            } else if (!std::strncmp(argv[i], "scratch=", 8)) {
                cfg.hasScratch = true;
                cfg.scratchBytes = u64(argv[i] + 8);
            } else if (!std::strncmp(argv[i], "slab=", 5)) {
                cfg.slabSizeBytes = u64(argv[i] + 5);
            }
        }
        const bool shrink = argc >= 5 && !std::strcmp(argv[argc - 1], "shrink");
        const double f = std::atof(argv[3]);
        const DeviceBudget b = shrink ? shrinkFractionForAlloc(u64(argv[2]), cfg, f) : computeBudget(u64(argv[2]), cfg, f);
        std::printf("fraction %.17g\nfreeUsed %" PRIu64 "\nbudget %" PRIu64 "\nscratch %" PRIu64 "\nheadroom %" PRIu64
                    "\nbacking %" PRIu64 "\nslabCount %" PRIu64 "\n",
                    b.fraction, b.freeUsed, b.budget, b.scratch, b.headroom, b.backing, b.slabCount);
        return 0;
    }
    if (cmd == "tier" && argc >= 6) {  // backing hostTier mapM slab
        const TierPlan t = planTier(u64(argv[2]), u64(argv[3]), u64(argv[4]), u64(argv[5]));
        std::printf("mapBytes %" PRIu64 "\ncapacity %" PRIu64 "\ngatePass %d\nspill %d\nringSlabBytes %" PRIu64
                    "\nring %d\nsearchMapValue %" PRIu64 "\n",
                    t.mapBytes, t.capacity, (int)t.gatePass, (int)t.spill, t.ringSlabBytes, t.ring, t.searchMapValue);
        return 0;
    }
    if (cmd == "human" && argc >= 3) {
        std::printf("%s\n", bytesToHuman(u64(argv[2])).c_str());
        return 0;
    }
    if (cmd == "size" && argc >= 3) {
        uint64_t v = 0;
        if (!readByteSize(argv[2], &v)) return 1;
        std::printf("%" PRIu64 "\n", v);
        return 0;
    }
    if (cmd == "fraction" && argc >= 3) {
        double v = 0;
        if (!readFraction(argv[2], &v)) return 1;
        std::printf("%.17g\n", v);
        return 0;
    }
    if (cmd == "devfrac" && argc >= 3) {
        std::map<std::string, double> m;
        if (!readFractionMap(argv[2], &m)) return 1;
        for (const auto &kv : m) std::printf("%s %.17g\n", kv.first.c_str(), kv.second);
        return 0;
    }
    if (cmd == "config") {
        // parse env + argv[2..] like ff_sieve, then validate
        Config cfg;
        std::vector<std::string> pos;
        std::vector<char *> av;
        av.push_back(argv[0]);
        for (int i = 2; i < argc; ++i) av.push_back(argv[i]);
        if (loadConfig((int)av.size(), av.data(), &cfg, &pos)) return 1;
        std::printf("fraction %.17g\nbudgetCap %" PRIu64 "\nscratch %" PRIu64 "\nslab %" PRIu64 "\nhostTier %" PRIu64
                    "\nprimemem %" PRIu64 "\nmapCap %" PRIu64 "\npositionals %zu\n",
                    cfg.globalFraction, cfg.hasBudgetCap ? cfg.budgetCapBytes : 0,
                    cfg.hasScratch ? cfg.scratchBytes : 0, cfg.slabSizeBytes, cfg.hostTierCapBytes, cfg.primeMem,
                    cfg.mapCap, pos.size());
        return 0;
    }
    if (cmd == "human" && argc >= 3) {
        std::printf("%s\n", bytesToHuman(u64(argv[2])).c_str());
        return 0;
    }
    if (cmd == "workspace" && argc >= 3) {
        std::printf("%" PRIu64 "\n", searchWorkspaceBytes(u64(argv[2])));
        return 0;
    }
    if (cmd == "format" && argc >= 7) {
        ffx_solution s;
        s.sum = (uint32_t)u64(argv[3]);
        s.low = (uint32_t)u64(argv[4]);
        s.high = s.sum - s.low;
        s.tags = (uint32_t)u64(argv[5]);
        std::fputs(formatLine(u64(argv[2]), s).c_str(), stdout);
        (void)argv[6];
        return 0;
    }
    if (cmd == "batch" && argc >= 4) {
        const std::vector<SumRange> r = batchRanges(u64(argv[2]), (uint32_t)u64(argv[3]),
                                                    argc >= 5 ? (uint32_t)u64(argv[4]) : (1u << 22), 8.0, 16.0);
        for (const SumRange &x : r) std::printf("%u %u\n", x.sumStart, x.sumLimit);
        return 0;
    }
    if (cmd == "served" && argc >= 4) {
        const double rate = argc >= 5 ? std::atof(argv[4]) : 1000.0;
        const std::vector<Request> r =
            servedRequests(u64(argv[2]), (uint32_t)u64(argv[3]), argc >= 6 ? (uint32_t)u64(argv[5]) : (1u << 18),
                           argc >= 7 ? (uint32_t)u64(argv[6]) : (1u << 23), argc >= 8 ? std::atof(argv[7]) : 6.0,
                           argc >= 9 ? std::atof(argv[8]) : 18.0, rate);
        for (const Request &x : r) std::printf("%u %u %.9f\n", x.sumStart, x.sumLimit, x.arrivalS);
        return 0;
    }
    if (cmd == "spans") {
        // stdin: "n m", n spans, m covered spans (odd lo <= hi).  Prints the
        // union of the n spans as "first last slotOffset", a "--" line, the
        // union minus the union of the covered spans, a "--" line, and the
        // slot of every odd sum in [1, max + 6] (-1 when not covered).
        size_t n = 0, m = 0;
        if (std::scanf("%zu %zu", &n, &m) != 2) return 2;
        std::vector<std::pair<uint32_t, uint32_t>> a(n), c(m);
        uint32_t top = 0;
        for (auto &x : a) {
            if (std::scanf("%u %u", &x.first, &x.second) != 2) return 2;
            if (x.second > top) top = x.second;
        }
        for (auto &x : c)
            if (std::scanf("%u %u", &x.first, &x.second) != 2) return 2;
        const IntervalSet U = unionOf(a), C = unionOf(c);
        std::vector<std::pair<uint32_t, uint32_t>> u, cv;
        for (const ffx_interval &x : U.iv) {
            std::printf("%u %u %" PRIu64 "\n", x.first, x.last, x.slotOffset);
            u.push_back({x.first, x.last});
        }
        for (const ffx_interval &x : C.iv) cv.push_back({x.first, x.last});
        std::puts("--");
        for (const auto &x : subtract(u, cv)) std::printf("%u %u\n", x.first, x.second);
        std::puts("--");
        for (uint32_t s = 1; s <= top + 6; s += 2) std::printf("%" PRId64 "\n", U.slotOf(s));
        return 0;
    }
    if (cmd == "servemap" && argc >= 6) {
        // serving admission: backing fixedBytes maxSum mapCap -> map value
        std::printf("%" PRIu64 "\n", serveMapValue(u64(argv[2]), u64(argv[3]), (uint32_t)u64(argv[4]), u64(argv[5])));
        return 0;
    }
    if (cmd == "job" && argc >= 4) {
        // makeJob: clamped range, chunk list and evaluated odd-sum span
        const RangeJob j = makeJob((uint32_t)u64(argv[2]), (uint32_t)u64(argv[3]), argc >= 5 ? u64(argv[4]) : kSegPrimeMem);
        std::printf("range %u %u\nfatal %d\nany %d\nspan %u %u\n", j.sumStart, j.sumLimit, (int)j.geom.fatal,
                    (int)j.any, j.lo, j.hi);
        for (const SumChunk &x : j.chunks) std::printf("chunk %u %u\n", x.lo, x.hi);
        return 0;
    }
    if (cmd == "sha256") {
        // SHA-256 of stdin
        Sha256 h;
        char buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, stdin)) > 0) h.update(buf, n);
        std::printf("%s\n", h.hex().c_str());
        return 0;
    }
    if (cmd == "rangedigest" && argc >= 4) {
        // stdin: "sum low high" per odd sum (low 0: not emitted).  Prints the
        // line count and the digest of the range's normalized stdout, built
        // by the emitter the batched and served modes hash with; the tags
        // come from trial division here.
        const RangeJob j = makeJob((uint32_t)u64(argv[2]), (uint32_t)u64(argv[3]), argc >= 5 ? u64(argv[4]) : kSegPrimeMem);
        std::vector<std::pair<uint32_t, uint32_t>> sol;  // (sum, slot)
        auto isPrime = [](uint32_t n) {
            if (n < 2) return false;
            for (uint32_t d = 2; d * d <= n; ++d)
                if (n % d == 0) return false;
            return true;
        };
        unsigned a, lo, hi;
        while (std::scanf("%u %u %u", &a, &lo, &hi) == 3) {
            if (!lo) continue;
            uint32_t v = lo;
            if (isPrime(lo)) v |= FFX_TAG_LOW_PRIME;
            else if (!(lo & (lo - 1))) v |= FFX_TAG_LOW_POW2;
            if (isPrime(hi)) v |= FFX_TAG_HIGH_PRIME;
            else if (!(hi & (hi - 1))) v |= FFX_TAG_HIGH_POW2;
            sol.push_back({a, v});
        }
        std::sort(sol.begin(), sol.end());
        std::string body;
        uint64_t lines = 0;
        if (!j.geom.fatal)
            lines = emitBody(body, j, [&](uint32_t x) {
                auto it = std::lower_bound(sol.begin(), sol.end(), std::make_pair(x, 0u));
                return it != sol.end() && it->first == x ? it->second : 0u;
            });
        std::printf("lines %" PRIu64 "\ndigest %s\n", lines, normalizedDigest(body).c_str());
        return 0;
    }
    std::fprintf(stderr, "ff_hostcheck: bad command\n");
    return 2;
}
