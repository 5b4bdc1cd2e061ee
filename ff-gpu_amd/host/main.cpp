// ff_sieve <sumStart> <sumLimit> -- the Freudenthal "Sum and Product" search
// on one MI355X, with ff_seg's stdout byte for byte.
//
// Host flow of source/main.cpp:144-573 with the GPU search made the default
// (the reference hard-wired `useGpu = false`, :415): config -> geometry ->
// device budget and run-gate -> header lines -> sieve on the device (the map
// stays in HBM) -> search on the device -> ordered compaction -> emission in
// ff_seg's chunk order -> timing lines.  STDOUT carries only ff_seg's bytes;
// diagnostics go to STDERR with an [ff_sieve] prefix.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "at_exit.h"
#include "budget.h"
#include "config.h"
#include "emission.h"
#include "ffx.h"
#include "geometry.h"
#include "modes.h"
#include "partition.h"
#include "scheduler.h"

using namespace ffg;

namespace {

double msSince(std::chrono::steady_clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

void phase(const char *name, double ms) { std::fprintf(stderr, "ff_sieve timing: %s = %.3f ms\n", name, ms); }

int err(const char *what, int rc)
{
    std::fprintf(stderr, "[ff_sieve] error: %s failed (%d): %s\n", what, rc, ffx_last_error());
    return 1;
}

void printDevice(const ffx_device_info &d, int idx)
{
    std::fprintf(stderr,
                 "[ff_sieve] device[%d] %s (%s vendor=%s bus=%s): free=%llu B (%s) total=%llu B (%s) CUs=%d "
                 "lds/cu=%llu B l2=%llu B clock=%d kHz\n",
                 idx, d.name, d.arch, d.vendor, d.busId, (unsigned long long)d.freeBytes,
                 bytesToHuman(d.freeBytes).c_str(), (unsigned long long)d.totalBytes,
                 bytesToHuman(d.totalBytes).c_str(), d.multiProcessorCount,
                 (unsigned long long)d.sharedMemPerMultiprocessor, (unsigned long long)d.l2Bytes, d.clockKHz);
}

int listDevices()
{
    int n = 0;
    int rc = ffx_device_count(&n);
    if (rc) return err("device enumeration", rc);
    if (n == 0) {
        std::fprintf(stderr, "[ff_sieve] error: no devices enumerated\n");
        return 1;
    }
    for (int i = 0; i < n; ++i) {
        ffx_device_info d;
        if ((rc = ffx_device_info_get(i, &d))) return err("device info", rc);
        printDevice(d, i);
    }
    std::fprintf(stderr, "%d logical device(s)\n", n);
    return 0;
}

// Device map size: the bitmap holds values up to mapM (ff_seg's M unless
// --map-cap lowers it); Miller-Rabin answers above it.
uint64_t deviceMapValue(const SegGeometry &g, uint64_t mapCap)
{
    if (!mapCap || mapCap >= g.maxPrimeMapValue) return g.maxPrimeMapValue;
    return ((mapCap + 1) & ~15ull) - 1 > 15 ? ((mapCap + 1) & ~15ull) - 1 : 15;
}

int runLeg(const Config &cfg, const std::vector<std::string> &pos)
{
    uint32_t sumStart, sumLimit;
    segPositionals(pos, &sumStart, &sumLimit);
    const LegGeometry lg = computeLegGeometry(sumStart, sumLimit);
    const SegGeometry sg = computeSegGeometry(sumStart, sumLimit, cfg.primeMem);
    std::fprintf(stderr,
                 "[ff_sieve] leg: sumStart=%u sumLimit=%u maxGeneratedPrime=%llu productLimit=%llu "
                 "mapBytes=%llu B (%s) maxPrimeMapValue=%llu%s\n",
                 sumStart, sumLimit, (unsigned long long)lg.maxGeneratedPrime, (unsigned long long)lg.productLimit,
                 (unsigned long long)sg.mapBytes, bytesToHuman(sg.mapBytes).c_str(),
                 (unsigned long long)sg.maxPrimeMapValue, sg.capped ? " (PRIMEMEM-capped)" : "");
    if (sg.fatal) {
        std::printf("%s\n", sg.headers[0].c_str());
        std::printf("\nThe input parameters would indicate that more memory is required. Please adjust input "
                    "parameters or Prime::maxPrimeAlloc.\n\n");
        return 255;
    }
    if (sumLimit > FFX_MAX_SUM) {
        std::fprintf(stderr, "[ff_sieve] error: sumLimit %u above the supported %u\n", sumLimit, FFX_MAX_SUM);
        return 1;
    }
    auto tEnum = std::chrono::steady_clock::now();
    int rc = ffx_init(cfg.device);
    if (rc) return err("device init", rc);
    ffx_device_info dev;
    if ((rc = ffx_device_info_get(cfg.device, &dev))) return err("device info", rc);
    phase("device enumeration", msSince(tEnum));
    printDevice(dev, cfg.device);

    // budget and run-gate (source/main.cpp:251-334)
    double f = cfg.globalFraction;
    auto it = cfg.deviceFractions.find("amd");
    if (it == cfg.deviceFractions.end()) it = cfg.deviceFractions.find(std::to_string(cfg.device));
    if (it != cfg.deviceFractions.end()) f = it->second;
    const DeviceBudget b = computeBudget(dev.freeBytes, cfg, f);
    std::fprintf(stderr,
                 "[ff_sieve] device[%d] budget: f=%.3g budget=%llu B (%s) scratch=%llu B headroom=%llu B "
                 "backing=%llu B (%s) slabs=%llu\n",
                 cfg.device, b.fraction, (unsigned long long)b.budget, bytesToHuman(b.budget).c_str(),
                 (unsigned long long)b.scratch, (unsigned long long)b.headroom, (unsigned long long)b.backing,
                 bytesToHuman(b.backing).c_str(), (unsigned long long)b.slabCount);
    if (cfg.slabSizeBytes > b.backing) {
        std::fprintf(stderr, "[ff_sieve] validation error: --slab-size %llu B > backing %llu B\n",
                     (unsigned long long)cfg.slabSizeBytes, (unsigned long long)b.backing);
        return 1;
    }
    const uint64_t mapM = deviceMapValue(sg, cfg.mapCap);
    const TierPlan tier = planTier(b.backing, cfg.hostTierCapBytes, mapM, cfg.slabSizeBytes);
    const uint64_t mapBytes = tier.mapBytes;
    std::fprintf(stderr,
                 "[ff_sieve] aggregate: backing=%llu B hostTier=%llu B capacity=%llu B (%s) mapBytes=%llu B (%s) "
                 "-> %s\n",
                 (unsigned long long)b.backing, (unsigned long long)cfg.hostTierCapBytes,
                 (unsigned long long)tier.capacity, bytesToHuman(tier.capacity).c_str(), (unsigned long long)mapBytes,
                 bytesToHuman(mapBytes).c_str(), tier.gatePass ? "GATE PASS" : "GATE FAIL");
    if (!tier.gatePass) {
        std::fprintf(stderr,
                     "[ff_sieve] ERROR: no device fits leg %u (map %llu B = %s): capacity %llu B (%s)\n"
                     "[ff_sieve] remediation: raise FF_VRAM_BUDGET and/or --host-tier-cap, or bound the device "
                     "map with --map-cap (Miller-Rabin answers above it; stdout is unchanged)\n",
                     sumLimit, (unsigned long long)mapBytes, bytesToHuman(mapBytes).c_str(),
                     (unsigned long long)tier.capacity, bytesToHuman(tier.capacity).c_str());
        return 1;
    }
    const bool dumping = !cfg.dumpMapFile.empty();
    // the search's device map: the whole map, or with a spilled map the
    // largest the backing holds (Miller-Rabin above it, same answers)
    const uint64_t searchM = dumping ? mapM : tier.searchMapValue;
    const uint64_t deviceBytes = dumping && tier.spill ? 0 : (searchM >> 4) + 1;
    if (tier.spill)
        std::fprintf(stderr,
                     "[ff_sieve] host tier: map %s > backing %s; %s\n", bytesToHuman(mapBytes).c_str(),
                     bytesToHuman(b.backing).c_str(),
                     dumping ? "the map streams through device slab buffers into host memory"
                             : "the search's device map holds values up to its bound, Miller-Rabin answers above");

    for (const std::string &h : sg.headers) std::printf("%s\n", h.c_str());
    std::fflush(stdout);

    // ---- sieve
    auto tSieve = std::chrono::steady_clock::now();
    void *stream = ffx_stream_create();
    ffx_sieve_plan *sp = nullptr;
    void *dMap = nullptr;
    uint8_t *pinnedMap = nullptr;
    AtExit release([&] {
        if (stream) ffx_stream_sync(stream);
        ffx_free(dMap);
        ffx_free_pinned(pinnedMap);
        ffx_sieve_plan_destroy(sp);
        ffx_stream_destroy(stream);
    });
    if (!stream) return err("stream", FFX_EIO);
    if ((rc = ffx_sieve_plan_create(&sp, deviceBytes ? searchM : mapM))) return err("sieve plan", rc);
    if (cfg.sieveSegmentBytes && (rc = ffx_sieve_plan_set(sp, FFX_SIEVE_SEGMENT_BYTES, cfg.sieveSegmentBytes)))
        return err("sieve tuning", rc);
    if (cfg.sieveBlocksPerCU && (rc = ffx_sieve_plan_set(sp, FFX_SIEVE_BLOCKS_PER_CU, cfg.sieveBlocksPerCU)))
        return err("sieve tuning", rc);
    std::vector<uint8_t> pageableMap;
    uint8_t *hostMap = nullptr;
    if (dumping) {
        // pinned, so the slab copies run at the link's rate; pageable if the
        // host cannot pin that much
        pinnedMap = (uint8_t *)ffx_alloc_pinned(mapBytes);
        if (!pinnedMap) {
            std::fprintf(stderr, "[ff_sieve] dump-map: %s of pinned host memory refused; using pageable memory\n",
                         bytesToHuman(mapBytes).c_str());
            pageableMap.resize(mapBytes);
        }
        hostMap = pinnedMap ? pinnedMap : pageableMap.data();
    }
    SieveTiming st;
    if (deviceBytes) {
        dMap = ffx_alloc(((deviceBytes + 255) & ~255ull));
        if (!dMap) return err("map allocation", FFX_ENOMEM);
        rc = sieveMap(sp, dMap, deviceBytes, dumping ? cfg.slabSizeBytes : 0, hostMap, stream, &st);
    } else {
        rc = sieveMapToHost(sp, mapBytes, tier.ringSlabBytes, tier.ring, hostMap, stream, &st);
    }
    if (rc) return err("sieve", rc);
    const double sieveWall = msSince(tSieve);
    std::fprintf(stderr, "[ff_sieve] sieve: %llu slab(s)%s, %.3f ms on the device, %.1f GB/s\n",
                 (unsigned long long)st.slabs, deviceBytes ? "" : " through the host tier", st.kernelMs,
                 (double)(deviceBytes ? deviceBytes : mapBytes) / (st.kernelMs * 1e6));
    phase("sieve phase", sieveWall);

    if (dumping) {
        FILE *fo = std::fopen(cfg.dumpMapFile.c_str(), "wb");
        if (!fo || std::fwrite(hostMap, 1, mapBytes, fo) != mapBytes) {
            std::fprintf(stderr, "[ff_sieve] ERROR: cannot write %s\n", cfg.dumpMapFile.c_str());
            if (fo) std::fclose(fo);
            return 1;
        }
        if (std::fclose(fo)) {
            std::fprintf(stderr, "[ff_sieve] ERROR: cannot write %s\n", cfg.dumpMapFile.c_str());
            return 1;
        }
        std::fprintf(stderr, "[ff_sieve] dump-map: wrote %s (%llu bytes)\n", cfg.dumpMapFile.c_str(),
                     (unsigned long long)mapBytes);
        return 0;
    }

    // ---- search
    auto tSearch = std::chrono::steady_clock::now();
    const std::vector<SumChunk> chunks = segRunChunks(sumStart, sumLimit, sg.maxPrimeMapValue);
    uint32_t lo = 0, hi = 0;
    std::vector<ffx_solution> sols;
    double searchKernelMs = 0;
    if (chunkSpan(chunks, &lo, &hi)) {
        if (hi > FFX_MAX_SUM) return err("search range", FFX_ERANGE);
        ffx_search_plan *plan = nullptr;
        ffx_solution *dOut = nullptr;
        uint32_t *dCount = nullptr;
        void *e0 = nullptr, *e1 = nullptr;
        AtExit releaseSearch([&] {
            ffx_stream_sync(stream);
            ffx_event_destroy(e0);
            ffx_event_destroy(e1);
            ffx_free(dOut);
            ffx_free(dCount);
            ffx_search_plan_destroy(plan);
        });
        if ((rc = ffx_search_plan_create(&plan, hi))) return err("search plan", rc);
        if ((cfg.searchGrab && (rc = ffx_search_plan_set(plan, FFX_SEARCH_GRAB, cfg.searchGrab))) ||
            (cfg.searchEager && (rc = ffx_search_plan_set(plan, FFX_SEARCH_EAGER_SPLITS, cfg.searchEager))) ||
            (cfg.searchBlocksPerCU && (rc = ffx_search_plan_set(plan, FFX_SEARCH_BLOCKS_PER_CU, cfg.searchBlocksPerCU))))
            return err("search tuning", rc);
        const uint64_t nslots = (uint64_t)(hi - lo) / 2 + 1;
        dOut = (ffx_solution *)ffx_alloc(nslots * sizeof(ffx_solution));
        dCount = (uint32_t *)ffx_alloc(4);
        e0 = ffx_event_create();
        e1 = ffx_event_create();
        if (!dOut || !dCount || !e0 || !e1) return err("search allocation", FFX_ENOMEM);
        ffx_event_record(e0, stream);
        if ((rc = ffx_search_async(plan, dMap, searchM, lo, hi, dOut, dCount, stream))) return err("search", rc);
        ffx_event_record(e1, stream);
        uint32_t count = 0;
        if ((rc = ffx_copy_d2h_async(&count, dCount, 4, stream)) || (rc = ffx_stream_sync(stream)))
            return err("search sync", rc);
        ffx_event_elapsed_ms(e0, e1, &searchKernelMs);
        sols.resize(count);
        if (count && (rc = ffx_copy_d2h(sols.data(), dOut, (uint64_t)count * sizeof(ffx_solution))))
            return err("solution copy", rc);
        std::fprintf(stderr, "[ff_sieve] search: sums [%u, %u] (%llu odd), %u solutions, kernels %.3f ms, %.0f sums/s\n",
                     lo, hi, (unsigned long long)nslots, count, searchKernelMs,
                     (double)nslots / (searchKernelMs * 1e-3));
    }
    uint64_t ordinal = 0;
    emitChunks(stdout, chunks, sols.data(), sols.size(), &ordinal);
    const double searchWall = msSince(tSearch);
    emitTimes(stdout, (uint64_t)(sieveWall * 1000.0), (uint64_t)(searchWall * 1000.0));
    std::fflush(stdout);
    phase("search phase", searchWall);
    return 0;
}

// key=value options of --batch / --serve (--opt key=value, repeatable)
bool modeOptions(const Config &cfg, BatchOptions *b, ServeOptions *sv, AutotuneOptions *at)
{
    for (const std::string &kv : cfg.modeArgs) {
        const size_t eq = kv.find('=');
        if (eq == std::string::npos) {
            std::fprintf(stderr, "[ff_sieve] error: --opt %s is not key=value\n", kv.c_str());
            return false;
        }
        const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
        const double d = std::atof(v.c_str());
        const unsigned long long u = std::strtoull(v.c_str(), nullptr, 10);
        if (k == "ranges") b->rangesFile = v;
        else if (k == "seed") b->seed = sv->seed = u;
        else if (k == "count") b->count = sv->count = (uint32_t)u;
        else if (k == "start_max") b->startMax = (uint32_t)u;
        else if (k == "wlo") b->wLo = sv->wLo = d;
        else if (k == "whi") b->wHi = sv->wHi = d;
        else if (k == "out_dir") b->outDir = sv->outDir = v;
        else if (k == "reps") b->reps = (uint32_t)u;
        else if (k == "graph") b->useGraph = u != 0;
        else if (k == "lim_lo") sv->limLo = (uint32_t)u;
        else if (k == "lim_hi") sv->limHi = (uint32_t)u;
        else if (k == "rate") sv->rate = d;
        else if (k == "batch_sums") sv->batchSums = u;
        else if (k == "streams") sv->streams = (uint32_t)u;
        else if (k == "plain_sample") sv->plainSample = (uint32_t)u;
        else if (k == "warm") sv->warm = u != 0;
        else if (k == "json") sv->jsonOut = v;
        else if (k == "hashes") b->hashOut = sv->hashOut = v;
        else if (k == "verify") b->verify = sv->verify = u != 0;
        else if (k == "slots_out") b->slotsOut = sv->slotsOut = v;
        else if (k == "sum_limit") at->sumLimit = (uint32_t)u;
        else if (k == "out") at->out = v;
        else if (k == "tune_reps") at->reps = (uint32_t)u;
        else {
            std::fprintf(stderr, "[ff_sieve] error: unknown --opt key '%s'\n", k.c_str());
            return false;
        }
    }
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    int ret = 0;
    {
        Config cfg;
        std::vector<std::string> pos;
        if (loadConfig(argc, argv, &cfg, &pos)) ret = 1;
        else if (cfg.listDevices) ret = listDevices();
        else if (cfg.mode == "batch" || cfg.mode == "serve" || cfg.mode == "autotune") {
            BatchOptions b;
            ServeOptions sv;
            AutotuneOptions at;
            if (!modeOptions(cfg, &b, &sv, &at)) ret = 1;
            else if (cfg.mode == "batch") ret = runBatch(cfg, b);
            else if (cfg.mode == "serve") ret = runServe(cfg, sv);
            else ret = runAutotune(cfg, at);
        } else ret = runLeg(cfg, pos);  // no positionals: ff_seg's default leg 5..2627
    }
    phase("total", msSince(t0));
    return ret;
}
