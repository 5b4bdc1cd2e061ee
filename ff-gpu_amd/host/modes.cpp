// Batched and served modes of ff_sieve.
//
// --batch: N sub-ranges, each answered exactly as `ff_sieve <a> <b>` would
// answer it, from ONE sieve and ONE persistent search launch over the union
// of their sums (captured in a hipGraph).
//
// --serve: a seeded Poisson stream of (sumStart, sumLimit) requests.  The
// bitmap stays resident and grows on demand; requests that have arrived are
// admitted while the device-memory budget allows, batched into persistent
// launches over the sums no earlier batch has computed, and up to `streams`
// batches run on the device at once.  Per-request latency is arrival to
// output; the same requests can be replayed one at a time as plain launches
// for comparison.
#include "modes.h"

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <sstream>
#include <climits>
#include <condition_variable>
#include <mutex>
#include <thread>

#include "at_exit.h"
#include "batch.h"
#include "budget.h"
#include "ffx.h"
#include "sha256.h"
#include "workload.h"

namespace ffg {

namespace {

using Clock = std::chrono::steady_clock;

double msSince(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }



int fail(const char *what, int rc)
{
    std::fprintf(stderr, "[ff_sieve] error: %s failed (%d): %s\n", what, rc, ffx_last_error());
    return 1;
}

int applyTuning(const Config &cfg, ffx_sieve_plan *sp, ffx_search_plan *qp)
{
    int rc = 0;
    if (sp && cfg.sieveSegmentBytes && (rc = ffx_sieve_plan_set(sp, FFX_SIEVE_SEGMENT_BYTES, cfg.sieveSegmentBytes)))
        return rc;
    if (sp && cfg.sieveBlocksPerCU && (rc = ffx_sieve_plan_set(sp, FFX_SIEVE_BLOCKS_PER_CU, cfg.sieveBlocksPerCU)))
        return rc;
    if (qp && cfg.searchGrab && (rc = ffx_search_plan_set(qp, FFX_SEARCH_GRAB, cfg.searchGrab))) return rc;
    if (qp && cfg.searchEager && (rc = ffx_search_plan_set(qp, FFX_SEARCH_EAGER_SPLITS, cfg.searchEager))) return rc;
    if (qp && cfg.searchBlocksPerCU && (rc = ffx_search_plan_set(qp, FFX_SEARCH_BLOCKS_PER_CU, cfg.searchBlocksPerCU)))
        return rc;
    return 0;
}

bool writeFile(const std::string &path, const std::string &data)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    return std::fclose(f) == 0 && ok;
}

std::string fatalText(const RangeJob &j)
{
    return j.geom.headers[0] +
           "\n\nThe input parameters would indicate that more memory is required. Please adjust input parameters or "
           "Prime::maxPrimeAlloc.\n\n";
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double idx = p * (double)(v.size() - 1);
    const size_t i = (size_t)idx;
    const double f = idx - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

// A range's output, hashed: its solution-line count and the SHA-256 of
// its stdout with normalized timing lines (the fatal text as printed).
struct Digest {
    uint64_t lines = 0;
    std::string hex;
};

template <class SlotAt>
Digest digestOf(const RangeJob &j, SlotAt slotAt, std::string &text)
{
    text.clear();
    Digest d;
    if (j.geom.fatal) {
        text = fatalText(j);
        d.hex = sha256Hex(text);
        return d;
    }
    d.lines = emitBody(text, j, slotAt);
    d.hex = normalizedDigest(text);
    return d;
}

// "index sumStart sumLimit lines sha256" per range.
bool writeDigests(const std::string &path, const std::vector<RangeJob> &jobs, const std::vector<Digest> &d)
{
    std::string out;
    char buf[160];
    for (size_t i = 0; i < jobs.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%zu %u %u %llu %s\n", i, jobs[i].sumStart, jobs[i].sumLimit,
                      (unsigned long long)d[i].lines, d[i].hex.c_str());
        out += buf;
    }
    return writeFile(path, out);
}

struct PlainCheck {
    uint64_t checked = 0, mismatched = 0;
    long firstBad = -1;
    double searchMs = 0;
};

// Every range's digest recomputed from ONE plain slot array over the odd
// sums [5, maxSum] (ffx_search_slots_async on the whole map: the plain path,
// not the batched or served one), formatted by the same emitter; a range
// whose digest or line count differs is a mismatch.  slotsOut, when set,
// receives the plain slot array (u32 per odd sum from 5, little-endian).
int plainCheck(ffx_search_plan *qp, const void *dMap, uint64_t M, uint32_t maxSum, const std::vector<RangeJob> &jobs,
               const std::vector<Digest> &got, const std::string &slotsOut, void *stream, PlainCheck *pc)
{
    if (maxSum < 5) maxSum = 5;
    const uint64_t nslots = (maxSum - 5) / 2 + 1;
    uint32_t *d = (uint32_t *)ffx_alloc(nslots * 4);
    void *e0 = ffx_event_create(), *e1 = ffx_event_create();
    AtExit release([&] {
        ffx_stream_sync(stream);
        ffx_free(d);
        ffx_event_destroy(e0);
        ffx_event_destroy(e1);
    });
    if (!d || !e0 || !e1) return fail("plain check buffers", FFX_ENOMEM);
    int rc;
    if ((rc = ffx_event_record(e0, stream)) || (rc = ffx_search_slots_async(qp, dMap, M, 5, maxSum, d, stream)) ||
        (rc = ffx_event_record(e1, stream)) || (rc = ffx_event_sync(e1)))
        return fail("plain check search", rc);
    ffx_event_elapsed_ms(e0, e1, &pc->searchMs);
    std::vector<uint32_t> plain(nslots);
    if ((rc = ffx_copy_d2h(plain.data(), d, nslots * 4))) return fail("plain check copy", rc);
    if (!slotsOut.empty() &&
        !writeFile(slotsOut, std::string(reinterpret_cast<const char *>(plain.data()), nslots * 4))) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write %s\n", slotsOut.c_str());
        return 1;
    }
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> bad{0};
    std::atomic<long> firstBad{LONG_MAX};
    auto worker = [&] {
        std::string text;
        for (size_t i; (i = next.fetch_add(1)) < jobs.size();) {
            const Digest want = digestOf(jobs[i], [&](uint32_t s) { return plain[(s - 5) >> 1]; }, text);
            if (want.hex != got[i].hex || want.lines != got[i].lines) {
                ++bad;
                long cur = firstBad.load();
                while ((long)i < cur && !firstBad.compare_exchange_weak(cur, (long)i)) {
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < 8; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
    pc->checked = jobs.size();
    pc->mismatched = bad;
    pc->firstBad = bad ? firstBad.load() : -1;
    if (bad)
        std::fprintf(stderr, "[ff_sieve] PARITY MISMATCH: %llu of %zu ranges differ from the plain slots (first: %ld, "
                             "%u..%u)\n",
                     (unsigned long long)pc->mismatched, jobs.size(), pc->firstBad, jobs[pc->firstBad].sumStart,
                     jobs[pc->firstBad].sumLimit);
    return 0;
}

std::string parityJson(bool verify, const PlainCheck &pc)
{
    if (!verify) return "\"parity\": null";
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "\"parity\": \"%s\", \"parity_checked\": %llu, \"parity_mismatches\": %llu, "
                  "\"plain_check_search_ms\": %.3f",
                  pc.mismatched ? "MISMATCH" : "per-range sha256 == plain slots", (unsigned long long)pc.checked,
                  (unsigned long long)pc.mismatched, pc.searchMs);
    return buf;
}

struct DeviceSetup {
    ffx_device_info info{};
    DeviceBudget budget;
};

int setupDevice(const Config &cfg, DeviceSetup *d)
{
    int rc = ffx_init(cfg.device);
    if (rc) return fail("device init", rc);
    if ((rc = ffx_device_info_get(cfg.device, &d->info))) return fail("device info", rc);
    d->budget = computeBudget(d->info.freeBytes, cfg, cfg.globalFraction);
    return 0;
}

}  // namespace

// --------------------------------------------------------------------------
// batched
// --------------------------------------------------------------------------

int runBatch(const Config &cfg, const BatchOptions &opt)
{
    std::vector<SumRange> ranges;
    if (!opt.rangesFile.empty()) {
        std::ifstream in(opt.rangesFile);
        if (!in) {
            std::fprintf(stderr, "[ff_sieve] error: cannot read %s\n", opt.rangesFile.c_str());
            return 1;
        }
        unsigned long long a, b;
        while (in >> a >> b) ranges.push_back({(uint32_t)a, (uint32_t)b});
    } else {
        ranges = batchRanges(opt.seed, opt.count, opt.startMax, opt.wLo, opt.wHi);
    }
    std::vector<RangeJob> jobs;
    jobs.reserve(ranges.size());
    std::vector<std::pair<uint32_t, uint32_t>> spans;
    uint64_t maxM = 15, requested = 0;
    for (const SumRange &r : ranges) {
        jobs.push_back(makeJob(r.sumStart, r.sumLimit, cfg.primeMem));
        const RangeJob &j = jobs.back();
        if (j.geom.fatal) continue;
        if (j.geom.maxPrimeMapValue > maxM) maxM = j.geom.maxPrimeMapValue;
        for (const SumChunk &c : j.chunks)
            if ((c.lo | 1u) <= c.hi) requested += (c.hi - (c.lo | 1u)) / 2 + 1;
        if (j.any) spans.push_back({j.lo, j.hi});
    }
    const IntervalSet U = unionOf(spans);
    if (U.lastSum() > FFX_MAX_SUM) {
        std::fprintf(stderr, "[ff_sieve] error: sums above %u\n", FFX_MAX_SUM);
        return 1;
    }
    DeviceSetup dev;
    if (setupDevice(cfg, &dev)) return 1;
    const uint64_t M = mapValueFor(maxM, cfg.mapCap);
    const uint64_t mapBytes = (M >> 4) + 1;
    const uint64_t need = mapBytes + U.totalSlots * 4 + U.iv.size() * sizeof(ffx_interval);
    std::fprintf(stderr,
                 "[ff_sieve] batch: %zu ranges, %llu requested odd sums, %llu distinct in %zu intervals; map M=%llu "
                 "(%s), backing %s\n",
                 jobs.size(), (unsigned long long)requested, (unsigned long long)U.totalSlots, U.iv.size(),
                 (unsigned long long)M, bytesToHuman(mapBytes).c_str(), bytesToHuman(dev.budget.backing).c_str());
    if (need > dev.budget.backing) {
        std::fprintf(stderr, "[ff_sieve] ERROR: batch needs %s > backing %s; bound the map with --map-cap\n",
                     bytesToHuman(need).c_str(), bytesToHuman(dev.budget.backing).c_str());
        return 1;
    }

    void *stream = ffx_stream_create();
    void *ev[4] = {ffx_event_create(), ffx_event_create(), ffx_event_create(), ffx_event_create()};
    ffx_sieve_plan *sp = nullptr;
    ffx_search_plan *qp = nullptr;
    void *dMap = nullptr;
    uint32_t *dSlots = nullptr;
    ffx_interval *dIv = nullptr;
    ffx_graph *graph = nullptr;
    AtExit release([&] {
        if (stream) ffx_stream_sync(stream);
        ffx_graph_destroy(graph);
        ffx_free(dIv);
        ffx_free(dSlots);
        ffx_free(dMap);
        ffx_search_plan_destroy(qp);
        ffx_sieve_plan_destroy(sp);
        for (void *e : ev) ffx_event_destroy(e);
        ffx_stream_destroy(stream);
    });
    if (!stream || !ev[0] || !ev[1] || !ev[2] || !ev[3]) return fail("stream/events", FFX_EIO);
    int rc;
    if ((rc = ffx_sieve_plan_create(&sp, M))) return fail("sieve plan", rc);
    if ((rc = ffx_search_plan_create(&qp, U.lastSum() > 5 ? U.lastSum() : 5))) return fail("search plan", rc);
    if ((rc = applyTuning(cfg, sp, qp))) return fail("tuning", rc);
    dMap = ffx_alloc((mapBytes + 255) & ~255ull);
    dSlots = (uint32_t *)ffx_alloc(U.totalSlots * 4 + 4);
    dIv = (ffx_interval *)ffx_alloc(U.iv.size() * sizeof(ffx_interval) + 16);
    if (!dMap || !dSlots || !dIv) return fail("allocation", FFX_ENOMEM);
    if (!U.iv.empty() && (rc = ffx_copy_h2d(dIv, U.iv.data(), U.iv.size() * sizeof(ffx_interval))))
        return fail("interval upload", rc);

    // sieve once
    const auto t0 = Clock::now();
    ffx_event_record(ev[0], stream);
    if ((rc = ffx_sieve_map_async(sp, dMap, mapBytes, stream))) return fail("sieve", rc);
    ffx_event_record(ev[1], stream);
    if ((rc = ffx_event_sync(ev[1]))) return fail("sieve sync", rc);
    double sieveMs = 0;
    ffx_event_elapsed_ms(ev[0], ev[1], &sieveMs);
    const double sieveWall = msSince(t0);

    // one persistent launch over every interval, captured once, replayed
    if (opt.useGraph && !U.iv.empty()) {
        if ((rc = ffx_graph_begin(stream))) return fail("graph capture", rc);
        rc = ffx_search_intervals_async(qp, dMap, M, dIv, (uint32_t)U.iv.size(), U.totalSlots, U.lastSum(), dSlots,
                                        stream);
        ffx_graph *g = nullptr;
        const int rc2 = ffx_graph_end(stream, &g);
        if (rc) return fail("batched search (capture)", rc);
        if (rc2) return fail("graph instantiate", rc2);
        graph = g;
    }
    std::vector<double> runs;
    const auto tS = Clock::now();
    for (uint32_t r = 0; r < (opt.reps ? opt.reps : 1); ++r) {
        ffx_event_record(ev[2], stream);
        if (graph) rc = ffx_graph_launch(graph, stream);
        else if (!U.iv.empty())
            rc = ffx_search_intervals_async(qp, dMap, M, dIv, (uint32_t)U.iv.size(), U.totalSlots, U.lastSum(),
                                            dSlots, stream);
        if (rc) return fail("batched search", rc);
        ffx_event_record(ev[3], stream);
        if ((rc = ffx_event_sync(ev[3]))) return fail("search sync", rc);
        double ms = 0;
        ffx_event_elapsed_ms(ev[2], ev[3], &ms);
        runs.push_back(ms);
    }
    const double searchWall = msSince(tS) / (double)(opt.reps ? opt.reps : 1);
    std::vector<uint32_t> slots(U.totalSlots + 1);
    if (U.totalSlots && (rc = ffx_copy_d2h(slots.data(), dSlots, U.totalSlots * 4))) return fail("slot copy", rc);
    const double searchMs = *std::min_element(runs.begin(), runs.end());

    // per-range stdout, each exactly a plain run's, formatted by 8 host
    // threads taking ranges in turn (hashed for --opt hashes / verify)
    const auto tE = Clock::now();
    if (!opt.outDir.empty()) mkdir(opt.outDir.c_str(), 0755);
    const bool wantDigest = opt.verify || !opt.hashOut.empty();
    std::vector<Digest> digests(wantDigest ? jobs.size() : 0);
    std::atomic<size_t> nextRange{0};
    std::atomic<uint64_t> lineCount{0};
    std::atomic<bool> writeFailed{false};
    auto emitter = [&] {
        std::string out;
        uint64_t nl = 0;
        for (size_t i; (i = nextRange.fetch_add(1)) < jobs.size();) {
            const Digest d = digestOf(jobs[i], [&](uint32_t s) {
                const int64_t k = U.slotOf(s);
                return k < 0 ? 0u : slots[(size_t)k];
            }, out);
            nl += d.lines;
            if (wantDigest) digests[i] = d;
            if (!opt.outDir.empty()) {
                if (!jobs[i].geom.fatal) appendTimes(out, (uint64_t)(sieveWall * 1000), (uint64_t)(searchWall * 1000));
                char name[64];
                std::snprintf(name, sizeof name, "/range_%05zu.txt", i);
                if (!writeFile(opt.outDir + name, out)) writeFailed = true;
            }
        }
        lineCount += nl;
    };
    std::vector<std::thread> emitters;
    for (int t = 0; t < 8; ++t) emitters.emplace_back(emitter);
    for (auto &t : emitters) t.join();
    if (writeFailed) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write the range files under %s\n", opt.outDir.c_str());
        return 1;
    }
    const uint64_t lines = lineCount;
    const double emitMs = msSince(tE);
    if (!opt.hashOut.empty() && !writeDigests(opt.hashOut, jobs, digests)) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write %s\n", opt.hashOut.c_str());
        return 1;
    }
    PlainCheck pc;
    if (opt.verify && (rc = plainCheck(qp, dMap, M, U.lastSum(), jobs, digests, opt.slotsOut, stream, &pc)))
        return rc;
    double searchMean = 0;
    for (double r : runs) searchMean += r / (double)runs.size();
    std::printf("{\"mode\": \"batch\", \"ranges\": %zu, \"requested_odd_sums\": %llu, \"distinct_odd_sums\": %llu, "
                "\"intervals\": %zu, \"solution_lines\": %llu, \"map_value\": %llu, \"map_bytes\": %llu, "
                "\"sieve_ms\": %.3f, \"search_ms\": %.3f, \"search_ms_mean\": %.3f, \"search_runs\": %zu, "
                "\"emit_ms\": %.3f, \"graph\": %s, \"distinct_sums_per_s\": %.1f, "
                "\"requested_sums_per_s\": %.1f, \"overlap\": %.2f, \"sieve_gbs\": %.1f, %s}\n",
                jobs.size(), (unsigned long long)requested, (unsigned long long)U.totalSlots, U.iv.size(),
                (unsigned long long)lines, (unsigned long long)M, (unsigned long long)mapBytes, sieveMs, searchMs,
                searchMean, runs.size(), emitMs, graph ? "true" : "false", (double)U.totalSlots / (searchMs * 1e-3),
                (double)requested / ((sieveMs + searchMs) * 1e-3),
                U.totalSlots ? (double)requested / (double)U.totalSlots : 0.0, (double)mapBytes / (sieveMs * 1e6),
                parityJson(opt.verify, pc).c_str());
    return pc.mismatched ? 3 : 0;
}

// --------------------------------------------------------------------------
// served
// --------------------------------------------------------------------------

namespace {

using Spans = std::vector<std::pair<uint32_t, uint32_t>>;

// Merges b into the sorted disjoint odd-sum span list a.
void mergeInto(Spans &a, const Spans &b)
{
    if (b.empty()) return;
    Spans all(a);
    all.insert(all.end(), b.begin(), b.end());
    std::sort(all.begin(), all.end());
    a.clear();
    for (const auto &x : all) {
        if (!a.empty() && (uint64_t)x.first <= (uint64_t)a.back().second + 2) a.back().second = std::max(a.back().second, x.second);
        else a.push_back(x);
    }
}

bool covers(const Spans &a, uint32_t lo, uint32_t hi)
{
    auto it = std::upper_bound(a.begin(), a.end(), std::make_pair(lo, 0xFFFFFFFFu));
    if (it == a.begin()) return false;
    --it;
    return it->first <= lo && it->second >= hi;
}

// Output formatting pool: completed requests are formatted off the
// scheduling thread; completion time is when the text is ready.
struct EmitPool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<size_t> q;
    bool done = false;
    std::vector<std::thread> th;
    std::function<void(size_t)> work;
    void start(int n, std::function<void(size_t)> w)
    {
        work = std::move(w);
        for (int i = 0; i < n; ++i)
            th.emplace_back([this] {
                for (;;) {
                    size_t id;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [this] { return done || !q.empty(); });
                        if (q.empty()) return;
                        id = q.front();
                        q.pop_front();
                    }
                    work(id);
                }
            });
    }
    void push(size_t id)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            q.push_back(id);
        }
        cv.notify_one();
    }
    ~EmitPool() { finish(); }
    void finish()
    {
        if (th.empty()) return;
        {
            std::lock_guard<std::mutex> lk(mu);
            done = true;
        }
        cv.notify_all();
        for (auto &t : th) t.join();
        th.clear();
    }
};

struct InFlight {
    int stream;
    Spans spans;
};

}  // namespace

int runServe(const Config &cfg, const ServeOptions &opt)
{
    const std::vector<Request> reqs =
        servedRequests(opt.seed, opt.count, opt.limLo, opt.limHi, opt.wLo, opt.wHi, opt.rate);
    const size_t n = reqs.size();
    std::vector<RangeJob> jobs(n);
    uint64_t requested = 0;
    uint32_t maxSum = 5;
    for (size_t i = 0; i < n; ++i) {
        jobs[i] = makeJob(reqs[i].sumStart, reqs[i].sumLimit, cfg.primeMem);
        const RangeJob &j = jobs[i];
        if (j.geom.fatal || !j.any) continue;
        maxSum = std::max(maxSum, j.hi);
        requested += (j.hi - j.lo) / 2 + 1;
    }
    if (maxSum > FFX_MAX_SUM) {
        std::fprintf(stderr, "[ff_sieve] error: sums above %u\n", FFX_MAX_SUM);
        return 1;
    }
    DeviceSetup dev;
    if (setupDevice(cfg, &dev)) return 1;
    const int S = (int)std::max(1u, std::min(opt.streams, 4u));
    const uint32_t base = 5;
    const uint64_t nslots = (maxSum - base) / 2 + 1;
    const uint32_t ivCap = 1u << 16;  // intervals per batch
    const uint64_t batchSums = std::max<uint64_t>(64, opt.batchSums);
    const uint64_t fixedBytes = nslots * 4 + (uint64_t)S * ivCap * sizeof(ffx_interval);
    // Admission against the device budget (source/budget.cpp:15-42): the
    // resident map gets what the backing holds after the fixed buffers.
    // Probe values above the map are answered by Miller-Rabin, so the bound
    // only trades speed for memory.
    const uint64_t M = serveMapValue(dev.budget.backing, fixedBytes, maxSum, cfg.mapCap);
    if (!M) {
        std::fprintf(stderr, "[ff_sieve] ERROR: budget %s cannot hold the serving buffers\n",
                     bytesToHuman(dev.budget.backing).c_str());
        return 1;
    }
    const uint64_t mapBytesMax = (M >> 4) + 1;
    std::fprintf(stderr,
                 "[ff_sieve] serve: %zu requests at %.0f/s, %llu requested odd sums up to %u; resident map up to "
                 "M=%llu (%s, grown on demand), budget backing %s, %d streams, %llu sums per batch\n",
                 n, opt.rate, (unsigned long long)requested, maxSum, (unsigned long long)M,
                 bytesToHuman(mapBytesMax).c_str(), bytesToHuman(dev.budget.backing).c_str(), S,
                 (unsigned long long)batchSums);

    int rc;
    ffx_sieve_plan *sp = nullptr;
    std::vector<ffx_search_plan *> qp(S, nullptr);
    std::vector<void *> st(S, nullptr), ev(S, nullptr);
    std::vector<ffx_interval *> dIv(S, nullptr), hIv(S, nullptr);
    void *gst = nullptr, *dMap = nullptr;
    uint32_t *dSlots = nullptr, *mirror = nullptr;
    std::vector<void *> growEvents;
    AtExit release([&] {
        ffx_device_sync();
        ffx_free_pinned(mirror);
        ffx_free(dSlots);
        ffx_free(dMap);
        ffx_stream_destroy(gst);
        for (void *e : growEvents) ffx_event_destroy(e);
        for (int k = 0; k < S; ++k) {
            ffx_search_plan_destroy(qp[k]);
            ffx_stream_destroy(st[k]);
            ffx_event_destroy(ev[k]);
            ffx_free(dIv[k]);
            ffx_free_pinned(hIv[k]);
        }
        ffx_sieve_plan_destroy(sp);
    });
    if ((rc = ffx_sieve_plan_create(&sp, M))) return fail("sieve plan", rc);
    for (int k = 0; k < S; ++k) {
        if ((rc = ffx_search_plan_create(&qp[k], maxSum))) return fail("search plan", rc);
        if ((rc = applyTuning(cfg, k == 0 ? sp : nullptr, qp[k]))) return fail("tuning", rc);
        st[k] = ffx_stream_create();
        ev[k] = ffx_event_create();
        dIv[k] = (ffx_interval *)ffx_alloc((uint64_t)ivCap * sizeof(ffx_interval));
        hIv[k] = (ffx_interval *)ffx_alloc_pinned((uint64_t)ivCap * sizeof(ffx_interval));
        if (!st[k] || !ev[k] || !dIv[k] || !hIv[k]) return fail("stream resources", FFX_ENOMEM);
    }
    gst = ffx_stream_create();
    dMap = ffx_alloc((mapBytesMax + 255) & ~255ull);
    dSlots = (uint32_t *)ffx_alloc(nslots * 4);
    mirror = (uint32_t *)ffx_alloc_pinned(nslots * 4);
    if (!gst || !dMap || !dSlots || !mirror) return fail("allocation", FFX_ENOMEM);

    // The first GiB of the map is sieved before the stream opens (small
    // values are always probed).  The rest grows on demand on its own stream,
    // queued in 16 GiB launches as soon as admitted requests need it, each
    // with an event, so the finished prefix advances while batches run.  A
    // batch waits until the map covers ff_seg's bound for its largest sum:
    // a search over a partial map would decide the missing probe values by
    // Miller-Rabin, several times slower than waiting for the sieve.
    const uint64_t growStep = 16ull << 30;
    const auto tm0 = Clock::now();
    uint64_t mapBytesDone = opt.warm ? mapBytesMax : std::min<uint64_t>(mapBytesMax, 1ull << 30), growths = 1;
    if ((rc = ffx_sieve_map_range_async(sp, dMap, 0, mapBytesDone, gst)) || (rc = ffx_stream_sync(gst)))
        return fail("map", rc);
    const double mapBuildS = std::chrono::duration<double>(Clock::now() - tm0).count();
    uint64_t wantBytes = mapBytesDone, issuedBytes = mapBytesDone;
    std::deque<std::pair<void *, uint64_t>> growQ;  // (event, map bytes done when it fires)
    auto bytesFor = [&](uint32_t L) { return (std::min(M, uncappedMapValue(L)) >> 4) + 1; };

    std::vector<double> doneAt(n, -1.0);
    Spans computed, claimed;  // claimed: computed, running or queued
    std::deque<std::pair<uint32_t, uint32_t>> work;  // queued spans, arrival order
    std::deque<InFlight> inflight;
    std::vector<bool> busy(S, false);
    std::vector<size_t> waiting;
    uint64_t batches = 0, searchedSlots = 0, lines = 0;
    std::mutex linesMu;
    const bool writeOut = !opt.outDir.empty();
    if (writeOut) mkdir(opt.outDir.c_str(), 0755);

    // search busy time: per batch, events around its upload and launch,
    // read against an origin event after the stream ends
    std::vector<std::pair<void *, void *>> batchEv;
    void *evOrigin = ffx_event_create();
    if (!evOrigin) return fail("origin event", FFX_ENOMEM);
    growEvents.push_back(evOrigin);
    if ((rc = ffx_event_record(evOrigin, gst))) return fail("origin event", rc);
    const auto t0 = Clock::now();
    auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    // a request is done when its stdout is formatted; its digest (when
    // asked for) is taken after the completion time is recorded
    const bool wantDigest = opt.verify || !opt.hashOut.empty();
    std::vector<Digest> digests(wantDigest ? n : 0);
    std::atomic<bool> writeFailed{false};
    EmitPool pool;
    pool.start(8, [&](size_t id) {
        std::string out;
        uint64_t nl = 0;
        if (jobs[id].geom.fatal) out = fatalText(jobs[id]);
        else nl = emitBody(out, jobs[id], [&](uint32_t s) { return mirror[(s - base) >> 1]; });
        if (writeOut) {
            if (!jobs[id].geom.fatal) appendTimes(out, 0, 0);
            char name[64];
            std::snprintf(name, sizeof name, "/request_%05zu.txt", id);
            if (!writeFile(opt.outDir + name, out)) writeFailed = true;
        }
        doneAt[id] = now();
        if (wantDigest) {
            Digest d;
            d.lines = nl;
            if (jobs[id].geom.fatal) {
                d.hex = sha256Hex(out);
            } else {
                if (writeOut) out.resize(out.size() - (sizeof "Prime time: 0 \xce\xbcs\nFreudenthal time: 0 \xce\xbcs\n" - 1));
                d.hex = normalizedDigest(out);
            }
            digests[id] = std::move(d);
        }
        std::lock_guard<std::mutex> lk(linesMu);
        lines += nl;
    });

    size_t next = 0, completed = 0;
    bool idleScan = false;  // the last scan found no queued sum the map covers
    while (completed < n) {
        bool progress = false, admitted = false, grew = false;
        // admission: arrived requests queue the sums nothing has claimed
        for (const double t = now(); next < n && reqs[next].arrivalS <= t; ++next) {
            const RangeJob &j = jobs[next];
            progress = admitted = true;
            if (j.geom.fatal || !j.any) {
                pool.push(next);
                ++completed;
                continue;
            }
            const Spans missing = subtract(Spans{{j.lo, j.hi}}, claimed);
            for (const auto &m : missing) work.push_back(m);
            mergeInto(claimed, missing);
            waiting.push_back(next);
            wantBytes = std::max(wantBytes, bytesFor(j.hi));
        }
        // map growth, concurrent with the batches
        while (issuedBytes < wantBytes) {
            const uint64_t to = std::min(wantBytes, issuedBytes + growStep);
            void *e = ffx_event_create();
            if (!e) return fail("growth event", FFX_ENOMEM);
            growEvents.push_back(e);
            if ((rc = ffx_sieve_map_range_async(sp, dMap, issuedBytes, to, gst)) || (rc = ffx_event_record(e, gst)))
                return fail("map growth", rc);
            growQ.push_back({e, to});
            issuedBytes = to;
            ++growths;
        }
        while (!growQ.empty()) {
            const int q = ffx_event_query(growQ.front().first);
            if (q < 0) return fail("map growth", q);
            if (q != 0) break;
            mapBytesDone = growQ.front().second;
            growQ.pop_front();
            progress = grew = true;
        }
        // batches: the oldest queued sums that the map covers, at most
        // batchSums per launch (sums still waiting for the map stay queued
        // without holding back the ones behind them)
        if (admitted || grew) idleScan = false;
        for (int k = 0; k < S && !work.empty() && !idleScan; ++k) {
            if (busy[k]) continue;
            std::vector<ffx_interval> iv;
            uint64_t total = 0;
            for (auto w = work.begin(); w != work.end() && total < batchSums && iv.size() < ivCap;) {
                if (bytesFor(w->second) > mapBytesDone) {
                    ++w;
                    continue;
                }
                const uint64_t cnt = (w->second - w->first) / 2 + 1, take = std::min(cnt, batchSums - total);
                ffx_interval x;
                x.first = w->first;
                x.last = w->first + 2 * (uint32_t)(take - 1);
                x.slotOffset = 0;
                iv.push_back(x);
                total += take;
                if (take == cnt) {
                    w = work.erase(w);
                } else {
                    w->first = x.last + 2;
                    ++w;
                }
            }
            if (iv.empty()) {
                idleScan = true;
                break;
            }
            std::sort(iv.begin(), iv.end(), [](const ffx_interval &a, const ffx_interval &b) { return a.first < b.first; });
            Spans spans;
            uint64_t off = 0;
            for (ffx_interval &x : iv) {
                x.slotOffset = off;
                off += (x.last - x.first) / 2 + 1;
                spans.push_back({x.first, x.last});
            }
            const uint64_t Muse = std::min(M, (mapBytesDone << 4) - 1);
            std::memcpy(hIv[k], iv.data(), iv.size() * sizeof(ffx_interval));
            void *eb = ffx_event_create(), *ee = ffx_event_create();
            if (!eb || !ee) return fail("batch events", FFX_ENOMEM);
            growEvents.push_back(eb);
            growEvents.push_back(ee);
            batchEv.push_back({eb, ee});
            if ((rc = ffx_event_record(eb, st[k])) ||
                (rc = ffx_copy_h2d_async(dIv[k], hIv[k], iv.size() * sizeof(ffx_interval), st[k])))
                return fail("interval upload", rc);
            if ((rc = ffx_search_intervals_by_sum_async(qp[k], dMap, Muse, dIv[k], (uint32_t)iv.size(), total,
                                                        iv.back().last, base, dSlots, nslots, st[k])) ||
                (rc = ffx_event_record(ee, st[k])))
                return fail("served search", rc);
            for (const ffx_interval &x : iv) {
                const uint64_t o = (x.first - base) >> 1, cnt = (x.last - x.first) / 2 + 1;
                if ((rc = ffx_copy_d2h_async(mirror + o, dSlots + o, cnt * 4, st[k]))) return fail("slot copy", rc);
            }
            if ((rc = ffx_event_record(ev[k], st[k]))) return fail("event", rc);
            busy[k] = true;
            InFlight f;
            f.stream = k;
            f.spans = std::move(spans);
            inflight.push_back(std::move(f));
            ++batches;
            searchedSlots += total;
            progress = true;
        }
        // finished batches
        bool landed = false;
        for (auto it = inflight.begin(); it != inflight.end();) {
            const int q = ffx_event_query(ev[it->stream]);
            if (q < 0) return fail("event query", q);
            if (q == 0) {
                mergeInto(computed, it->spans);
                busy[it->stream] = false;
                it = inflight.erase(it);
                landed = true;
            } else {
                ++it;
            }
        }
        // requests whose sums are all in the mirror go to the formatting pool
        if (landed || admitted) {
            progress = true;
            for (size_t i = 0; i < waiting.size();) {
                const size_t id = waiting[i];
                if (covers(computed, jobs[id].lo, jobs[id].hi)) {
                    pool.push(id);
                    ++completed;
                    waiting[i] = waiting.back();
                    waiting.pop_back();
                } else {
                    ++i;
                }
            }
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    pool.finish();
    const double makespan = now();
    if (writeFailed) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write the request files under %s\n", opt.outDir.c_str());
        return 1;
    }
    if (!growQ.empty()) {
        if ((rc = ffx_stream_sync(gst))) return fail("map growth", rc);
        mapBytesDone = growQ.back().second;
        growQ.clear();
    }
    std::vector<double> lat(n);
    for (size_t i = 0; i < n; ++i) lat[i] = (doneAt[i] - reqs[i].arrivalS) * 1000.0;
    const double p50 = percentile(lat, 0.50), p99 = percentile(lat, 0.99);
    // device time spent searching: the union of the batches' intervals
    std::vector<std::pair<double, double>> busyIv;
    for (const auto &b : batchEv) {
        double t0ms = 0, t1ms = 0;
        if ((rc = ffx_event_elapsed_ms(evOrigin, b.first, &t0ms)) || (rc = ffx_event_elapsed_ms(evOrigin, b.second, &t1ms)))
            return fail("batch event times", rc);
        busyIv.push_back({t0ms, t1ms});
    }
    std::sort(busyIv.begin(), busyIv.end());
    double busyMs = 0, curLo = 0, curHi = -1;
    for (const auto &b : busyIv) {
        if (b.first > curHi) {
            if (curHi > curLo) busyMs += curHi - curLo;
            curLo = b.first;
            curHi = b.second;
        } else {
            curHi = std::max(curHi, b.second);
        }
    }
    if (curHi > curLo) busyMs += curHi - curLo;

    // Sequential plain launches of plainSample requests spread evenly over
    // the stream (every n/plainSample-th), replayed at their arrival times
    // relative to the first of them: FIFO, one request at a time through the
    // plain entry point (ffx_search_slots_async), nothing shared between
    // requests, the map resident at its full size.  The served latencies of
    // the same requests are reported beside them.
    double pp50 = 0, pp99 = 0, sp50 = 0, sp99 = 0, plainBusyRate = 0, plainSpan = 0, plainBusyS = 0;
    uint64_t plainSums = 0;
    if (opt.plainSample) {
        const size_t m = std::min<size_t>(opt.plainSample, n);
        std::vector<size_t> pick;
        for (size_t j = 0; j < m; ++j) pick.push_back(j * n / m);
        std::vector<double> plat, slat;
        uint32_t maxCnt = 1;
        for (size_t i : pick)
            if (jobs[i].any && !jobs[i].geom.fatal) maxCnt = std::max<uint32_t>(maxCnt, (jobs[i].hi - jobs[i].lo) / 2 + 1);
        uint32_t *dS = (uint32_t *)ffx_alloc((uint64_t)maxCnt * 4);
        AtExit freeS([&] {
            ffx_device_sync();
            ffx_free(dS);
        });
        if (!dS) return fail("plain slots", FFX_ENOMEM);
        if (mapBytesDone < mapBytesMax) {
            if ((rc = ffx_sieve_map_range_async(sp, dMap, mapBytesDone, mapBytesMax, gst)) ||
                (rc = ffx_stream_sync(gst)))
                return fail("map", rc);
            mapBytesDone = mapBytesMax;
        }
        std::vector<uint32_t> hostSlots(maxCnt);
        const auto p0 = Clock::now();
        const double a0 = reqs[pick[0]].arrivalS;
        for (size_t i : pick) {
            const RangeJob &j = jobs[i];
            const double arr = reqs[i].arrivalS - a0;
            const double nowS = std::chrono::duration<double>(Clock::now() - p0).count();
            if (nowS < arr) std::this_thread::sleep_for(std::chrono::duration<double>(arr - nowS));
            const auto ts = Clock::now();
            if (!j.geom.fatal && j.any) {
                const uint64_t cnt = (j.hi - j.lo) / 2 + 1;
                if ((rc = ffx_search_slots_async(qp[0], dMap, M, j.lo, j.hi, dS, st[0])) ||
                    (rc = ffx_copy_d2h_async(hostSlots.data(), dS, cnt * 4, st[0])) || (rc = ffx_stream_sync(st[0])))
                    return fail("plain search", rc);
                plainSums += cnt;
            }
            std::string out;
            if (!j.geom.fatal) emitJob(out, j, [&](uint32_t x) { return hostSlots[(x - j.lo) >> 1]; }, 0, 0);
            plainBusyS += std::chrono::duration<double>(Clock::now() - ts).count();
            plat.push_back((std::chrono::duration<double>(Clock::now() - p0).count() - arr) * 1000.0);
            slat.push_back(lat[i]);
        }
        plainSpan = std::chrono::duration<double>(Clock::now() - p0).count();
        pp50 = percentile(plat, 0.5);
        pp99 = percentile(plat, 0.99);
        sp50 = percentile(slat, 0.5);
        sp99 = percentile(slat, 0.99);
        plainBusyRate = plainBusyS > 0 ? (double)plainSums / plainBusyS : 0;
    }
    if (!opt.hashOut.empty() && !writeDigests(opt.hashOut, jobs, digests)) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write %s\n", opt.hashOut.c_str());
        return 1;
    }
    PlainCheck pc;
    if (opt.verify) {
        if (mapBytesDone < mapBytesMax) {
            if ((rc = ffx_sieve_map_range_async(sp, dMap, mapBytesDone, mapBytesMax, gst)) ||
                (rc = ffx_stream_sync(gst)))
                return fail("map", rc);
            mapBytesDone = mapBytesMax;
        }
        if ((rc = plainCheck(qp[0], dMap, M, maxSum, jobs, digests, opt.slotsOut, st[0], &pc))) return rc;
    }
    char js[2560];
    std::snprintf(js, sizeof js,
                  "{\"mode\": \"serve\", \"requests\": %zu, \"rate_per_s\": %.1f, \"requested_odd_sums\": %llu, "
                  "\"searched_odd_sums\": %llu, \"batches\": %llu, \"batch_sums\": %llu, \"streams\": %d, "
                  "\"makespan_s\": %.3f, \"offered_sums_per_s\": %.1f, \"search_busy_s\": %.3f, "
                  "\"searched_sums_per_busy_s\": %.1f, \"searched_sums_per_s\": %.1f, "
                  "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"map_value\": %llu, \"map_bytes\": %llu, "
                  "\"map_growths\": %llu, \"solution_lines\": %llu, \"plain_sample\": %u, \"plain_span_s\": %.3f, "
                  "\"plain_busy_s\": %.3f, \"plain_sums_per_busy_s\": %.1f, \"plain_p50_ms\": %.3f, "
                  "\"plain_p99_ms\": %.3f, \"served_sample_p50_ms\": %.3f, \"served_sample_p99_ms\": %.3f, "
                  "\"warm\": %s, \"map_prebuilt_s\": %.3f, %s}\n",
                  n, opt.rate, (unsigned long long)requested, (unsigned long long)searchedSlots,
                  (unsigned long long)batches, (unsigned long long)batchSums, S, makespan,
                  (double)requested / makespan, busyMs * 1e-3,
                  busyMs > 0 ? (double)searchedSlots / (busyMs * 1e-3) : 0.0, (double)searchedSlots / makespan, p50,
                  p99, (unsigned long long)M, (unsigned long long)mapBytesMax, (unsigned long long)growths,
                  (unsigned long long)lines, opt.plainSample, plainSpan, plainBusyS, plainBusyRate, pp50, pp99, sp50,
                  sp99, opt.warm ? "true" : "false", mapBuildS, parityJson(opt.verify, pc).c_str());
    std::fputs(js, stdout);
    if (!opt.jsonOut.empty() && !writeFile(opt.jsonOut, js)) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write %s\n", opt.jsonOut.c_str());
        return 1;
    }
    return pc.mismatched ? 3 : 0;
}

// --------------------------------------------------------------------------
// autotune
// --------------------------------------------------------------------------

// Times every candidate of the kernels' launch parameters on one workload
// (odd sums [5, sumLimit] over ff_seg's map for it) and writes the fastest
// as a tune file for --tune.  Every candidate's output is compared with the
// default configuration's: the sieve on a smaller map byte for byte, the
// search slot for slot.
int runAutotune(const Config &cfg, const AutotuneOptions &opt)
{
    DeviceSetup dev;
    if (setupDevice(cfg, &dev)) return 1;
    const RangeJob job = makeJob(5, opt.sumLimit, cfg.primeMem);
    if (job.geom.fatal || !job.any) {
        std::fprintf(stderr, "[ff_sieve] autotune: empty workload\n");
        return 1;
    }
    const uint64_t M = mapValueFor(job.geom.maxPrimeMapValue, cfg.mapCap);
    const uint64_t mapBytes = (M >> 4) + 1;
    const uint64_t nslots = (job.hi - 5) / 2 + 1;
    const uint32_t reps = std::max(1u, opt.reps);
    // the sieve's byte-for-byte check runs on a map of at most 256 MiB
    const uint64_t checkBytes = std::min<uint64_t>(mapBytes, 256ull << 20);
    if (mapBytes + checkBytes * 2 + nslots * 8 > dev.budget.backing) {
        std::fprintf(stderr, "[ff_sieve] autotune: workload needs more than the budget; bound it with --map-cap\n");
        return 1;
    }
    int rc;
    void *stream = ffx_stream_create();
    void *e0 = ffx_event_create(), *e1 = ffx_event_create();
    uint8_t *dMap = (uint8_t *)ffx_alloc((mapBytes + 255) & ~255ull);
    uint8_t *dChk = (uint8_t *)ffx_alloc((checkBytes + 255) & ~255ull);
    uint32_t *dSlots = (uint32_t *)ffx_alloc(nslots * 4);
    ffx_sieve_plan *sp = nullptr;
    ffx_search_plan *qp = nullptr;
    AtExit release([&] {
        if (stream) ffx_stream_sync(stream);
        ffx_search_plan_destroy(qp);
        ffx_sieve_plan_destroy(sp);
        ffx_free(dSlots);
        ffx_free(dChk);
        ffx_free(dMap);
        ffx_event_destroy(e0);
        ffx_event_destroy(e1);
        ffx_stream_destroy(stream);
    });
    if (!stream || !e0 || !e1 || !dMap || !dChk || !dSlots) return fail("autotune buffers", FFX_ENOMEM);
    std::vector<uint8_t> chkRef(checkBytes), chk(checkBytes);
    std::vector<uint32_t> slotsRef(nslots), slots(nslots);

    auto timeIt = [&](auto &&launch, double *best) -> int {
        *best = 1e30;
        for (uint32_t r = 0; r < reps; ++r) {
            int q;
            if ((q = ffx_event_record(e0, stream)) || (q = launch()) || (q = ffx_event_record(e1, stream)) ||
                (q = ffx_event_sync(e1)))
                return q;
            double ms = 0;
            ffx_event_elapsed_ms(e0, e1, &ms);
            *best = std::min(*best, ms);
        }
        return 0;
    };

    // sieve: LDS tile (segment bytes) x workgroups per CU
    struct SieveTrial {
        uint32_t seg, bpc;
        double ms;
    };
    std::vector<SieveTrial> st;
    if ((rc = ffx_sieve_plan_create(&sp, M))) return fail("sieve plan", rc);
    if ((rc = ffx_sieve_map_async(sp, dChk, checkBytes, stream)) || (rc = ffx_stream_sync(stream)) ||
        (rc = ffx_copy_d2h(chkRef.data(), dChk, checkBytes)))
        return fail("sieve reference", rc);
    for (uint32_t seg : {32768u, 114688u})
        for (uint32_t bpc : {1u, 2u, 3u, 4u}) {
            if ((rc = ffx_sieve_plan_set(sp, FFX_SIEVE_SEGMENT_BYTES, seg)) ||
                (rc = ffx_sieve_plan_set(sp, FFX_SIEVE_BLOCKS_PER_CU, bpc)))
                return fail("sieve set", rc);
            if ((rc = ffx_memset_async(dChk, 0, checkBytes, stream)) ||
                (rc = ffx_sieve_map_async(sp, dChk, checkBytes, stream)) || (rc = ffx_stream_sync(stream)) ||
                (rc = ffx_copy_d2h(chk.data(), dChk, checkBytes)))
                return fail("sieve candidate", rc);
            if (chk != chkRef) {
                std::fprintf(stderr, "[ff_sieve] autotune: sieve segment %u x %u per CU differs from the default\n",
                             seg, bpc);
                return 1;
            }
            double ms;
            if ((rc = timeIt([&] { return ffx_sieve_map_async(sp, dMap, mapBytes, stream); }, &ms)))
                return fail("sieve timing", rc);
            st.push_back({seg, bpc, ms});
            std::fprintf(stderr, "[ff_sieve] autotune sieve segment=%u blocks/CU=%u: %.3f ms\n", seg, bpc, ms);
        }
    const SieveTrial bs = *std::min_element(st.begin(), st.end(), [](auto &a, auto &b) { return a.ms < b.ms; });
    // the search reads the map the default configuration wrote
    if ((rc = ffx_sieve_plan_set(sp, FFX_SIEVE_SEGMENT_BYTES, 114688)) ||
        (rc = ffx_sieve_map_async(sp, dMap, mapBytes, stream)) || (rc = ffx_stream_sync(stream)))
        return fail("sieve", rc);

    // search: sums per atomic grab x eager-drain threshold x workgroups per CU
    struct SearchTrial {
        uint32_t grab, eager, bpc;
        double ms;
    };
    std::vector<SearchTrial> qt;
    if ((rc = ffx_search_plan_create(&qp, job.hi))) return fail("search plan", rc);
    uint32_t g0 = 0, e0v = 0, b0 = 0;
    ffx_search_plan_get(qp, FFX_SEARCH_GRAB, &g0);
    ffx_search_plan_get(qp, FFX_SEARCH_EAGER_SPLITS, &e0v);
    ffx_search_plan_get(qp, FFX_SEARCH_BLOCKS_PER_CU, &b0);
    if ((rc = ffx_search_slots_async(qp, dMap, M, 5, job.hi, dSlots, stream)) || (rc = ffx_stream_sync(stream)) ||
        (rc = ffx_copy_d2h(slotsRef.data(), dSlots, nslots * 4)))
        return fail("search reference", rc);
    double defaultMs = 0;
    for (uint32_t grab : {1u, 2u, 4u})
        for (uint32_t eager : {256u, 1024u, 4096u})
            for (uint32_t bpc : {4u, 5u, 6u, 8u}) {
                if ((rc = ffx_search_plan_set(qp, FFX_SEARCH_GRAB, grab)) ||
                    (rc = ffx_search_plan_set(qp, FFX_SEARCH_EAGER_SPLITS, eager)) ||
                    (rc = ffx_search_plan_set(qp, FFX_SEARCH_BLOCKS_PER_CU, bpc)))
                    return fail("search set", rc);
                double ms;
                if ((rc = timeIt([&] { return ffx_search_slots_async(qp, dMap, M, 5, job.hi, dSlots, stream); }, &ms)))
                    return fail("search timing", rc);
                if ((rc = ffx_copy_d2h(slots.data(), dSlots, nslots * 4))) return fail("slot copy", rc);
                if (slots != slotsRef) {
                    std::fprintf(stderr, "[ff_sieve] autotune: search grab=%u eager=%u blocks/CU=%u differs\n", grab,
                                 eager, bpc);
                    return 1;
                }
                if (grab == g0 && eager == e0v && bpc == b0) defaultMs = ms;
                qt.push_back({grab, eager, bpc, ms});
                std::fprintf(stderr, "[ff_sieve] autotune search grab=%u eager=%u blocks/CU=%u: %.3f ms\n", grab,
                             eager, bpc, ms);
            }
    const SearchTrial bq = *std::min_element(qt.begin(), qt.end(), [](auto &a, auto &b) { return a.ms < b.ms; });

    // the chosen values first (the tune-file reader takes the first match)
    std::string js;
    char buf[1024];
    std::snprintf(buf, sizeof buf,
                  "{\"search_grab\": %u, \"search_eager_splits\": %u, \"search_blocks_per_cu\": %u, "
                  "\"sieve_segment_bytes\": %u, \"sieve_blocks_per_cu\": %u,\n \"workload\": \"odd sums [5, %u], map "
                  "M=%llu\", \"device\": \"%s %s, %d CUs\",\n \"search_ms\": %.3f, \"search_default_ms\": %.3f, "
                  "\"sieve_ms\": %.3f,\n",
                  bq.grab, bq.eager, bq.bpc, bs.seg, bs.bpc, job.hi, (unsigned long long)M, dev.info.name[0] ? dev.info.name : "AMD",
                  dev.info.arch, dev.info.multiProcessorCount, bq.ms, defaultMs, bs.ms);
    js += buf;
    js += " \"trials\": {\"sieve\": [";
    for (size_t i = 0; i < st.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%s[%u, %u, %.3f]", i ? ", " : "", st[i].seg, st[i].bpc, st[i].ms);
        js += buf;
    }
    js += "],\n  \"search\": [";
    for (size_t i = 0; i < qt.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%s[%u, %u, %u, %.3f]", i ? ", " : "", qt[i].grab, qt[i].eager, qt[i].bpc,
                      qt[i].ms);
        js += buf;
    }
    js += "],\n  \"columns\": {\"sieve\": [\"segment bytes\", \"blocks per CU\", \"ms\"], \"search\": [\"grab\", "
          "\"eager splits\", \"blocks per CU\", \"ms\"]}}}\n";
    if (!writeFile(opt.out, js)) {
        std::fprintf(stderr, "[ff_sieve] error: cannot write %s\n", opt.out.c_str());
        return 1;
    }
    std::fputs(js.c_str(), stdout);
    return 0;
}

}  // namespace ffg
