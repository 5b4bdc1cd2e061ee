// Reference-side adapter: ff-gpu's per-architecture dispatch entry points,
// implemented over libffx for gfx950.
//
// A maintainer adds this file to the reference's host build (g++, linked with
// -lffx) and selects the gfx950 table where the reference picks a vendor
// table (source/pull_scheduler.cpp:329-332).  It provides:
//   SievePoolGet_gfx950            -> const SievePoolOps*  (source/sieve_slab_engine.h:49-101)
//   SieveSlabEngineGetLaunchFn_gfx950 -> SieveSlabEngineRunFn (:27-37)
//   SearchKernelGetLaunchFn_gfx950 -> SearchKernelRunFn (source/m4/gpu_search_kernel.cpp:59-66)
// Nothing here launches a kernel itself: every operation is one or two libffx
// calls.  It compiles against the reference's own headers (-I<ref>/source).
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ffx.h"
#include "m4/gpu_search_kernel.h"  // GpuRecord
#include "sieve_slab_engine.h"     // SievePoolOps, SieveSlabEngineRunFn

static_assert(sizeof(GpuRecord) == sizeof(ffx_gpu_record), "GpuRecord is 32 bytes with padding");
static_assert(offsetof(GpuRecord, low) == offsetof(ffx_gpu_record, low), "GpuRecord.low at 8");
static_assert(offsetof(GpuRecord, high) == offsetof(ffx_gpu_record, high), "GpuRecord.high at 16");
static_assert(offsetof(GpuRecord, tag) == offsetof(ffx_gpu_record, tag), "GpuRecord.tag at 24");

namespace {

int sel(int vendorIndex) { return ffx_init(vendorIndex) == FFX_OK ? 0 : -1; }
int rc(int r) { return r == FFX_OK ? 0 : -1; }

// The reference hands the sieve its small-prime list on every slab call
// (with or without a leading 2) and never the map's top value.  A plan needs
// a bound M with every sieving prime <= isqrt(M) and every slab value <= M:
// for primes up to pmax, (next prime after pmax)^2 - 1 is the largest map the
// list can sieve, and the reference's lists always end at isqrt(M).
uint64_t planBound(const uint32_t *p, uint32_t n)
{
    uint64_t q = n ? p[n - 1] + 1 : 3;
    for (;; ++q) {
        bool prime = q >= 2;
        for (uint64_t f = 2; f * f <= q && prime; ++f) prime = q % f != 0;
        if (prime) return q * q - 1;
    }
}

std::mutex g_mu;
std::map<std::tuple<int, const uint32_t *, uint32_t>, ffx_sieve_plan *> g_plans;

ffx_sieve_plan *planFor(int vi, const uint32_t *h_primes, uint32_t n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(vi, h_primes, n);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) return it->second;
    ffx_sieve_plan *pl = nullptr;
    if (ffx_sieve_plan_create_from_primes(&pl, h_primes, n, planBound(h_primes, n)) != FFX_OK) return nullptr;
    g_plans[key] = pl;
    return pl;
}

void *opAlloc(int vi, uint64_t bytes) { return sel(vi) ? nullptr : ffx_alloc(bytes); }
int opMemset(int vi, void *d, int v, uint64_t bytes)
{
    return sel(vi) || ffx_memset_async(d, v, bytes, nullptr) || ffx_device_sync() ? -1 : 0;
}
int opFree(int vi, void *d) { return sel(vi) ? -1 : rc(ffx_free(d)); }
int opH2D(int vi, void *d, const void *h, uint64_t b) { return sel(vi) ? -1 : rc(ffx_copy_h2d(d, h, b)); }
int opD2H(int vi, void *h, const void *d, uint64_t b) { return sel(vi) ? -1 : rc(ffx_copy_d2h(h, d, b)); }

// Priming (0xFF fill, value 1 cleared) is part of the libffx kernel, so
// primeSlab needs no memset here.
int opSlab(int vi, const uint32_t *h_primes, uint32_t n, uint64_t segLo, uint64_t segHi, void *d_slab,
           uint64_t slabBytes, int primeSlab)
{
    (void)primeSlab;
    if (sel(vi) || (segHi - segLo + 15) / 16 > slabBytes) return -1;
    ffx_sieve_plan *pl = planFor(vi, h_primes, n);
    if (!pl) return -1;
    return ffx_sieve_slab_async(pl, segLo, segHi, d_slab, nullptr) || ffx_device_sync() ? -1 : 0;
}

void *opStream(int vi) { return sel(vi) ? nullptr : ffx_stream_create(); }
int opStreamDestroy(int vi, void *s) { return sel(vi) ? -1 : rc(ffx_stream_destroy(s)); }
void *opEvent(int vi) { return sel(vi) ? nullptr : ffx_event_create(); }
int opEventDestroy(int vi, void *e) { return sel(vi) ? -1 : rc(ffx_event_destroy(e)); }
int opRecord(int vi, void *e, void *s) { return sel(vi) ? -1 : rc(ffx_event_record(e, s)); }
int opWait(int vi, void *e, void *s) { return sel(vi) ? -1 : rc(ffx_event_wait(e, s)); }
int opSync(int vi, void *e) { return sel(vi) ? -1 : rc(ffx_event_sync(e)); }
int opElapsed(int vi, void *a, void *b, double *ms) { return sel(vi) ? -1 : rc(ffx_event_elapsed_ms(a, b, ms)); }
int opPinned(int vi, void **h, uint64_t bytes)
{
    if (sel(vi)) return -1;
    *h = ffx_alloc_pinned(bytes);
    return *h ? 0 : -1;
}
int opPinnedFree(int vi, void *h) { return sel(vi) ? -1 : rc(ffx_free_pinned(h)); }
int opH2DAsync(int vi, void *d, const void *h, uint64_t b, void *s)
{
    return sel(vi) ? -1 : rc(ffx_copy_h2d_async(d, h, b, s));
}
int opD2HAsync(int vi, void *h, const void *d, uint64_t b, void *s)
{
    return sel(vi) ? -1 : rc(ffx_copy_d2h_async(h, d, b, s));
}
int opSlabAsync(int vi, const uint32_t *h_primes, uint32_t n, uint64_t segLo, uint64_t segHi, void *d_slab,
                uint64_t slabBytes, void *stream, void *waitEvent)
{
    if (sel(vi) || (segHi - segLo + 15) / 16 > slabBytes) return -1;
    ffx_sieve_plan *pl = planFor(vi, h_primes, n);
    if (!pl) return -1;
    if (waitEvent && ffx_event_wait(waitEvent, stream)) return -1;
    return rc(ffx_sieve_slab_async(pl, segLo, segHi, d_slab, stream));
}

const SievePoolOps kOps = {opAlloc,   opMemset,       opFree,     opH2D,          opD2H,    opSlab,
                           opStream,  opStreamDestroy, opEvent,   opEventDestroy, opRecord, opWait,
                           opSync,    opElapsed,      opPinned,   opPinnedFree,   opH2DAsync, opD2HAsync,
                           opSlabAsync};

// SieveSlabEngineRunFn: the whole map into a host buffer.
int engineRun(int dev, const uint32_t *h_primes, uint32_t n, uint64_t maxPrimeMapValue, uint8_t *h_out)
{
    if (sel(dev)) return -1;
    ffx_sieve_plan *pl = nullptr;
    if (ffx_sieve_plan_create_from_primes(&pl, h_primes, n, maxPrimeMapValue) != FFX_OK) return -1;
    const uint64_t bytes = (maxPrimeMapValue >> 4) + 1;
    void *d = ffx_alloc(bytes);
    int r = d ? 0 : -1;
    if (!r) r = ffx_sieve_map_async(pl, d, bytes, nullptr) || ffx_copy_d2h(h_out, d, bytes) ? -1 : 0;
    if (d) ffx_free(d);
    ffx_sieve_plan_destroy(pl);
    return r;
}

// SearchKernelRunFn: zeroed 32-byte records filled per odd sum, count added.
int searchRun(int dev, const uint8_t *d_map, uint64_t M, uint64_t sumStart, uint64_t sumLimit, uint32_t *d_count,
              GpuRecord *d_records)
{
    return rc(ffx_search_records(dev, d_map, M, sumStart, sumLimit, d_count,
                                 reinterpret_cast<ffx_gpu_record *>(d_records)));
}

bool haveDevice(int i)
{
    int n = 0;
    return ffx_device_count(&n) == FFX_OK && i >= 0 && i < n;
}

}  // namespace

extern "C" const SievePoolOps *SievePoolGet_gfx950(int vendorIndex) { return haveDevice(vendorIndex) ? &kOps : nullptr; }

extern "C" SieveSlabEngineRunFn SieveSlabEngineGetLaunchFn_gfx950(int deviceIndex)
{
    return haveDevice(deviceIndex) ? &engineRun : nullptr;
}

typedef int (*SearchKernelRunFnGfx950)(int, const uint8_t *, uint64_t, uint64_t, uint64_t, uint32_t *, GpuRecord *);
extern "C" SearchKernelRunFnGfx950 SearchKernelGetLaunchFn_gfx950(int deviceIndex)
{
    return haveDevice(deviceIndex) ? &searchRun : nullptr;
}
