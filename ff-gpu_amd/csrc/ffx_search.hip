// libffx search: F(s) for every odd sum, one wavefront per sum.
//
// Replaces FFSearchKernel_<arch> and its launcher
// (source/m4/gpu_search_kernel.h:229-373, gpu_search_kernel.cpp:31-66,
// gpu_search_launcher.cpp:36-57).  Semantics are those of ff_seg's RunIt and
// DoesSammyKnow (reference/segmentedSieve.C:800-823, :529-639):
//   emit s iff  !PTP(s)  and  (3 !| s or IsPrime(s/3))  and  exactly one
//   split s = e + o (e even) is "valid":
//     e = 2^k, o prime                                   (Power2Prime)
//     e = 2^k, o composite, AB1(k, o)                    (Power2Composite)
//     e = d*2^k (k odd >= 3, d >= 5 odd), 3 !| d*o,
//       PTP(2^k+d*o), (d==o or PTP(d+o*2^k)), AB1(k, d*o)  (Power2Odd)
//     e = d*2^k (k even >= 2, d >= 3 odd), (3 | d*o or o prime),
//       same three tests                                 (Power2Even)
// with PTP(i) = IsPrime(i-2) for odd i, and AB1(k, q) true iff exactly one
// divisor f <= isqrt(q*2^k) of q*2^k, f in {2^k} u {d|q, d>1} u
// {2^k*d : d|q, d>1}, has !PTP(f + q*2^k/f)  (:424-477).
// The count of valid splits, its early exit at two and the unique split's
// terms do not depend on the order splits are visited in, so the 64 lanes
// of a wavefront visit 64 splits per round and count with ballots.
//
// Per-sum work differs by five orders of magnitude (most sums stop after a
// few rounds; emitted sums visit all ~s/4 splits), so waves pull sums from
// one atomic counter, largest sums first.  Splits that pass both PTP probes
// go to a per-wave LDS queue and are resolved 64 at a time, one per lane, so
// the rare factor-pair enumerations do not serialize the probe rounds.
// Terms d, o < s are factored with a smallest-prime-factor table, so no
// 64-bit trial division runs on the device.
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ffx_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kQueueCap = 128;
constexpr uint32_t kGrab = 4;     // sums taken per atomic
constexpr int kMaxPrimes = 12;    // distinct odd primes of q < 2^50
constexpr uint32_t kDBits = 27;   // queue entry: d | k << 27

struct SearchArgs {
    const uint8_t *map;
    uint64_t M;               // bitmap answers odd n <= M
    const uint16_t *spf;      // smallest prime factor of odd n (index n>>1), 0 if n prime or 1
    uint32_t spfLimit;        // spf covers odd n < spfLimit
    uint32_t first;           // first odd sum
    uint32_t nsums;           // odd sums first, first+2, ...
    uint32_t *slots;
    uint32_t *counter;
};

// ---------------------------------------------------------------- primality

// Montgomery product a*b/2^64 mod n for odd n < 2^63, a, b < n.
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, uint64_t n, uint64_t nneg)
{
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    const uint64_t m = lo * nneg;
    const uint64_t t = hi + __umul64hi(m, n) + (lo != 0);
    return t >= n ? t - n : t;
}

// Deterministic Miller-Rabin with ff_seg's bases {2,3,5,7,11,13,17} and
// base count by magnitude (AskMillerRabin, segmentedSieve.C:364-387).
// n odd, 17 < n < 2^63.
__device__ __noinline__ bool mr_is_prime(uint64_t n)
{
    uint64_t inv = n;  // n * n == 1 mod 8
#pragma unroll
    for (int i = 0; i < 5; ++i) inv *= 2 - n * inv;
    const uint64_t nneg = 0 - inv;
    const uint64_t one = (0 - n) % n;  // 2^64 mod n
    uint64_t r2 = one;
    for (int i = 0; i < 64; ++i) {
        r2 <<= 1;
        if (r2 >= n) r2 -= n;
    }
    const uint64_t mone = n - one;
    const uint32_t r = (uint32_t)__builtin_ctzll(n - 1);
    const uint64_t d = (n - 1) >> r;
    const int nb = n < 3215031751ull ? 4 : n < 2152302898747ull ? 5 : n < 3474749660383ull ? 6 : 7;
    const uint32_t bases[7] = {2, 3, 5, 7, 11, 13, 17};
    for (int bi = 0; bi < nb; ++bi) {
        uint64_t b = mont_mul(bases[bi], r2, n, nneg);
        uint64_t y = one;
        for (uint64_t e = d; e; e >>= 1) {
            if (e & 1) y = mont_mul(y, b, n, nneg);
            b = mont_mul(b, b, n, nneg);
        }
        if (y == one || y == mone) continue;
        bool pass = false;
        for (uint32_t j = 1; j < r; ++j) {
            y = mont_mul(y, y, n, nneg);
            if (y == mone) {
                pass = true;
                break;
            }
            if (y == one) return false;
        }
        if (!pass) return false;
    }
    return true;
}

__device__ __forceinline__ bool map_bit(const uint8_t *map, uint64_t n)
{
    return (map[n >> 4] >> (7 - ((n >> 1) & 7))) & 1;
}

// IsPrime for odd n >= 1 (segmentedSieve.C:283-296).
__device__ __forceinline__ bool is_prime_odd(const SearchArgs &a, uint64_t n)
{
    if (n <= a.M) return map_bit(a.map, n);
    return mr_is_prime(n);
}

// IsPrime for any n (tags of the even term).
__device__ __forceinline__ bool is_prime_any(const SearchArgs &a, uint64_t n)
{
    if (!(n & 1)) return n == 2;
    return is_prime_odd(a, n);
}

__device__ __forceinline__ uint64_t isqrt53(uint64_t v)
{
    uint64_t r = (uint64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// ---------------------------------------------------------------- AB1

// Adds the factorization of odd n (< spfLimit) to (pr, ex, nf), merging
// primes already present.
__device__ __forceinline__ void add_factors(const SearchArgs &a, uint32_t n, uint32_t *pr, uint32_t *ex, int &nf)
{
    while (n > 1) {
        uint32_t p = a.spf[n >> 1];
        if (p == 0) p = n;
        uint32_t e = 0;
        do {
            n /= p;
            ++e;
        } while (n % p == 0);
        bool merged = false;
#pragma unroll
        for (int i = 0; i < kMaxPrimes; ++i)
            if (i < nf && pr[i] == p) {
                ex[i] += e;
                merged = true;
            }
        if (!merged) {
#pragma unroll
            for (int i = 0; i < kMaxPrimes; ++i)
                if (i == nf) {
                    pr[i] = p;
                    ex[i] = e;
                }
            ++nf;
        }
    }
}

// AllButOneProductOfTermPairsOfSumOfFactorPairsHasSingleFactorPair(k, d*o)
// (segmentedSieve.C:424-477), with q = d*o factored from d and o.
__device__ __noinline__ bool ab1(const SearchArgs &a, uint32_t k, uint32_t d, uint32_t o)
{
    const uint64_t q = (uint64_t)d * o;
    const uint64_t P = q << k;
    const uint64_t L = isqrt53(P);
    const uint64_t Lk = L >> k;
    int bad = 0;
    if ((1ull << k) <= L && !is_prime_odd(a, (1ull << k) + q - 2)) ++bad;
    uint32_t pr[kMaxPrimes], ex[kMaxPrimes], e[kMaxPrimes];
    uint64_t pp[kMaxPrimes + 1];
    int nf = 0;
#pragma unroll
    for (int i = 0; i < kMaxPrimes; ++i) {
        pr[i] = 1;
        ex[i] = 0;
        e[i] = 0;
        pp[i] = 1;
    }
    pp[kMaxPrimes] = 1;
    add_factors(a, d, pr, ex, nf);
    add_factors(a, o, pr, ex, nf);
    const double qd = (double)q;
    // odometer over exponent vectors; pp[i] = prod_{j>=i} pr[j]^e[j]
    for (;;) {
        int i = 0;
        while (i < nf && e[i] == ex[i]) ++i;
        if (i == nf) break;
        uint64_t top = 1;
#pragma unroll
        for (int j = 0; j < kMaxPrimes; ++j)
            if (j == i) {
                e[j] += 1;
                pp[j] *= pr[j];
                top = pp[j];
            }
#pragma unroll
        for (int j = 0; j < kMaxPrimes; ++j)
            if (j < i) {
                e[j] = 0;
                pp[j] = top;
            }
        const uint64_t dv = top;
        const uint64_t c = (uint64_t)(qd / (double)dv);  // exact: dv | q, q < 2^53
        if (dv <= L && !is_prime_odd(a, dv + (c << k) - 2))
            if (++bad > 1) return false;
        if (dv <= Lk && !is_prime_odd(a, (dv << k) + c - 2))
            if (++bad > 1) return false;
    }
    return bad == 1;
}

// ---------------------------------------------------------------- per sum

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Resolves queue entries [from, from+nb) of this wave's queue, one per lane;
// updates the wave's valid count and, on the first valid split, its terms.
__device__ __forceinline__ void drain(const SearchArgs &a, const uint2 *queue, int from, int nb, int lane, int &cnt,
                                      uint32_t &te, uint32_t &to)
{
    bool valid = false;
    uint32_t ee = 0, oo = 0;
    if (lane < nb) {
        const uint2 ent = queue[from + lane];
        const uint32_t d = ent.x & ((1u << kDBits) - 1), k = ent.x >> kDBits;
        oo = ent.y;
        ee = d << k;
        valid = ab1(a, k, d, oo);
    }
    const uint64_t vm = __ballot(valid);
    if (vm) {
        const int l = __ffsll((unsigned long long)vm) - 1;
        const uint32_t e1 = __shfl(ee, l), o1 = __shfl(oo, l);
        if (cnt == 0) {
            te = e1;
            to = o1;
        }
        cnt += __popcll(vm);
    }
}

__device__ uint32_t solve_sum_wave(const SearchArgs &a, uint32_t s, int lane, uint2 *queue)
{
    // statement 1 (RunIt :806-807)
    if (is_prime_odd(a, s - 2)) return 0;
    if (s % 3 == 0 && !is_prime_odd(a, s / 3)) return 0;

    int cnt = 0, qn = 0;
    uint32_t te = 0, to = 0;
    const uint64_t lt = lanes_below(lane);

    // Power2Prime (:547-564): lane l tests k = l + 2
    {
        const uint32_t k = 2 + lane;
        const bool act = k < 32 && (1u << k) < s - 2;
        const bool pr = act && is_prime_odd(a, s - (1u << k));
        const uint64_t pm = __ballot(pr), cm = __ballot(act && !pr);
        cnt = __popcll(pm);
        if (cnt > 1) return 0;
        if (cnt == 1) {
            te = 1u << (2 + __ffsll((unsigned long long)pm) - 1);
            to = s - te;
        }
        // composite odd terms feed Power2Composite (:591-605) through the queue
        if (act && !pr) queue[__popcll(cm & lt)] = make_uint2(1u | (k << kDBits), s - (1u << k));
        qn = __popcll(cm);
    }
    wave_sync_lds();

    // Power2Odd / Power2Even (:566-589, :607-630), k ascending
    for (uint32_t k = 2; k < 32; ++k) {
        const uint32_t d0 = (k & 1) ? 5 : 3;
        if (((uint64_t)d0 << k) > (uint64_t)(s - 3)) break;
        const uint32_t D = (s - 3) >> k;  // d*2^k < s-2  <=>  d <= (s-3)>>k
        for (uint32_t db = d0; db <= D; db += 128) {
            const uint32_t d = db + 2 * lane;
            bool cand = false;
            uint32_t o = 0;
            if (d <= D) {
                o = s - (d << k);
                const bool m3 = (d % 3 == 0) || (o % 3 == 0);
                const bool pass = (k & 1) ? !m3 : (m3 || is_prime_odd(a, o));
                if (pass) {
                    const uint64_t q = (uint64_t)d * o;
                    if (is_prime_odd(a, (1ull << k) + q - 2))
                        cand = d == o || is_prime_odd(a, (uint64_t)d + ((uint64_t)o << k) - 2);
                }
            }
            const uint64_t cm = __ballot(cand);
            if (cm) {
                if (cand) queue[qn + __popcll(cm & lt)] = make_uint2(d | (k << kDBits), o);
                qn += __popcll(cm);
                wave_sync_lds();
                if (qn >= 64) {
                    drain(a, queue, qn - 64, 64, lane, cnt, te, to);
                    qn -= 64;
                    if (cnt > 1) return 0;
                }
            }
        }
    }
    while (qn > 0) {
        const int nb = qn < 64 ? qn : 64;
        drain(a, queue, qn - nb, nb, lane, cnt, te, to);
        qn -= nb;
        if (cnt > 1) return 0;
    }
    if (cnt != 1) return 0;

    const uint32_t lo = te < to ? te : to, hi = te < to ? to : te;
    uint32_t slot = lo;
    // tags, prime first as OutputTags does (:663-677)
    if (is_prime_any(a, lo)) slot |= FFX_TAG_LOW_PRIME;
    else if ((lo & (lo - 1)) == 0) slot |= FFX_TAG_LOW_POW2;
    if (is_prime_any(a, hi)) slot |= FFX_TAG_HIGH_PRIME;
    else if ((hi & (hi - 1)) == 0) slot |= FFX_TAG_HIGH_POW2;
    return slot;
}

__global__ void zero_u32(uint32_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kBlock) void search_sums(SearchArgs a)
{
    __shared__ uint2 queues[kWavesPerBlock][kQueueCap];
    const int lane = threadIdx.x & 63;
    uint2 *queue = queues[threadIdx.x >> 6];
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.counter, kGrab);
        base = __shfl(base, 0);
        if (base >= a.nsums) break;
        const uint32_t end = base + kGrab < a.nsums ? base + kGrab : a.nsums;
        for (uint32_t idx = base; idx < end; ++idx) {
            const uint32_t slotIdx = a.nsums - 1 - idx;  // largest sums first
            const uint32_t s = a.first + 2 * slotIdx;
            const uint32_t res = solve_sum_wave(a, s, lane, queue);
            if (lane == 0) a.slots[slotIdx] = res;
        }
    }
}

// ---------------------------------------------------------------- compaction

constexpr int kCompactBlock = 256;
constexpr int kPerThread = 16;
constexpr uint32_t kTile = kCompactBlock * kPerThread;

__global__ __launch_bounds__(kCompactBlock) void tile_count(const uint32_t *slots, uint64_t n, uint32_t *tileCounts)
{
    __shared__ uint32_t part[kCompactBlock / 64];
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPerThread;
    uint32_t c = 0;
    for (int i = 0; i < kPerThread; ++i)
        if (t0 + i < n && slots[t0 + i]) ++c;
    for (int off = 32; off; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kCompactBlock / 64; ++w) s += part[w];
        tileCounts[blockIdx.x] = s;
    }
}

// Exclusive scan of tile counts in one block; total to *count.
__global__ __launch_bounds__(1024) void tile_scan(uint32_t *tileCounts, uint32_t ntiles, uint32_t *count)
{
    __shared__ uint32_t sums[1024];
    const uint32_t per = (ntiles + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b + i < ntiles) local += tileCounts[b + i];
    sums[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (uint32_t)off ? sums[threadIdx.x - off] : 0;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = sums[threadIdx.x] - local;
    for (uint32_t i = 0; i < per; ++i)
        if (b + i < ntiles) {
            const uint32_t c = tileCounts[b + i];
            tileCounts[b + i] = run;
            run += c;
        }
    if (threadIdx.x == 1023) *count = sums[1023];
}

__global__ __launch_bounds__(kCompactBlock) void tile_scatter(const uint32_t *slots, uint64_t n, uint32_t firstSum,
                                                              const uint32_t *tileOffsets, ffx_solution *out)
{
    __shared__ uint32_t part[kCompactBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPerThread;
    uint32_t c = 0;
    for (int i = 0; i < kPerThread; ++i)
        if (t0 + i < n && slots[t0 + i]) ++c;
    // inclusive scan across the wave
    uint32_t inc = c;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
    }
    if (lane == 63) part[w] = inc;
    __syncthreads();
    uint32_t waveBase = 0;
    for (int j = 0; j < w; ++j) waveBase += part[j];
    uint32_t pos = tileOffsets[blockIdx.x] + waveBase + inc - c;
    for (int i = 0; i < kPerThread; ++i) {
        const uint64_t idx = t0 + i;
        if (idx >= n) break;
        const uint32_t v = slots[idx];
        if (!v) continue;
        ffx_solution r;
        r.sum = firstSum + 2u * (uint32_t)idx;
        r.low = v & FFX_SLOT_LOW_MASK;
        r.high = r.sum - r.low;
        r.tags = v & ~FFX_SLOT_LOW_MASK;
        out[pos++] = r;
    }
}

__global__ void slots_to_records(const uint32_t *slots, uint64_t n, uint32_t firstSum, ffx_gpu_record *rec,
                                 uint32_t *count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = slots[i];
    if (!v) return;
    ffx_gpu_record r;
    memset(&r, 0, sizeof r);
    r.sum = firstSum + 2u * (uint32_t)i;
    r.low = v & FFX_SLOT_LOW_MASK;
    r.high = r.sum - r.low;
    rec[i] = r;
    atomicAdd(count, 1u);
}

}  // namespace

struct ffx_search_plan {
    uint32_t sumLimitMax;
    uint16_t *dSpf;
    uint32_t spfLimit;
    uint32_t *dCounter;
    uint32_t *dTile;      // tile counts / offsets
    uint32_t tileCap;
    uint32_t *dSlots;     // scratch slots for ffx_search_async
    uint64_t slotCap;
    int grid;
    int device;
};

using namespace ffx;

static int ensure_scratch(ffx_search_plan *pl, uint64_t nslots, bool slots)
{
    const uint64_t ntiles = (nslots + kTile - 1) / kTile;
    if (ntiles > pl->tileCap) {
        (void)hipFree(pl->dTile);
        pl->dTile = nullptr;
        pl->tileCap = 0;
        FFX_HIP(hipMalloc(&pl->dTile, (ntiles + 1) * 4));
        pl->tileCap = (uint32_t)ntiles;
    }
    if (slots && nslots > pl->slotCap) {
        (void)hipFree(pl->dSlots);
        pl->dSlots = nullptr;
        pl->slotCap = 0;
        FFX_HIP(hipMalloc(&pl->dSlots, nslots * 4));
        pl->slotCap = nslots;
    }
    return FFX_OK;
}

extern "C" {

int ffx_search_plan_create(ffx_search_plan **out, uint32_t sumLimitMax)
{
    if (!out) return fail(FFX_EINVAL, "search plan: null out");
    *out = nullptr;
    if (sumLimitMax > FFX_MAX_SUM) return fail(FFX_ERANGE, "search plan: sumLimitMax %u > %u", sumLimitMax, FFX_MAX_SUM);
    if (sumLimitMax < 5) sumLimitMax = 5;
    // smallest prime factor of odd n < spfLimit (0 when n is 1 or prime)
    const uint32_t spfLimit = sumLimitMax + 2;
    std::vector<uint16_t> spf((spfLimit >> 1) + 1, 0);
    for (uint64_t p = 3; p * p < spfLimit; p += 2) {
        if (spf[p >> 1]) continue;
        for (uint64_t m = p * p; m < spfLimit; m += 2 * p)
            if (!spf[m >> 1]) spf[m >> 1] = (uint16_t)p;
    }
    ffx_search_plan *pl = new (std::nothrow) ffx_search_plan();
    if (!pl) return fail(FFX_ENOMEM, "search plan: host alloc");
    pl->sumLimitMax = sumLimitMax;
    pl->spfLimit = spfLimit;
    hipError_t e = hipGetDevice(&pl->device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
    if (e == hipSuccess) e = hipMalloc(&pl->dSpf, spf.size() * 2);
    if (e == hipSuccess) e = hipMemcpy(pl->dSpf, spf.data(), spf.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&pl->dCounter, 64);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "search plan upload");
        ffx_search_plan_destroy(pl);
        return rc;
    }
    pl->grid = (cus > 0 ? cus : 256) * 8;  // 32 waves per CU
    *out = pl;
    return FFX_OK;
}

int ffx_search_plan_destroy(ffx_search_plan *plan)
{
    if (!plan) return FFX_OK;
    (void)hipFree(plan->dSpf);
    (void)hipFree(plan->dCounter);
    (void)hipFree(plan->dTile);
    (void)hipFree(plan->dSlots);
    delete plan;
    return FFX_OK;
}

int ffx_search_slots_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                           uint32_t sumLimit, uint32_t *d_slots, void *stream)
{
    if (!plan || !d_map || !d_slots) return fail(FFX_EINVAL, "search: null argument");
    const uint32_t first = (sumStart | 1u) < 5 ? 5 : (sumStart | 1u);
    if (sumLimit < first) return FFX_OK;
    if (sumLimit > plan->sumLimitMax)
        return fail(FFX_ERANGE, "search: sumLimit %u above the plan's %u", sumLimit, plan->sumLimitMax);
    if (maxPrimeMapValue < 15) return fail(FFX_EINVAL, "search: map too small");
    SearchArgs a;
    a.map = static_cast<const uint8_t *>(d_map);
    a.M = maxPrimeMapValue;
    a.spf = plan->dSpf;
    a.spfLimit = plan->spfLimit;
    a.first = first;
    a.nsums = (sumLimit - first) / 2 + 1;
    a.slots = d_slots;
    a.counter = plan->dCounter;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, st, plan->dCounter, 16u);
    FFX_HIP(hipGetLastError());
    const uint32_t need = (a.nsums + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t grid = need < (uint32_t)plan->grid ? need : (uint32_t)plan->grid;
    hipLaunchKernelGGL(search_sums, dim3(grid), dim3(kBlock), 0, st, a);
    FFX_HIP(hipGetLastError());
    return FFX_OK;
}

int ffx_compact_async(ffx_search_plan *plan, const uint32_t *d_slots, uint64_t nslots, uint32_t firstSum,
                      ffx_solution *d_out, uint32_t *d_count, void *stream)
{
    if (!plan || !d_slots || !d_out || !d_count) return fail(FFX_EINVAL, "compact: null argument");
    hipStream_t st = as_stream(stream);
    if (nslots == 0) {
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, st, d_count, 1u);
        FFX_HIP(hipGetLastError());
        return FFX_OK;
    }
    int rc = ensure_scratch(plan, nslots, false);
    if (rc) return rc;
    const uint32_t ntiles = (uint32_t)((nslots + kTile - 1) / kTile);
    hipLaunchKernelGGL(tile_count, dim3(ntiles), dim3(kCompactBlock), 0, st, d_slots, nslots, plan->dTile);
    hipLaunchKernelGGL(tile_scan, dim3(1), dim3(1024), 0, st, plan->dTile, ntiles, d_count);
    hipLaunchKernelGGL(tile_scatter, dim3(ntiles), dim3(kCompactBlock), 0, st, d_slots, nslots, firstSum, plan->dTile,
                       d_out);
    FFX_HIP(hipGetLastError());
    return FFX_OK;
}

int ffx_search_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                     uint32_t sumLimit, ffx_solution *d_out, uint32_t *d_count, void *stream)
{
    if (!plan) return fail(FFX_EINVAL, "search: null plan");
    const uint32_t first = (sumStart | 1u) < 5 ? 5 : (sumStart | 1u);
    const uint64_t nslots = sumLimit >= first ? (sumLimit - first) / 2 + 1 : 0;
    int rc = ensure_scratch(plan, nslots ? nslots : 1, true);
    if (rc) return rc;
    rc = ffx_search_slots_async(plan, d_map, maxPrimeMapValue, sumStart, sumLimit, plan->dSlots, stream);
    if (rc) return rc;
    return ffx_compact_async(plan, plan->dSlots, nslots, first, d_out, d_count, stream);
}

int ffx_search_records(int device, const uint8_t *d_map, uint64_t maxPrimeMapValue, uint64_t sumStart,
                       uint64_t sumLimit, uint32_t *d_count, ffx_gpu_record *d_records)
{
    static std::mutex mu;
    static ffx_search_plan *cached = nullptr;
    static int cachedDev = -1;
    std::lock_guard<std::mutex> lock(mu);
    if (sumLimit > FFX_MAX_SUM) return fail(FFX_ERANGE, "search records: sumLimit too large");
    int rc = ffx_init(device);
    if (rc) return rc;
    if (!cached || cachedDev != device || cached->sumLimitMax < sumLimit) {
        ffx_search_plan_destroy(cached);
        cached = nullptr;
        rc = ffx_search_plan_create(&cached, (uint32_t)sumLimit);
        if (rc) return rc;
        cachedDev = device;
    }
    const uint32_t first = (uint32_t)sumStart;
    if ((first & 1) == 0 || first < 5) return fail(FFX_EINVAL, "search records: sumStart must be odd and >= 5");
    if (sumLimit < first) return FFX_OK;
    const uint64_t n = (sumLimit - first) / 2 + 1;
    rc = ensure_scratch(cached, n, true);
    if (rc) return rc;
    rc = ffx_search_slots_async(cached, d_map, maxPrimeMapValue, first, (uint32_t)sumLimit, cached->dSlots, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(slots_to_records, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, nullptr, cached->dSlots, n,
                       first, d_records, d_count);
    FFX_HIP(hipGetLastError());
    FFX_HIP(hipDeviceSynchronize());
    return FFX_OK;
}

}  // extern "C"
