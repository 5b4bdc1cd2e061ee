// libffx sieve: the odd-only prime bitmap, written by LDS-resident segments.
//
// Replaces SieveSlab_<arch> (source/sieve_slab_kernel.h:40-107), its launch
// and prime cache (source/sieve_slab_engine.cpp:115-202, :279-312, :409-442)
// and the host small-prime generation of SieveEngine
// (source/sieve_engine.cpp:31-47, :114-133).  The bitmap it writes is the
// one Prime::Prime + SegmentFill build (reference/segmentedSieve.C:177-222,
// :255-270): odd n in byte n>>4 under mask 0x80>>((n>>1)&7), 1 = prime for
// odd n <= M, untouched (1) for odd n > M, value 1 cleared.
//
// Kernel design (gfx950, wave64, 160 KiB LDS per CU):
//  * One workgroup of 1024 threads owns one map segment at a time in LDS
//    (32 KiB = 2^18 odd values below maps of 2^32 values, 112 KiB from there)
//    and sieves a contiguous run of segments.
//  * Bits are kept "struck" = 1 in LDS, in natural order (bit b of word w is
//    odd index 32w+b); the write-out inverts them and reorders each byte to
//    the reference's MSB-first layout (bit position b^7 within the word).
//  * Primes 3..61 never strike: their multiples are OR-ed in as periodic word
//    patterns (eight groups, periods 105..3599 words) kept in LDS.
//  * Primes 67..segOdds/192 (more than 192 strikes per segment) are struck
//    cooperatively: two primes per wavefront, lane l of a half striking
//    every 32nd multiple (a fixed bit, an advancing word).  Larger primes
//    are struck one prime per lane, carrying their next strike across the
//    workgroup's run of segments.  Both skip the odd multiples of 3 (a
//    mod-6 wheel over the multiplier).  Strikes are LDS atomic ORs
//    (ds_or_b32).
//  * The segment leaves LDS through 16-byte coalesced stores.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ffx_internal.h"

namespace {

// Defaults measured on MI355X (ff_sieve --autotune and direct timing):
// maps whose sieving primes stay below 2^16 (M < 2^32) are fastest with
// 32 KiB segments and four 1024-thread workgroups per CU (two resident with
// the 43 KiB of patterns): 64 MiB map 0.121 ms against 0.148 with 112 KiB,
// but 256 MiB 0.404 against 0.348, 1 GiB 1.44 against 1.25 (since the
// cooperative primes are taken from a counter; before, 1 GiB took 1.64
// with 32 KiB against 1.70 and the switch was at 2^35 values).
// Larger maps, where most primes strike a segment only a few times and
// each segment pays every prime's set-up, are faster with the largest
// segment the LDS holds beside the patterns, one workgroup resident per
// CU: 2 GiB map 3.32 ms against 3.41 with 32 KiB, 4 GiB 6.68 against 7.31,
// 16 GiB 39.5 ms against 40.1 (96 KiB), 41.2 (80 KiB) and 41.8 (64 KiB, two
// per CU) when the kernel had no wheels.  The grid holds two workgroups per
// CU (one: 16 GiB 28.49 ms, two: 28.30, three: 28.33, four: 28.31).
constexpr uint32_t kSmallMapSegBytes = 32768, kSmallMapBlocksPerCU = 4;
constexpr uint32_t kLargeMapSegBytes = 114688, kLargeMapBlocksPerCU = 2;
constexpr uint64_t kLargeMapFrom = 1ull << 32;
constexpr int kThreads = 1024;
// Cooperative primes up to segOdds / kWheelDiv strike through a mod-6 wheel
// (a third fewer LDS atomics; above it a lane has too few strikes per
// progression to pay for the second pass).  16 GiB map: 39.3 ms without the
// wheel, 36.5 / 35.6 / 35.4 / 35.6 / 36.0 ms with divisors 96 / 192 / 384 /
// 768 / 1536; 70 GB map 198 -> 177 ms.
#ifndef FFX_SIEVE_WHEEL_DIV
#define FFX_SIEVE_WHEEL_DIV 384
#endif
constexpr uint32_t kWheelDiv = FFX_SIEVE_WHEEL_DIV;
// Those up to segOdds / kWheel30Div also skip their multiples of 5 (a fifth
// fewer strikes again); above it a lane has too few strikes per segment to
// pay for the wheel's set-up and partial last step.
#ifndef FFX_SIEVE_WHEEL30_DIV
#define FFX_SIEVE_WHEEL30_DIV 1536
#endif
constexpr uint32_t kWheel30Div = FFX_SIEVE_WHEEL30_DIV;

constexpr int kGroups = 8;
// Small-prime groups whose word patterns replace strikes.  Each period (the
// product of the group's primes) is coprime to 32, so the pattern of odd
// indices repeats every `period` 32-bit words.
constexpr uint32_t kGroupPrimes[kGroups][3] = {{3, 5, 7}, {11, 13, 0}, {17, 19, 0}, {23, 29, 0},
                                               {31, 37, 0}, {41, 43, 0}, {47, 53, 0}, {59, 61, 0}};
constexpr uint32_t kLargestPatternPrime = 61;

constexpr uint32_t group_period(int g)
{
    return kGroupPrimes[g][0] * kGroupPrimes[g][1] * (kGroupPrimes[g][2] ? kGroupPrimes[g][2] : 1);
}
// A thread fills up to kFillWords words of a segment at once, 32 words
// apart (lane i of a half-wave on words i, i+32, i+64, ... of the
// half-wave's block, so each read and write instruction has 32 lanes on 32
// consecutive words: no LDS bank conflicts), with one pattern cursor per
// group.  Each group's pattern is stored with its first kPatWrap words
// repeated after its period, so those words never wrap inside one read.
constexpr uint32_t kFillWords = 4;
constexpr uint32_t kPatWrap = 32 * (kFillWords - 1);
constexpr uint32_t pattern_base(int g)
{
    uint32_t s = 0;
    for (int i = 0; i < g; ++i) s += group_period(i) + kPatWrap;
    return s;
}
constexpr uint32_t kPatWords = pattern_base(kGroups);  // 11006 words, 43 KiB

// Primes below segOdds / coop_div(segOdds) (more than coop_div strikes per
// segment) strike cooperatively, the rest one per lane through the mod-30
// wheel.  Since the cooperative primes are
// taken from a counter after each wave's lane primes, the boundary sits
// higher: 112 KiB segments, divisor 192 / 256 / 320 / 384 / 512 / 640 /
// 768 / 896 / 1024: 16 GiB map 24.6 / 23.5 / 22.9 / 22.4 / 21.9 / 21.4 /
// 21.4 / 22.9 / 24.5 ms (4 GiB 5.69 ... 4.88 at 768 ... 5.72 at 1024, 70
// GB 128 ... 113 ... 115); 32 KiB segments, whose threads carry only 12
// lane primes: 1 GiB map 1.54 / 1.49 / 1.48 / 1.50 / 1.66 ms at 192 /
// 256 / 320 / 384 / 512.  (Before the counter, 192 was best: 16 GiB map
// 30.9 ms against 31.2 / 31.8 / 33.2 at 256 / 384 / 512.)
constexpr uint32_t coop_div(uint32_t segOdds) { return segOdds == 32768u * 8 ? 320u : 768u; }

// Odd primes in (kLargestPatternPrime, x): the cooperative primes of a
// segment of coop_div * x odd values, counted at compile time to size their
// table.
constexpr uint32_t coop_primes_below(uint32_t x)
{
    uint32_t n = 0;
    for (uint32_t c = kLargestPatternPrime + 2; c < x; c += 2) {
        bool prime = true;
        for (uint32_t d = 3; d * d <= c; d += 2)
            if (c % d == 0) {
                prime = false;
                break;
            }
        n += prime;
    }
    return n;
}
// table entries for a segment of segOdds odd values, rounded up to 64
constexpr uint32_t coop_capacity(uint32_t segOdds) { return (coop_primes_below(segOdds / coop_div(segOdds)) + 63) & ~63u; }

struct SieveArgs {
    uint8_t *dst;          // receives map bytes [byteLo, byteHi) at dst[byte - dstBase]
    uint64_t dstBase;
    uint64_t byteLo, byteHi;
    uint64_t oddLimit;     // odd indices j >= oddLimit (values > M) are never struck
    const uint32_t *primes;  // sieving primes > 61, ascending
    const double *invp;      // 1.0 / primes[i]
    uint32_t nprimes;
    uint32_t ncoop;          // primes[0, ncoop) strike wave-cooperatively (> coop_div strikes per segment)
    uint32_t nwheel;         // primes[0, nwheel) of those skip their multiples of 3 (mod-6 wheel)
    uint32_t nwheel30;       // primes[0, nwheel30) of those also skip their multiples of 5 (mod-30 wheel)
    uint32_t ncarry;         // lane primes carried per thread (<= the kernel's kCarry)
    const uint32_t *patterns;  // kPatWords words: each group's period, then its first kPatWrap words again
    uint64_t segFirst, segCount;
};

// First odd index >= j0 struck by prime p (its odd multiples from p*p on,
// i.e. odd indices congruent to (p-1)/2 mod p, from (p*p-1)/2), relative to
// j0.  j0 < 2^52 so the double quotient is off by at most one.
__device__ __forceinline__ uint64_t first_strike(uint64_t j0, uint32_t p, double invp)
{
    int64_t q = (int64_t)((double)j0 * invp);
    int64_t m = (int64_t)j0 - q * (int64_t)p;
    if (m < 0) m += p;
    if (m >= (int64_t)p) m -= p;
    uint32_t r = (p - 1) >> 1;
    uint64_t off = (uint32_t)m <= r ? r - (uint32_t)m : r + p - (uint32_t)m;
    uint64_t sq = ((uint64_t)p * p - 1) >> 1;
    if (j0 + off < sq) off = sq - j0;
    return off;
}

// The same first strike in double arithmetic, clamped: the offset when it
// is below limit, else some value >= limit.  Exact because every quantity is
// an integer below 2^53: x = (p-1)/2 - j0, x - t*p from one fma, and
// (p*p-1)/2 - j0; the quotient estimate t is off by at most one, so the
// remainder lands in (-p, 2p) and two unsigned minima bring it into [0, p).
__device__ __forceinline__ uint32_t first_strike_below(double j0d, uint32_t p, double invp, uint32_t limit)
{
    const double pd = (double)p;
    const double x = (double)((p - 1) >> 1) - j0d;
    const double t = floor(x * invp);
    uint32_t u = (uint32_t)(int32_t)fma(-t, pd, x);
    u = min(u, u + p);
    u = min(u, u - p);
    // not before p*p (odd index (p*p-1)/2, congruent to u mod p)
    const double below = fma(pd, 0.5 * pd, -0.5) - j0d;
    return below > (double)u ? (uint32_t)fmin(below, (double)limit) : u;
}

// Reorders a word of natural-order bits (bit b = odd index b) into the
// reference byte layout: within each little-endian byte the smallest odd
// value is the most significant bit, i.e. bit b moves to b ^ 7.
__device__ __forceinline__ uint32_t to_map_order(uint32_t v)
{
    return __builtin_bswap32(__builtin_bitreverse32(v));
}

// Lane primes carried across a run: thread t keeps, for k < kCarry, the
// offset of the next strike of primes[ncoop + t + k*kThreads] in a register,
// so only the run's first segment divides; the primes beyond them start
// each segment with first_strike.  kHoldPrimes keeps the primes themselves
// in registers too; otherwise they are re-read each segment (L2-resident),
// which halves the registers per carried prime.  32 KiB segments run eight
// waves per SIMD (64 registers per lane), the large segments four (128).
// Measured on MI355X with 64 KiB large segments (4 GiB / 16 GiB / 70 GB
// maps): 32 KiB with 12 re-read carries 9.45 ms (8: 9.7, none: 10.1);
// 64 KiB with 32 held 43.4 ms on
// 16 GiB (64 re-read: 44.9) but 214 ms on 70 GB, where 64 re-read take 202.
// Carried lane primes strike through a mod-30 wheel over the multiplier m:
// the odd multiples with 3 | m or 5 | m are the patterns' (3 and 5 are
// pattern primes), so only m coprime to 30 are struck -- m = 1, 7, 11, 13,
// 17, 19, 23, 29 (mod 30), odd-index steps 3p, 2p, p, 2p, p, 2p, 3p, p from
// one to the next (nibble i of kWheel30Steps for position i).  A carried
// strike is its offset (bits 0..28) and the wheel position of its m (bits
// 29..31); the launch keeps a run at most kRunOddsLimit odd values, so the
// offsets fit.
constexpr uint32_t kWheel30Steps = 0x13212123u;
constexpr bool wheel30_tables_agree()
{
    constexpr uint32_t m[9] = {1, 7, 11, 13, 17, 19, 23, 29, 31};
    constexpr uint64_t pos = 0x0788685483281880ull;
    for (uint32_t i = 0; i < 8; ++i)
        if (((kWheel30Steps >> (4 * i)) & 15u) != (m[i + 1] - m[i]) / 2) return false;
    for (uint32_t r = 1; r < 30; r += 2) {
        uint32_t want = 8;
        for (uint32_t i = 0; i < 8; ++i)
            if (m[i] == r) want = i;
        if (((pos >> (4 * (r >> 1))) & 15u) != want) return false;
    }
    return true;
}
static_assert(wheel30_tables_agree(), "mod-30 wheel steps and positions");
constexpr int kWheelPosShift = 29;
constexpr uint32_t kOffsetMask = (1u << kWheelPosShift) - 1;
constexpr uint64_t kRunOddsLimit = 1ull << kWheelPosShift;
// the wheel position of m mod 30, or 8 when 3 | m or 5 | m (odd m)
__device__ __forceinline__ uint32_t wheel30_pos(uint32_t r30)
{
    // nibble r30 / 2 (odd r30 in 1..29)
    constexpr uint64_t kPos = 0x0788685483281880ull;
    return (uint32_t)(kPos >> (4 * (r30 >> 1))) & 15u;
}
constexpr int kCarrySmall = 12;      // 32 KiB segments, primes re-read
constexpr int kCarryHeld = 32;       // large segments, primes held
constexpr int kCarryReread = 64;     // large segments, primes re-read
// The large kernels hold their carried primes by default: 32 held against
// 64 re-read, 4 GiB map 5.85 / 6.66 ms, 16 GiB 25.9 / 27.9, 70 GB 133.8 /
// 136.8, 256 GiB 764.7 / 770.0 (more primes carried does not make up for the
// registers the re-read kernel spends on them).

template <uint32_t kSegBytes, int kCarry, bool kHoldPrimes>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kSegBytes == 32768 ? 8 : 4))) void sieve_segments(SieveArgs a)
{
    constexpr uint32_t kSegWords = kSegBytes / 4;
    constexpr uint32_t kSegOdds = kSegBytes * 8;  // odd values per segment
    constexpr uint32_t kCoopMax = coop_capacity(kSegOdds);  // the odd primes 67..kSegOdds/64
    // words per fill chunk: four in the 112 KiB kernels (two ds_read2 per
    // pattern group), two in the 32 KiB kernel, whose 64 registers hold no
    // more words in flight beside its carried strikes
    constexpr uint32_t kFill = kSegBytes == 32768 ? 2 : kFillWords;
    // the first word of this thread's chunk in a fill step of
    // kFill * kThreads words: the wave's block, its half, its lane
    const uint32_t fillWord0 = kFill * 64u * ((uint32_t)threadIdx.x >> 6) + kFill * 32u * (((uint32_t)threadIdx.x >> 5) & 1u) +
                               ((uint32_t)threadIdx.x & 31u);
    // One LDS block: the patterns first (their group offsets and the
    // segment's own offset then fit the 16-bit offset field of the DS
    // instructions, so neither costs an address add), then the segment
    // (16-byte aligned for the four-word reads and writes), then the
    // cooperative primes' next strikes.
    constexpr uint32_t kSegAt = (kPatWords + 3) & ~3u;
    __shared__ __attribute__((aligned(16))) uint32_t lds[kSegAt + kSegWords + kCoopMax];
    __shared__ uint32_t coopTaken;  // cooperative primes taken in this segment
    uint32_t *const pat = lds;
    uint32_t *const seg = lds + kSegAt;
    uint32_t *const coopNext = lds + kSegAt + kSegWords;  // first strike of each cooperative prime, from the segment's start
    const int tid = threadIdx.x;
    const int lane = tid & 63;

    for (uint32_t i = tid; i < kPatWords; i += kThreads) pat[i] = a.patterns[i];

    // Per-thread pattern cursors: a thread fills the kFill words
    // fillWord0 + 32k (k < kFill) of each step of kFill * kThreads words;
    // its cursor into group g is the byte offset of its first word's place
    // in the period, advanced by the step per chunk and by kSegWords per
    // segment, and wrapped with an unsigned minimum (no division inside the
    // loops).
    uint32_t period[kGroups], stepFill[kGroups], stepSeg[kGroups], cur[kGroups];
    // Each workgroup sieves a contiguous run of segments, so a cooperative
    // prime's first strike in the next segment follows from this one's
    // (kept in LDS) instead of a 64-bit division per segment.
    // (uniform: moved to scalar registers, so the 64-bit division's result
    // does not hold vector registers through the run)
    uint64_t perBlock = (a.segCount + gridDim.x - 1) / gridDim.x;
    perBlock = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(perBlock >> 32)) << 32 |
               __builtin_amdgcn_readfirstlane((uint32_t)perBlock);
    const uint64_t runLo = (uint64_t)blockIdx.x * perBlock;
    const uint64_t runHi = runLo + perBlock < a.segCount ? runLo + perBlock : a.segCount;
    {
        const uint64_t w0 = (a.segFirst + runLo) * (uint64_t)kSegWords + fillWord0;
        const uint64_t segStride = kSegWords;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            period[g] = group_period(g);
            cur[g] = 4u * (uint32_t)(w0 % period[g]);
            stepFill[g] = (kFill * kThreads) % period[g];
            stepSeg[g] = (uint32_t)(segStride % period[g]);
        }
    }
    // The pattern words of one fill chunk (words 32k from the cursors),
    // OR-ed group by group into dst[32k]; each cursor then moves on by one
    // chunk step.
    auto fill_chunk = [&](uint32_t (&c)[kGroups], uint32_t *dst) {
        uint32_t v[kFill] = {};
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint32_t *q = reinterpret_cast<const uint32_t *>(
                reinterpret_cast<const char *>(pat + pattern_base(g)) + c[g]);
#pragma unroll
            for (uint32_t k = 0; k < kFill; ++k) v[k] |= q[32 * k];
            const uint32_t n = c[g] + 4u * stepFill[g];
            c[g] = min(n, n - 4u * period[g]);
        }
#pragma unroll
        for (uint32_t k = 0; k < kFill; ++k) dst[32 * k] = v[k];
    };
    static_assert(kFill <= kFillWords, "the pattern wrap covers kFillWords words");
    constexpr uint32_t kChunks = kSegWords / 4;  // write-out chunks: four words, 16 map bytes
    constexpr bool kFastWriteOut = kSegBytes == 32768 || kHoldPrimes;
    static_assert(kChunks % kThreads == 0 && kSegWords % (kFill * kThreads) == 0,
                  "every thread fills and writes the same number of chunks");
    // carried lane primes: next strike relative to the current segment's
    // start, at most the run's length (at most kRunOddsLimit odd values), so
    // a prime whose p*p lies beyond the run never strikes in it
    uint32_t cn[kCarry], cp[kHoldPrimes ? kCarry : 1];
    if (runLo < runHi) {
        const uint64_t j0 = (a.segFirst + runLo) * (uint64_t)kSegOdds;
        const uint64_t runOdds = (runHi - runLo) * (uint64_t)kSegOdds;
#pragma unroll
        for (int k = 0; k < kCarry; ++k) {
            const uint32_t pi = a.ncoop + (uint32_t)tid + (uint32_t)k * kThreads;
            cn[k] = 0;
            if (kHoldPrimes) cp[kHoldPrimes ? k : 0] = 0;
            if ((uint32_t)k < a.ncarry && pi < a.nprimes) {
                const uint32_t p = a.primes[pi];
                uint64_t f = first_strike(j0, p, a.invp[pi]);
                // mod-30 wheel: its multiple m (exact, as for the cooperative
                // primes), moved on by 2 (p odd indices) at most twice to
                // the next m prime to 30
                const uint64_t m = (uint64_t)rint((2.0 * (double)(j0 + f) + 1.0) * a.invp[pi]);
                uint32_t r30 = (uint32_t)(m % 30), pos = wheel30_pos(r30);
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    if (pos == 8) {
                        f += p;
                        r30 = r30 + 2 == 31 ? 1u : r30 + 2;
                        pos = wheel30_pos(r30);
                    }
                cn[k] = (uint32_t)(f < runOdds ? f : runOdds) | pos << kWheelPosShift;
                if (kHoldPrimes) cp[kHoldPrimes ? k : 0] = p;
            }
        }
    }
    __syncthreads();

    for (uint64_t si = runLo; si < runHi; ++si) {
        const uint64_t segIdx = a.segFirst + si;
        const uint64_t j0 = segIdx * (uint64_t)kSegOdds;
        const uint64_t segByte0 = segIdx * (uint64_t)kSegBytes;

        if (tid == 0) coopTaken = 0;
        // 1. small-prime patterns, one chunk of four words per thread and step
        {
            uint32_t c[kGroups];
#pragma unroll
            for (int g = 0; g < kGroups; ++g) c[g] = cur[g];
#pragma unroll 1
            for (uint32_t w = fillWord0; w < kSegWords; w += kFill * kThreads) fill_chunk(c, seg + w);
        }
        __syncthreads();

        // 2. medium primes
        const uint64_t limit64 = a.oddLimit > j0 ? a.oddLimit - j0 : 0;
        const uint32_t limit = limit64 < kSegOdds ? (uint32_t)limit64 : kSegOdds;
        if (limit) {
            // Each wave strikes its own lane primes first, then takes the
            // cooperative primes two at a time (one per half) from a counter
            // until they run out, so the waves that finish their lane
            // primes first strike more of them and the workgroup reaches
            // the barrier together (dealt out two per wave in turn, the
            // waves' shares differed: 16 GiB map 26.0 -> 24.5 ms, 4 GiB
            // 5.84 -> 5.67).  A prime is struck by one half-wave per
            // segment, whichever it is: coopNext[pi] carries its next strike
            // to the next segment, past the barriers between.
            const uint32_t half = (uint32_t)lane >> 5, hl = (uint32_t)lane & 31;
            // Carried lane primes.  After a prime's last strike j >= limit;
            // with limit == kSegOdds, j - kSegOdds is the next segment's first
            // strike (a segment cut short by oddLimit is the last one the run
            // strikes).  Re-read primes: the empty asm keeps the compiler
            // from hoisting their loads out of the segment loop (holding them
            // after all); the pointer is then cast back to the global address
            // space, since a flat load also counts on the LDS counter and its
            // wait inside the strike loop would wait out every ds_or as well
            // (4 GiB map 10.2 -> 9.9 ms, 70 GB map 241 -> 215 ms together with
            // the one-ahead prefetch).
            const uint32_t *cprGeneric = a.primes + a.ncoop + tid;
            if (!kHoldPrimes) asm volatile("" : "+v"(cprGeneric));
            const __attribute__((address_space(1))) uint32_t *cpr =
                (const __attribute__((address_space(1))) uint32_t *)cprGeneric;
            const uint32_t lane0 = a.ncoop + (uint32_t)tid;
            uint32_t pnext = 0;
            if (!kHoldPrimes && 0 < a.ncarry && lane0 < a.nprimes) pnext = cpr[0];
#pragma unroll
            for (int k = 0; k < kCarry; ++k) {
                // re-read primes: prime k+1 is loaded while prime k strikes
                uint32_t p = pnext;
                if (kHoldPrimes) {
                    p = cp[kHoldPrimes ? k : 0];
                } else if (k + 1 < kCarry && (uint32_t)(k + 1) < a.ncarry &&
                           lane0 + (uint32_t)(k + 1) * kThreads < a.nprimes) {
                    pnext = cpr[(k + 1) * kThreads];
                }
                if ((uint32_t)k < a.ncarry && lane0 + (uint32_t)k * kThreads < a.nprimes) {
                    // mod-30 wheel: pos4 is four times the wheel position of
                    // the next strike (the bit-field extract takes it mod 32)
                    uint32_t j = cn[k] & kOffsetMask, pos4 = (cn[k] >> kWheelPosShift) << 2;
                    uint32_t st = __umul24(p, __builtin_amdgcn_ubfe(kWheel30Steps, pos4, 4));
                    while (j < limit) {
                        atomicOr(&seg[j >> 5], 1u << (j & 31));
                        j += st;
                        pos4 += 4;
                        st = __umul24(p, __builtin_amdgcn_ubfe(kWheel30Steps, pos4, 4));
                    }
                    cn[k] = (j - kSegOdds) | ((pos4 >> 2) & 7u) << kWheelPosShift;
                }
            }
            // primes beyond the carried ones: first strike from j0 in
            // doubles, the next prime loaded while this one strikes
            const double j0d = (double)j0;
            uint32_t pi = a.ncoop + a.ncarry * kThreads + tid, pn = 0;
            // opaque per segment, so the compiler forms the prime and
            // reciprocal addresses here instead of holding them (spilled, in
            // the 64-register kernel) through the whole run
            asm volatile("" : "+v"(pi));
            double ivn = 0.0;
            if (pi < a.nprimes) {
                pn = a.primes[pi];
                ivn = a.invp[pi];
            }
            for (; pi < a.nprimes; pi += kThreads) {
                const uint32_t p = pn;
                const double iv = ivn;
                if (pi + kThreads < a.nprimes) {
                    pn = a.primes[pi + kThreads];
                    ivn = a.invp[pi + kThreads];
                }
                for (uint32_t j = first_strike_below(j0d, p, iv, limit); j < limit; j += p)
                    atomicOr(&seg[j >> 5], 1u << (j & 31));
            }
            for (;;) {
                uint32_t pb = 0;
                if (lane == 0) pb = atomicAdd(&coopTaken, 2u);
                pb = __builtin_amdgcn_readfirstlane(pb);
                if (pb >= a.ncoop) break;
                const uint32_t pi = pb + half;
                if (pi < a.ncoop && pi < a.nwheel30) {
                    // Mod-30 wheel: the odd multiples m*p with 3 | m or
                    // 5 | m are already struck (3 and 5 are pattern
                    // primes).  Two progressions of odd-index stride 3p
                    // hold m = 1 and m = 5 (mod 6); lane l of the half
                    // takes elements l + 32t of each, a stride of 96p odd
                    // values = 3p whole words, so its bit stays put.  Along
                    // a lane m grows by 192 = 2 (mod 5) per element, so one
                    // element in five is a multiple of 5 and is stepped
                    // over: each step of the lane's loop strikes 4 of 5
                    // elements at offsets fixed for the lane.  coopNext
                    // keeps the first element at or past the segment's
                    // start as offset << 4 | (m mod 5) << 1 | (m = 5 mod 6).
                    // Larger wheel primes take the mod-6 block below: its
                    // set-up is shorter, and the kernel is issue-bound.
                    const uint32_t p = a.primes[pi];
                    uint32_t first, five, first5;
                    if (si == runLo) {
                        uint64_t f = first_strike(j0, p, a.invp[pi]);  // < (p*p-1)/2 < 2^32
                        // its multiple m (exact: 2(j0+f)+1 < 2^53 is m*p)
                        const uint64_t m = (uint64_t)rint((2.0 * (double)(j0 + f) + 1.0) * a.invp[pi]);
                        uint32_t r30 = (uint32_t)(m % 30);
                        if (r30 % 6 == 3) {  // m + 2 = 5 (mod 6): one multiple on
                            f += p;
                            r30 += 2;
                        }
                        first = (uint32_t)f;
                        five = r30 % 6 == 5;
                        first5 = r30 % 5;
                    } else {
                        const uint32_t v = coopNext[pi];
                        first = v >> 4;
                        first5 = (v >> 1) & 7u;
                        five = v & 1u;
                    }
                    const uint32_t p3 = 3 * p;  // < 2^16
                    const uint32_t stepBytes = 4u * p3, endBytes = ((limit + 31) >> 5) << 2;
                    char *segBytes = reinterpret_cast<char *>(seg);
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        // the progression of first (r = 0), then the other
                        // residue: from m = 1 the next m = 5 is m + 4
                        // (index + 2p), from m = 5 the next m = 1 is m + 2
                        // (index + p)
                        const uint32_t fr = r == 0 ? first : first + (five ? p : 2 * p);
                        const uint32_t fiveR = r == 0 ? five : 1u - five;
                        const uint32_t fr5 = r == 0 ? first5 : first5 + (five ? 2u : 4u);
                        // m mod 5 of the lane's element 0 (m + 6l), then of
                        // element t: c5 + 2t, a multiple of 5 at t = 2 c5
                        // (mod 5)
                        const uint32_t c5 = (fr5 + hl) % 5u;
                        const uint32_t jr = fr + __umul24(hl, p3);
                        uint32_t je = jr, t = 0;
                        if (jr < limit) {
                            const uint32_t mask = 1u << (jr & 31);
                            const uint32_t skip = (2u * c5) % 5u;
                            uint32_t ad = (jr >> 5) << 2;
                            const uint32_t o0 = skip == 0 ? stepBytes : 0u, o1 = o0 + (skip == 1 ? 2u : 1u) * stepBytes,
                                           o2 = o1 + (skip == 2 ? 2u : 1u) * stepBytes,
                                           o3 = o2 + (skip == 3 ? 2u : 1u) * stepBytes;
                            // a step covers five elements and runs only when
                            // the fifth is below the end too, even when it is
                            // the one stepped over (skip == 4, o3 = 3 steps),
                            // so the lane stops on its first element past it
                            const uint32_t endGroup =
                                skip == 4 ? (endBytes > stepBytes ? endBytes - stepBytes : 0u) : endBytes;
                            uint32_t a0 = ad + o0, a1 = ad + o1, a2 = ad + o2, a3 = ad + o3;
                            for (; a3 < endGroup; a0 += 5 * stepBytes, a1 += 5 * stepBytes, a2 += 5 * stepBytes,
                                                  a3 += 5 * stepBytes) {
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + a0), mask);
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + a1), mask);
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + a2), mask);
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + a3), mask);
                                // four addresses, each advanced once per step
                                // (left to itself the compiler recomputes
                                // them from a common offset, six more VALU)
                                asm("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3));
                            }
                            // the last, partial step: at most four elements
                            // below the end, struck in offset order (a3 is
                            // below it only when skip == 4), and t of them
                            ad = a0 - o0;
                            if (a0 < endBytes) atomicOr(reinterpret_cast<uint32_t *>(segBytes + a0), mask);
                            if (a1 < endBytes) atomicOr(reinterpret_cast<uint32_t *>(segBytes + a1), mask);
                            if (a2 < endBytes) atomicOr(reinterpret_cast<uint32_t *>(segBytes + a2), mask);
                            if (a3 < endBytes) atomicOr(reinterpret_cast<uint32_t *>(segBytes + a3), mask);
                            t = (ad < endBytes) + (ad + stepBytes < endBytes) + (ad + 2 * stepBytes < endBytes) +
                                (ad + 3 * stepBytes < endBytes);
                            je = ((ad + t * stepBytes) << 3) | (jr & 31);
                        }
                        // the 32 lanes end on 32 consecutive elements at or
                        // past kSegOdds; the one ending below kSegOdds + 3p
                        // (lane 0 when the progression did not strike) holds
                        // its first in the next segment, and the smaller of
                        // the two progressions' firsts is kept
                        const uint32_t d = je - kSegOdds;
                        if (limit == kSegOdds && (d < p3 || (hl == 0 && fr >= kSegOdds))) {
                            const uint32_t v = d << 4 | ((c5 + 2u * t) % 5u) << 1 | fiveR;
                            if (r == 0) coopNext[pi] = v;
                            else atomicMin(&coopNext[pi], v);
                        }
                    }
                } else if (pi < a.ncoop && pi < a.nwheel) {
                    // Mod-6 wheel: the odd multiples m*p with 3 | m are
                    // already struck (3 is a pattern prime), so only
                    // m = 1 and m = 5 (mod 6) are struck, as two progressions
                    // of odd-index stride 3p; lane l of the half strikes
                    // elements l + 32t of each, a stride of 96p odd values =
                    // 3p whole words, so its bit stays put.  coopNext keeps
                    // the first wheel multiple at or past the segment's start
                    // as offset << 1 | (m = 5 mod 6).
                    const uint32_t p = a.primes[pi];
                    uint32_t first, five;
                    if (si == runLo) {
                        uint64_t f = first_strike(j0, p, a.invp[pi]);  // < (p*p-1)/2 < 2^32
                        // its multiple m (exact: 2(j0+f)+1 < 2^53 is m*p)
                        const uint64_t m = (uint64_t)rint((2.0 * (double)(j0 + f) + 1.0) * a.invp[pi]);
                        uint32_t r6 = (uint32_t)(m % 6);
                        if (r6 == 3) {  // m + 2 = 5 (mod 6): one multiple on
                            f += p;
                            r6 = 5;
                        }
                        first = (uint32_t)f;
                        five = r6 == 5;
                    } else {
                        const uint32_t v = coopNext[pi];
                        first = v >> 1;
                        five = v & 1u;
                    }
                    const uint32_t p3 = 3 * p;  // < 2^16
                    const uint32_t stepBytes = 4u * p3, endBytes = ((limit + 31) >> 5) << 2;
                    char *segBytes = reinterpret_cast<char *>(seg);
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        // the progression of first (r = 0), then the other
                        // residue: from m = 1 the next m = 5 is m + 4
                        // (index + 2p), from m = 5 the next m = 1 is m + 2
                        const uint32_t fr = r == 0 ? first : first + (five ? p : 2 * p);
                        const uint32_t fiveR = r == 0 ? five : 1u - five;
                        const uint32_t jr = fr + __umul24(hl, p3);
                        uint32_t je = jr;
                        if (jr < limit) {
                            const uint32_t mask = 1u << (jr & 31);
                            uint32_t ad = (jr >> 5) << 2;
                            if (kSegBytes == 32768) {
                                const uint32_t endPair = endBytes > stepBytes ? endBytes - stepBytes : 0u;
                                for (; ad < endPair; ad += 2 * stepBytes) {
                                    atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad), mask);
                                    atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad + stepBytes), mask);
                                }
                            }
                            for (; ad < endBytes; ad += stepBytes)
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad), mask);
                            je = (ad << 3) | (jr & 31);
                        }
                        // the 32 lanes end on 32 consecutive elements at or
                        // past kSegOdds; the one ending below kSegOdds + 3p
                        // (lane 0 when the progression did not strike) holds
                        // its first in the next segment, and the smaller of
                        // the two progressions' firsts is kept
                        const uint32_t d = je - kSegOdds;
                        if (limit == kSegOdds && (d < p3 || (hl == 0 && fr >= kSegOdds))) {
                            const uint32_t v = d << 1 | fiveR;
                            if (r == 0) coopNext[pi] = v;
                            else atomicMin(&coopNext[pi], v);
                        }
                    }
                } else if (pi < a.ncoop) {
                    const uint32_t p = a.primes[pi];
                    uint32_t first;
                    if (si == runLo) {
                        const uint64_t f = first_strike(j0, p, a.invp[pi]);  // < (p*p-1)/2 < 2^32
                        first = (uint32_t)f;
                    } else {
                        first = coopNext[pi];
                    }
                    // lane l of the half strikes multiples first + (l + 32t)p:
                    // a stride of 32p odd values is p whole words, so the bit
                    // stays put and only the word address advances (bits past
                    // limit in its last word are cleared again at write-out)
                    const uint32_t j = first + __umul24(hl, p);  // p < 2^16
                    uint32_t je = j;  // the lane's first multiple at or past limit
                    if (j < limit) {
                        const uint32_t mask = 1u << (j & 31);
                        const uint32_t stepBytes = 4u * p, endBytes = ((limit + 31) >> 5) << 2;
                        char *segBytes = reinterpret_cast<char *>(seg);
                        uint32_t ad = (j >> 5) << 2;
                        if (kSegBytes == 32768) {
                            // two strikes per step while both are below the
                            // end (4 GiB map 9.08 -> 8.93 ms; the large-segment
                            // kernels measured slower with it)
                            const uint32_t endPair = endBytes > stepBytes ? endBytes - stepBytes : 0u;
                            for (; ad < endPair; ad += 2 * stepBytes) {
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad), mask);
                                atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad + stepBytes), mask);
                            }
                        }
                        for (; ad < endBytes; ad += stepBytes)
                            atomicOr(reinterpret_cast<uint32_t *>(segBytes + ad), mask);
                        je = (ad << 3) | (j & 31);
                    }
                    // first strike from the next segment's start: the 32
                    // lanes end on 32 consecutive multiples at or past
                    // kSegOdds, and the one lane that ends below kSegOdds + p
                    // holds the first (lane 0 when the prime did not strike)
                    const uint32_t d = je - kSegOdds;
                    if (limit == kSegOdds && (d < p || (hl == 0 && first >= kSegOdds))) coopNext[pi] = d;
                }
            }
        }
        __syncthreads();

        // 3. write-out: 16 bytes (4 words) per thread per chunk.  A segment
        // wholly inside the written range, past value 1 and below M, with
        // an aligned destination (every one but the first and last of a
        // map), needs no per-word checks (16 GiB map 28.07 -> 27.64 ms,
        // 4 GiB 6.63 -> 6.50); the kernel that re-reads its lane primes has
        // no registers to spare for the second loop (70 GB map 144.5 ->
        // 146.4 ms with it) and keeps the checked one.
        const uint64_t lo = a.byteLo > segByte0 ? a.byteLo : segByte0;
        const uint64_t hiSeg = segByte0 + kSegBytes;
        const uint64_t hi = a.byteHi < hiSeg ? a.byteHi : hiSeg;
        uint8_t *const segDst = a.dst + (segByte0 - a.dstBase);
        const bool interior = segByte0 >= a.byteLo && hiSeg <= a.byteHi && j0 != 0 && j0 + kSegOdds <= a.oddLimit &&
                              ((uintptr_t)segDst & 15) == 0;
        if (kFastWriteOut && interior) {
            for (uint32_t ch = tid; ch < kChunks; ch += kThreads) {
                const uint4 sv = *reinterpret_cast<const uint4 *>(seg + 4 * ch);
                *reinterpret_cast<uint4 *>(segDst + 16ull * ch) =
                    make_uint4(to_map_order(~sv.x), to_map_order(~sv.y), to_map_order(~sv.z), to_map_order(~sv.w));
            }
        } else {
            for (uint32_t ch = tid; ch < kChunks; ch += kThreads) {
                const uint4 sv = *reinterpret_cast<const uint4 *>(seg + kFillWords * ch);
                const uint64_t b0 = segByte0 + 16ull * ch;
                if (b0 + 16 <= lo || b0 >= hi) continue;
                uint32_t out[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t w = 4 * ch + k;
                    uint32_t v = out[k];
                    const uint64_t jw = j0 + 32ull * w;  // odd index of bit 0
                    if (jw == 0) {
                        // value 1 is not prime; the pattern primes themselves are
                        v |= 1u;
                        for (int g = 0; g < kGroups; ++g)
                            for (int t = 0; t < 3; ++t)
                                if (kGroupPrimes[g][t]) v &= ~(1u << ((kGroupPrimes[g][t] - 1) >> 1));
                    }
                    if (jw + 32 > a.oddLimit) {
                        // values above M keep their initial 1 (never struck)
                        const uint64_t keep = a.oddLimit > jw ? a.oddLimit - jw : 0;
                        v &= keep >= 32 ? 0xFFFFFFFFu : ((1u << keep) - 1u);
                    }
                    out[k] = to_map_order(~v);
                }
                uint8_t *d = a.dst + (b0 - a.dstBase);
                if (b0 >= lo && b0 + 16 <= hi && ((uintptr_t)d & 15) == 0) {
                    *reinterpret_cast<uint4 *>(d) = make_uint4(out[0], out[1], out[2], out[3]);
                } else {
                    const uint8_t *ob = reinterpret_cast<const uint8_t *>(out);
                    for (int k = 0; k < 16; ++k)
                        if (b0 + k >= lo && b0 + k < hi) d[k] = ob[k];
                }
            }
        }
        __syncthreads();

#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const uint32_t n = cur[g] + 4u * stepSeg[g];
            cur[g] = min(n, n - 4u * period[g]);
        }
    }
}

uint64_t isqrt_u64(uint64_t n)
{
    uint64_t r = (uint64_t)std::sqrt((double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}

}  // namespace

struct ffx_sieve_plan {
    uint64_t maxPrimeMapValue;
    uint32_t nprimes;
    std::vector<uint32_t> hostPrimes;  // the struck primes (> 61)
    uint32_t segBytes, blocksPerCU, cus;
    uint32_t laneCarry;   // at most this many carried lane primes per thread (FFX_SIEVE_LANE_CARRY)
    uint32_t holdPrimes;  // large segments: 0 (held) or 1 held, 2 re-read
    uint32_t *dPrimes;
    double *dInv;
    uint32_t *dPatterns;
    int device;
};

using namespace ffx;

static int plan_build(ffx_sieve_plan **out, std::vector<uint32_t> &oddPrimes, uint64_t M)
{
    if (!out) return fail(FFX_EINVAL, "sieve plan: null out");
    *out = nullptr;
    // keep the primes the kernel strikes: 61 < p <= isqrt(M)
    const uint64_t sq = isqrt_u64(M);
    std::vector<uint32_t> med;
    for (uint32_t p : oddPrimes)
        if (p > kLargestPatternPrime && p <= sq) med.push_back(p);
    std::vector<double> inv(med.size());
    for (size_t i = 0; i < med.size(); ++i) inv[i] = 1.0 / (double)med[i];
    // word patterns: bit b of word i of group g is set when odd index 32i+b
    // is congruent to (p-1)/2 mod some prime p of the group
    std::vector<uint32_t> pat(kPatWords, 0);
    for (int g = 0; g < kGroups; ++g) {
        const uint32_t P = group_period(g), off = pattern_base(g);
        for (uint64_t j = 0; j < 32ull * P; ++j)
            for (int t = 0; t < 3; ++t) {
                const uint32_t p = kGroupPrimes[g][t];
                if (p && j % p == (p - 1) / 2) pat[off + j / 32] |= 1u << (j % 32);
            }
        for (uint32_t i = 0; i < kPatWrap; ++i) pat[off + P + i] = pat[off + i];
    }
    ffx_sieve_plan *pl = new (std::nothrow) ffx_sieve_plan();
    if (!pl) return fail(FFX_ENOMEM, "sieve plan: host alloc");
    pl->maxPrimeMapValue = M;
    pl->nprimes = (uint32_t)med.size();
    pl->hostPrimes = med;
    const bool large = M >= kLargeMapFrom;
    pl->segBytes = large ? kLargeMapSegBytes : kSmallMapSegBytes;
    pl->blocksPerCU = large ? kLargeMapBlocksPerCU : kSmallMapBlocksPerCU;
    pl->laneCarry = kCarryReread;
    pl->holdPrimes = 0;
    hipError_t e = hipGetDevice(&pl->device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
    if (e == hipSuccess) e = hipMalloc(&pl->dPrimes, (med.size() + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&pl->dInv, (med.size() + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&pl->dPatterns, kPatWords * 4);
    if (e == hipSuccess && !med.empty()) e = hipMemcpy(pl->dPrimes, med.data(), med.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && !med.empty()) e = hipMemcpy(pl->dInv, inv.data(), inv.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(pl->dPatterns, pat.data(), kPatWords * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "sieve plan upload");
        ffx_sieve_plan_destroy(pl);
        return rc;
    }
    pl->cus = cus > 0 ? (uint32_t)cus : 256;
    *out = pl;
    return FFX_OK;
}

static int launch(const ffx_sieve_plan *pl, uint8_t *dst, uint64_t dstBase, uint64_t byteLo, uint64_t byteHi,
                  void *stream)
{
    if (!pl) return fail(FFX_EINVAL, "sieve: null plan");
    if (!dst) return fail(FFX_EINVAL, "sieve: null destination");
    if (byteHi <= byteLo) return FFX_OK;
    SieveArgs a;
    a.dst = dst;
    a.dstBase = dstBase;
    a.byteLo = byteLo;
    a.byteHi = byteHi;
    a.oddLimit = (pl->maxPrimeMapValue + 1) >> 1;
    a.primes = pl->dPrimes;
    a.invp = pl->dInv;
    a.nprimes = pl->nprimes;
    const uint64_t segBytes = pl->segBytes, segOdds = segBytes * 8;
    // primes with more than coop_div strikes per segment strike wave-cooperatively
    a.ncoop = (uint32_t)(std::lower_bound(pl->hostPrimes.begin(), pl->hostPrimes.end(),
                                          (uint32_t)(segOdds / coop_div((uint32_t)segOdds))) -
                         pl->hostPrimes.begin());
    const uint32_t coopCap = segBytes == kSmallMapSegBytes ? coop_capacity(kSmallMapSegBytes * 8) : coop_capacity(kLargeMapSegBytes * 8);
    if (a.ncoop > coopCap) return fail(FFX_EINVAL, "sieve: %u cooperative primes exceed the table", a.ncoop);
    // the mod-6 wheel for the cooperative primes up to segOdds / kWheelDiv
    a.nwheel = (uint32_t)(std::lower_bound(pl->hostPrimes.begin(), pl->hostPrimes.end(), (uint32_t)(segOdds / kWheelDiv)) -
                          pl->hostPrimes.begin());
    a.nwheel30 = std::min(a.nwheel, (uint32_t)(std::lower_bound(pl->hostPrimes.begin(), pl->hostPrimes.end(),
                                                                 (uint32_t)(segOdds / kWheel30Div)) -
                                               pl->hostPrimes.begin()));
    a.patterns = pl->dPatterns;
    a.segFirst = byteLo / segBytes;
    a.segCount = (byteHi + segBytes - 1) / segBytes - a.segFirst;
    if (a.segFirst * segOdds >= (1ull << 52)) return fail(FFX_ERANGE, "sieve: map offset beyond 2^53 values");
    const uint64_t maxGrid = (uint64_t)pl->cus * pl->blocksPerCU;
    uint64_t grid = a.segCount < maxGrid ? a.segCount : maxGrid;
    // each workgroup's run of segments stays within kRunOddsLimit odd values
    // (its carried strike offsets have 29 bits); when that takes more
    // workgroups than the grid, a whole multiple of the CU count keeps the
    // last wave from running a few workgroups alone (70 GB map, one 112 KiB
    // workgroup per CU, runs of at most 2^31 values: 261 runs took 357 ms,
    // 512 take 188)
    const uint64_t maxRun = kRunOddsLimit / segOdds;  // segments per run
    const uint64_t minGrid = (a.segCount + maxRun - 1) / maxRun;
    if (grid < minGrid) grid = std::min<uint64_t>((minGrid + pl->cus - 1) / pl->cus * pl->cus, a.segCount);
    const bool hold = pl->holdPrimes != 2;
    if (segBytes == 32768) {
        a.ncarry = std::min<uint32_t>(pl->laneCarry, kCarrySmall);
        hipLaunchKernelGGL((sieve_segments<32768, kCarrySmall, false>), dim3((uint32_t)grid), dim3(kThreads), 0,
                           as_stream(stream), a);
    } else if (hold) {
        a.ncarry = std::min<uint32_t>(pl->laneCarry, kCarryHeld);
        hipLaunchKernelGGL((sieve_segments<kLargeMapSegBytes, kCarryHeld, true>), dim3((uint32_t)grid), dim3(kThreads), 0,
                           as_stream(stream), a);
    } else {
        a.ncarry = std::min<uint32_t>(pl->laneCarry, kCarryReread);
        hipLaunchKernelGGL((sieve_segments<kLargeMapSegBytes, kCarryReread, false>), dim3((uint32_t)grid), dim3(kThreads), 0,
                           as_stream(stream), a);
    }
    FFX_HIP(hipGetLastError());
    return FFX_OK;
}

extern "C" {

int ffx_sieve_plan_create(ffx_sieve_plan **out, uint64_t maxPrimeMapValue)
{
    const uint64_t sq = isqrt_u64(maxPrimeMapValue);
    if (sq > 0xFFFFFFFFull) return fail(FFX_ERANGE, "sieve plan: maxPrimeMapValue too large");
    std::vector<uint8_t> comp(sq + 1, 0);
    std::vector<uint32_t> primes;
    for (uint64_t p = 3; p <= sq; p += 2) {
        if (comp[p]) continue;
        primes.push_back((uint32_t)p);
        for (uint64_t m = p * p; m <= sq; m += 2 * p) comp[m] = 1;
    }
    return plan_build(out, primes, maxPrimeMapValue);
}

int ffx_sieve_plan_create_from_primes(ffx_sieve_plan **out, const uint32_t *h_primes, uint32_t numPrimes,
                                      uint64_t maxPrimeMapValue)
{
    if (!h_primes && numPrimes) return fail(FFX_EINVAL, "sieve plan: null prime list");
    std::vector<uint32_t> primes;
    uint32_t prev = 0;
    for (uint32_t i = 0; i < numPrimes; ++i) {
        const uint32_t p = h_primes[i];
        if (p == 2) continue;
        if (p < 3 || !(p & 1) || p <= prev) return fail(FFX_EINVAL, "sieve plan: prime list not odd ascending");
        primes.push_back(p);
        prev = p;
    }
    const uint64_t sq = isqrt_u64(maxPrimeMapValue);
    if (sq >= 3 && (primes.empty() || primes.back() < sq)) {
        // the list must reach isqrt(M): the largest odd prime <= sq must be in it
        uint64_t q = sq | 1;
        if (q > sq) q -= 2;
        for (; q >= 3; q -= 2) {
            bool isp = true;
            for (uint64_t d = 3; d * d <= q; d += 2)
                if (q % d == 0) { isp = false; break; }
            if (isp) break;
        }
        if (q >= 3 && (primes.empty() || primes.back() < q))
            return fail(FFX_EINVAL, "sieve plan: prime list stops below isqrt(maxPrimeMapValue)");
    }
    return plan_build(out, primes, maxPrimeMapValue);
}

int ffx_sieve_plan_set(ffx_sieve_plan *plan, int param, uint32_t value)
{
    if (!plan) return fail(FFX_EINVAL, "sieve plan: null");
    switch (param) {
    case FFX_SIEVE_SEGMENT_BYTES:
        if (value != kSmallMapSegBytes && value != kLargeMapSegBytes)
            return fail(FFX_ERANGE, "segment bytes %u: %u or %u", value, kSmallMapSegBytes, kLargeMapSegBytes);
        plan->segBytes = value;
        return FFX_OK;
    case FFX_SIEVE_BLOCKS_PER_CU:
        if (value < 1 || value > 8) return fail(FFX_ERANGE, "sieve blocks per CU %u outside 1..8", value);
        plan->blocksPerCU = value;
        return FFX_OK;
    case FFX_SIEVE_LANE_CARRY:
        if (value > (uint32_t)kCarryReread) return fail(FFX_ERANGE, "sieve lane carry %u outside 0..%d", value, kCarryReread);
        plan->laneCarry = value;
        return FFX_OK;
    case FFX_SIEVE_HOLD_PRIMES:
        if (value > 2) return fail(FFX_ERANGE, "sieve hold primes %u outside 0..2", value);
        plan->holdPrimes = value;
        return FFX_OK;
    default:
        return fail(FFX_EINVAL, "sieve plan: unknown parameter %d", param);
    }
}

int ffx_sieve_plan_get(const ffx_sieve_plan *plan, int param, uint32_t *value)
{
    if (!plan || !value) return fail(FFX_EINVAL, "sieve plan: null");
    switch (param) {
    case FFX_SIEVE_SEGMENT_BYTES: *value = plan->segBytes; return FFX_OK;
    case FFX_SIEVE_BLOCKS_PER_CU: *value = plan->blocksPerCU; return FFX_OK;
    case FFX_SIEVE_LANE_CARRY: *value = plan->laneCarry; return FFX_OK;
    case FFX_SIEVE_HOLD_PRIMES: *value = plan->holdPrimes; return FFX_OK;
    default: return fail(FFX_EINVAL, "sieve plan: unknown parameter %d", param);
    }
}

int ffx_sieve_plan_destroy(ffx_sieve_plan *plan)
{
    if (!plan) return FFX_OK;
    (void)hipFree(plan->dPrimes);
    (void)hipFree(plan->dInv);
    (void)hipFree(plan->dPatterns);
    delete plan;
    return FFX_OK;
}

int ffx_sieve_slab_async(const ffx_sieve_plan *plan, uint64_t segLo, uint64_t segHi, void *d_slab, void *stream)
{
    if (segLo & 15) return fail(FFX_EINVAL, "sieve slab: segLo %llu not a multiple of 16", (unsigned long long)segLo);
    if (segHi <= segLo) return FFX_OK;
    const uint64_t byteLo = segLo >> 4, byteHi = (segHi + 15) >> 4;
    return launch(plan, static_cast<uint8_t *>(d_slab), byteLo, byteLo, byteHi, stream);
}

int ffx_sieve_map_async(const ffx_sieve_plan *plan, void *d_map, uint64_t mapBytes, void *stream)
{
    return launch(plan, static_cast<uint8_t *>(d_map), 0, 0, mapBytes, stream);
}

int ffx_sieve_map_range_async(const ffx_sieve_plan *plan, void *d_map, uint64_t byteLo, uint64_t byteHi,
                              void *stream)
{
    return launch(plan, static_cast<uint8_t *>(d_map), 0, byteLo, byteHi, stream);
}

}  // extern "C"
