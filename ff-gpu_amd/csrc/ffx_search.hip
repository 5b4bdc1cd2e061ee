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
// of a wavefront visit splits in parallel and count with ballots.
//
// Per-sum work differs by five orders of magnitude (most sums stop after a
// few hundred splits; emitted sums visit all ~s/4 splits), so waves pull sums
// from one atomic counter, largest sums first.  Splits are visited k
// descending as the reference does (it stops far earlier on non-solutions
// than k ascending), flattened over all k so every lane is busy; each k
// lists only the classes of d mod 30 that can hold a valid split (the
// Power2Odd family, both probe values coprime to 15).  Every listed split
// first goes through the filters that need no map read: the Power2Even
// family test and either, in long k-segments, one bit of a sieve over d by
// the primes 7..251, or a residue filter (a probe value with a prime factor
// in 7..23 is composite; the residues of both probe values come from those
// of d, o and 2^k).
// The few survivors are compacted into a per-wave LDS probe queue and probed
// 128 at a time, so the random map reads run on full waves; splits that pass
// both PTP probes go to a second queue for AllButOne, resolved 64 at a time,
// one per lane, after a pre-check that tests the terms of the first divisor
// pairs for small factors, reads those that pass and rejects almost all
// splits.  While a sum is young both queues
// are emptied every round (early exit at two valid splits).  Terms d, o < s
// are factored with a smallest-prime-factor table, so no 64-bit trial
// division runs on the device.
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ffx_internal.h"

// Work counters, compiled in only for the diagnostic library
// (make -C ff-gpu_amd counters -> lib/libffx_counters.so, FFX_PROBE_COUNTERS):
// map reads by kind, AllButOne calls, Miller-Rabin tests and splits, summed
// over a launch with device atomics (which slow the kernel; the counts, not
// the time, are what that build is for).
#ifdef FFX_PROBE_COUNTERS
__device__ unsigned long long g_ffxCounters[16];
#define FFX_COUNT(i, v) atomicAdd(&g_ffxCounters[i], (unsigned long long)(v))
#else
#define FFX_COUNT(i, v) ((void)0)
#endif

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kQueueCap = 192;     // < 64 left after a drain + 128 from one probe batch
// Two splits per lane per filter round: with long segments scanned off the
// sieve over d, the filter rounds are a small share and two measure as fast
// as four (odd sums to 2^21: 27.81 ms both), with a smaller probe queue (four
// won when every split went through the rounds: 26.4 against 28.8 ms).
#ifndef FFX_ROUND_SPLITS
#define FFX_ROUND_SPLITS 2
#endif
constexpr int kRoundSplits = FFX_ROUND_SPLITS;  // splits per lane per filter round
constexpr int kProbeCap = 128 + 64 * kRoundSplits;  // < 128 left after the batches + one round
static_assert(kRoundSplits >= 1 && kRoundSplits <= 4, "1..4 splits per lane per round");
constexpr uint32_t kNeedO = 1u << 31;  // probe entry: o must be prime too (Power2Even, 3 !| d*o)
constexpr uint32_t kGrab = 4;      // default sums taken per atomic (the autotuned choice)
constexpr int kMaxPrimes = 12;     // distinct odd primes of q < 3^31 (q <= s^2/16, s < 2^26)
constexpr uint32_t kDBits = 27;    // queue entry: d | k << 27
constexpr int kMaxSegs = 32;
// Default: while a sum has visited fewer splits than this, every queued
// candidate is resolved after each step (most sums stop within a few hundred
// splits); beyond it candidates are resolved 64 at a time.
// Long k-segments are scanned off the sieve over d only past this point too.
constexpr uint32_t kEagerSplits = 256;  // the autotuned choice for odd sums to 2^20 and 2^21
constexpr uint32_t kBlocksPerCU = 4;  // default persistent grid: CUs x this (all resident)

// Residue filter: a probe value that shares a factor with one of the moduli
// in use and is not that factor itself is composite without touching the
// map.  One bit table per modulus, bit r set when gcd(r, m) == 1 or r is a
// prime factor of m: a value with a clear bit is then composite at any size
// (a set bit only means "not ruled out": a value congruent to a prime factor
// of m goes on to the map read).  Each modulus is below 2^16, so
// residues of the two probe values 2^k + d*o - 2 and d + o*2^k - 2 come from
// the 32-bit residues of d, o and 2^k with 32-bit products.
constexpr uint32_t kMods[5] = {3 * 5 * 7 * 11 * 13, 17 * 19 * 23, 29 * 31, 37 * 41, 43 * 47};
// Two moduli measured fastest on MI355X with the mod-30 split classes, the
// sieve over d and the AllButOne pre-check: [5, 2^20) search 11.4 ms with
// two, 11.9 with one or three; [5, 2^21) 45.2 ms with two, 46.3 with three,
// 48.9 with one (before the sieve over d, three were fastest: 19.7 ms
// against 20.2 with two).  The classes already exclude 3 and 5; the tables
// add 7..23.
#ifndef FFX_RESIDUE_MODULI
#define FFX_RESIDUE_MODULI 2
#endif
constexpr int kNMods = FFX_RESIDUE_MODULI;
static_assert(kNMods >= 1 && kNMods <= 5, "1..5 residue moduli");
constexpr uint32_t tab_offset(int i) { return i == 0 ? 0 : tab_offset(i - 1) + (kMods[i - 1] + 31) / 32; }
constexpr uint32_t kTabWords = tab_offset(kNMods);

// Sieve over d: in a long k-segment, a chunk of kSieveBits consecutive odd d
// is marked where 2^k + d*o - 2 (quadratic in d: two residue classes mod q
// at most) or d + o*2^k - 2 (linear: one class) is divisible by a prime q
// of kSievePrimes; splits on marked d are dropped before any map read, with
// no per-split residue arithmetic (the split classes mod 30 already exclude
// 3 and 5).  Both values exceed 2*sqrt(s-1) - 2 > 251 for every split of a
// sum s >= kSieveMinSum, so a marked value is composite.  The primes 7..61
// mark many bits each and are laid in as 64-bit word masks, every lane
// building its own 64 bits from all their roots; the primes from 67 strike
// their bits one lane per (prime, value, root).  The chunk's bits live in
// 512 bytes of LDS per wave.
constexpr uint32_t kSievePrimes[135] = {
       7,   11,   13,   17,   19,   23,   29,   31,   37,   41,   43,   47,   53,
      59,   61,   67,   71,   73,   79,   83,   89,   97,  101,  103,  107,  109,
     113,  127,  131,  137,  139,  149,  151,  157,  163,  167,  173,  179,  181,
     191,  193,  197,  199,  211,  223,  227,  229,  233,  239,  241,  251,  257,
     263,  269,  271,  277,  281,  283,  293,  307,  311,  313,  317,  331,  337,
     347,  349,  353,  359,  367,  373,  379,  383,  389,  397,  401,  409,  419,
     421,  431,  433,  439,  443,  449,  457,  461,  463,  467,  479,  487,  491,
     499,  503,  509,  521,  523,  541,  547,  557,  563,  569,  571,  577,  587,
     593,  599,  601,  607,  613,  617,  619,  631,  641,  643,  647,  653,  659,
     661,  673,  677,  683,  691,  701,  709,  719,  727,  733,  739,  743,  751,
     757,  761,  769,  773,  787};
// Masked 7..61 and struck 67..167 measured fastest (odd sums [5, 2^21)
// search: 33.2 ms, against 34.2 with 7..43 / 47..139, 34.9 with 7..31 /
// 37..131 and 34.5 with 7..61 / 67..131).
constexpr int kNSmallSieve = 15;  // 7..61: word masks
static_assert(kSievePrimes[kNSmallSieve - 1] < 64 && kSievePrimes[kNSmallSieve] > 64, "masked primes below 64");
constexpr int kNLargeSieve = 21;  // 67..167 struck: (prime, value, root) items, 3 per prime
constexpr int kNFarSieve = 15;    // 173..251 struck, one item per lane (few bits each): odd sums to
                                  // 2^21 27.8 -> 25.5 ms, map reads 1.75 G -> 1.53 G
// Beyond: kNFar2Sets sets of 21 primes (257..787), one item per lane in
// each, their state in LDS (q | s mod q << 10 | next bit << 20 | kind << 30),
// the next bit carried from chunk to chunk by the strike loop itself; for
// sums from 2^18 (odd sums to 2^21: 25.5 -> 21.9 ms, map reads 1.53 G ->
// 1.12 G).
#ifndef FFX_FAR2_SETS
#define FFX_FAR2_SETS 4
#endif
constexpr int kNFar2Sets = FFX_FAR2_SETS;
constexpr int kNSievePrimes = kNSmallSieve + kNLargeSieve + kNFarSieve + 21 * kNFar2Sets;
constexpr int kFar2First = kNSmallSieve + kNLargeSieve + kNFarSieve;  // index of the first far-2 prime
constexpr uint32_t kSieveBits = 4096;            // odd d per chunk (512 bytes of LDS per wave)
constexpr uint32_t kSieveMinSum = 1u << 15;
constexpr uint32_t kSieveMinListed = 512;  // segments shorter than this are not sieved (256..1024 measured alike, 2048 0.4 % slower)
// square roots mod each sieve prime: 256 bytes per prime (an address from
// the prime's index and the residue alone, no offset table to read first)
constexpr uint32_t kSqrtStride = 1024;  // entries per prime (primes below 1024)
constexpr uint32_t kSieveSqrtBytes = kNSievePrimes * kSqrtStride * 2;
static_assert(2 * 181 - 2 > kSievePrimes[kFar2First - 1], "sieve primes below the smallest probe value of a sum >= kSieveMinSum");
// the far-2 sets only for sums from 2^18: their probe values exceed
// 2 sqrt(2^18 - 1) - 2 = 1022 > every far-2 prime
constexpr uint32_t kFar2MinSum = 1u << 18;
static_assert(kNFar2Sets == 0 || 1022 > kSievePrimes[kNSievePrimes - 1], "far-2 primes below the probe values of a sum >= kFar2MinSum");
static_assert(kSievePrimes[kNSievePrimes - 1] < kSqrtStride, "square roots and inverses are 16-bit, one row per prime");

struct SearchArgs {
    const uint8_t *map;
    uint64_t M;               // bitmap answers odd n <= M
    const uint16_t *spf;      // smallest prime factor of odd n (index n>>1), 0 if n prime or 1
    uint32_t spfLimit;        // spf covers odd n < spfLimit
    uint32_t first;           // first odd sum (single range)
    uint32_t nsums;           // slots to fill: odd sums first, first+2, ... or all intervals
    uint32_t *slots;
    uint32_t *counter;
    const uint32_t *coprime;  // kTabWords words: the residue tables
    const ffx_interval *iv;   // when niv > 0: slot g is sum iv[i].first + 2(g - iv[i].slotOffset)
    uint32_t niv;
    uint32_t grab;            // sums taken per atomic
    uint32_t eager;           // splits visited before queued candidates wait for full batches
    uint32_t bySum;           // slot of sum s is slots[(s - sumBase) / 2] instead of slot g
    uint32_t sumBase;
    uint64_t slotCap;         // bySum: slots outside [0, slotCap) are not written
    const uint32_t *sieveK;   // per sieve prime i and k < 32: (1-4^k)^-1 (0 if none) | (2^(k+1))^-1 << 16, mod q
    const uint16_t *sieveSqrt;  // per sieve prime, kSqrtStride entries: a square root of each residue, 0xFFFF for non-residues
};

// What the AllButOne check reads: the map and the smallest-factor table.  It
// is passed by value to the out-of-line check, so the kernel argument block
// itself never needs an address (an address would copy it to scratch).
struct TermTables {
    const uint8_t *map;
    uint64_t M;
    const uint16_t *spf;
    uint32_t spfLimit;
};

// Sum of slot g: one range, or the interval holding g (binary search; the
// wave's lanes agree on g, so the loads are uniform).
__device__ __forceinline__ uint32_t sum_of_slot(const SearchArgs &a, uint32_t g)
{
    if (a.niv == 0) return a.first + 2 * g;
    uint32_t lo = 0, hi = a.niv - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.iv[mid].slotOffset <= g) lo = mid;
        else hi = mid - 1;
    }
    return a.iv[lo].first + 2 * (uint32_t)(g - a.iv[lo].slotOffset);
}

// ---------------------------------------------------------------- primality

// a*b mod n for a, b < n < 2^50, all exact integers held in doubles: h is
// the rounded product and l = a*b - h exactly (one fused multiply-add); the
// quotient floor(h/n) from the reciprocal is off by at most one (relative
// error below 2^-51 on a quotient below 2^50), and h - q*n + l is an integer
// below 2^52 in magnitude, so every step is exact and one correction each
// way lands in [0, n).
__device__ __forceinline__ double mulmod_f64(double a, double b, double n, double ninv)
{
    const double h = a * b;
    const double l = fma(a, b, -h);
    const double q = floor(h * ninv);
    double r = fma(-q, n, h) + l;
    if (r < 0) r += n;
    if (r >= n) r -= n;
    return r;
}

// Deterministic Miller-Rabin with ff_seg's bases {2,3,5,7,11,13,17} and
// base count by magnitude (AskMillerRabin, segmentedSieve.C:364-387), for
// odd 17 <= n < 2^50, in double-precision arithmetic: a modular product is
// about ten fused floating-point operations.  Every value the search tests
// is below 2^50 for sums up to FFX_MAX_SUM (2^26 - 1): the probe values
// 2^k + d*o - 2 <= s^2/16 + 4 and d + o*2^k - 2 < s, the AllButOne terms
// f + P/f <= P/2 + 2 with P <= s^2/4, and s - 2, s/3 and the terms of s.
//
// For 2^20 <= n < 2^47 (every probe value of sums below 2^25) residues are
// kept balanced, |y| < 0.53 n, and a product is reduced by the nearest
// quotient: h = a*b rounded, l = a*b - h exactly (fma), q = rint(h/n) is
// within 0.52 of a*b/n (|a*b| < 0.29 n^2 and n < 2^47 put the reciprocal's
// error below 0.01), so h - q*n (one exact fma: an integer below 2^48) plus
// l (|l| <= |h| 2^-52 < 0.01 n) is again below 0.53 n -- six operations, no
// corrections.  The exponent is taken left to right, so the other factor
// of each step is the base itself: b*y < 17 * 0.53 n < 2^51 is exact and
// its reduction is a rint and one fma.  y = 1 and y = -1 are then the only
// residues of 1 and n - 1, so every test reads exactly as in the general
// path below.
__device__ __forceinline__ double mulmod_bal(double a, double b, double n, double ninv)
{
    const double h = a * b;
    const double l = fma(a, b, -h);
    return fma(-rint(h * ninv), n, h) + l;
}

__device__ __noinline__ bool mr_is_prime(uint64_t n)
{
    FFX_COUNT(8, 1);
    const double nd = (double)n, ninv = 1.0 / nd, mone = nd - 1.0;
    const uint32_t r = (uint32_t)__builtin_ctzll(n - 1);
    const uint64_t d = (n - 1) >> r;
    const int nb = n < 3215031751ull ? 4 : n < 2152302898747ull ? 5 : n < 3474749660383ull ? 6 : 7;
    const double bases[7] = {2, 3, 5, 7, 11, 13, 17};
    if (n >= (1ull << 20) && n < (1ull << 47)) {
        const int top = 63 - __builtin_clzll(d);  // d is odd: bit 0 set
        for (int bi = 0; bi < nb; ++bi) {
            const double b = bases[bi];
            double y = b;
            for (int i = top - 1; i >= 0; --i) {
                y = mulmod_bal(y, y, nd, ninv);
                if ((d >> i) & 1) {
                    const double h = y * b;  // exact
                    y = fma(-rint(h * ninv), nd, h);
                }
            }
            if (y == 1.0 || y == -1.0) continue;
            bool pass = false;
            for (uint32_t j = 1; j < r; ++j) {
                y = mulmod_bal(y, y, nd, ninv);
                if (y == -1.0) {
                    pass = true;
                    break;
                }
                if (y == 1.0) return false;
            }
            if (!pass) return false;
        }
        return true;
    }
    for (int bi = 0; bi < nb; ++bi) {
        double b = bases[bi], y = 1.0;
        for (uint64_t e = d; e; e >>= 1) {
            if (e & 1) y = mulmod_f64(y, b, nd, ninv);
            b = mulmod_f64(b, b, nd, ninv);
        }
        if (y == 1.0 || y == mone) continue;
        bool pass = false;
        for (uint32_t j = 1; j < r; ++j) {
            y = mulmod_f64(y, y, nd, ninv);
            if (y == mone) {
                pass = true;
                break;
            }
            if (y == 1.0) return false;
        }
        if (!pass) return false;
    }
    return true;
}

__device__ __forceinline__ bool byte_bit(uint32_t byte, uint64_t n) { return (byte >> (7 - ((n >> 1) & 7))) & 1; }

// Map bytes through a global-address-space pointer: global_load, not flat.
__device__ __forceinline__ uint32_t map_byte(const uint8_t *map, uint64_t n)
{
    return ((const __attribute__((address_space(1))) uint8_t *)map)[n >> 4];
}

__device__ __forceinline__ bool map_bit(const uint8_t *map, uint64_t n) { return byte_bit(map_byte(map, n), n); }

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

__device__ __forceinline__ bool tab_bit(const uint32_t *t, uint32_t r) { return (t[r >> 5] >> (r & 31)) & 1; }

// n mod m for n < 2^28 with one high multiply: c = ceil(2^(32+sh) / m) and
// e = c*m - 2^(32+sh) < 2^(sh+4) make floor(n*c / 2^(32+sh)) = floor(n/m)
// for every n < 2^28 (n*e < 2^(32+sh)).  The general 32-bit division by a
// constant needs three more instructions.
constexpr int floor_log2(uint32_t m) { return m <= 1 ? 0 : 1 + floor_log2(m >> 1); }

template <uint32_t m>
struct Mod28 {
    static constexpr int sh = floor_log2(m);  // c < 2^32 for m not a power of two
    static constexpr uint64_t c = ((1ull << (32 + sh)) + m - 1) / m;
    static_assert(c < (1ull << 32), "multiplier must fit 32 bits");
    static_assert(c * m - (1ull << (32 + sh)) < (1ull << (sh + 4)), "quotient must be exact below 2^28");
};

template <uint32_t m>
__device__ __forceinline__ uint32_t mod28(uint32_t n)
{
    const uint32_t q = __umulhi(n, (uint32_t)Mod28<m>::c) >> Mod28<m>::sh;
    return n - q * m;
}

// Coprimality to m of X = 2^k + d*o - 2 and Y = d + o*2^k - 2 (pk = 2^k),
// from residues; d, o, pk < 2^26 and m^2 + 2m < 2^28 keep every reduced
// value below 2^28.
template <uint32_t m>
__device__ __forceinline__ void residue_pair(const uint32_t *t, uint32_t d, uint32_t o, uint32_t pm, bool &rx, bool &ry,
                                             bool &ro)
{
    static_assert((uint64_t)m * m + 2 * m < (1ull << 28), "residue products must stay below 2^28");
    const uint32_t dm = mod28<m>(d), om = mod28<m>(o);
    rx = rx & tab_bit(t, mod28<m>(dm * om + pm + m - 2));
    ry = ry & tab_bit(t, mod28<m>(dm + om * pm + m - 2));
    ro = ro & tab_bit(t, om);
}

// 2^k mod the first two moduli, packed (both below 2^16); the split
// segment table carries it per k.
__device__ __forceinline__ uint32_t pow2_residues(uint32_t k)
{
    const uint32_t pk = 1u << k;
    uint32_t r = mod28<kMods[0]>(pk);
    if constexpr (kNMods > 1) r |= mod28<kMods[1]>(pk) << 16;
    return r;
}

template <int I>
__device__ __forceinline__ void residue_all(const uint32_t *tab, uint32_t d, uint32_t o, uint32_t k, uint32_t pms,
                                            bool &rx, bool &ry, bool &ro)
{
    if constexpr (I < kNMods) {
        const uint32_t pm = I == 0 ? (pms & 0xFFFF) : I == 1 ? (pms >> 16) : mod28<kMods[I]>(1u << k);
        residue_pair<kMods[I]>(tab + tab_offset(I), d, o, pm, rx, ry, ro);
        residue_all<I + 1>(tab, d, o, k, pms, rx, ry, ro);
    }
}

// rx (ry, ro) false when X (Y, o) is composite by a prime factor of the
// moduli (a multiple of that prime other than the prime itself).  pms = pow2_residues(k).
__device__ __forceinline__ void probe_residues(const uint32_t *tab, uint32_t d, uint32_t o, uint32_t k, uint32_t pms,
                                               bool &rx, bool &ry, bool &ro)
{
    rx = ry = ro = true;
    residue_all<0>(tab, d, o, k, pms, rx, ry, ro);
}

// n mod m for n < 2^52, from two 26-bit halves.
template <uint32_t m>
__device__ __forceinline__ uint32_t mod52(uint64_t n)
{
    constexpr uint32_t r26 = (1u << 26) % m;
    return mod28<m>(mod28<m>((uint32_t)(n >> 26)) * r26 + mod28<m>((uint32_t)n & ((1u << 26) - 1)));
}

// Coprimality to 3*5*7*11*13 (the first residue table) of the two terms of
// the divisor pairs (f, g 2^k) and (f 2^k, g) of q 2^k, g = q/f < 2^52,
// f < 2^26, k <= 25: ra for f + g 2^k - 2, rb for f 2^k + g - 2.  (Also
// testing the other two moduli saves a few more reads but costs more issue:
// measured slower.)
__device__ __forceinline__ void pair_residues(const uint32_t *tab, uint32_t f, uint64_t g, uint32_t k, bool &ra,
                                              bool &rb)
{
    constexpr uint32_t m = kMods[0];
    const uint32_t pk = mod28<m>(1u << k), fm = mod28<m>(f), gm = mod52<m>(g);
    ra = tab_bit(tab, mod28<m>(fm + gm * pk + m - 2));
    rb = tab_bit(tab, mod28<m>(fm * pk + gm + m - 2));
}

// ---------------------------------------------------------------- AB1

// Adds the factorization of odd n (< spfLimit) to the prime list pr[0..nf)
// (LDS, this lane's row) and the 5-bit exponent fields of ex.  Quotients by
// the factor p are exact, so they are products with p^-1 mod 2^32, and
// p | n  <=>  n * p^-1 mod 2^32 <= (2^32-1)/p.
__device__ __forceinline__ void factor_into(const TermTables &a, uint32_t n, uint32_t *pr, uint64_t &ex, int &nf)
{
    while (n > 1) {
        uint32_t p;
        if (n < a.spfLimit) {
            p = a.spf[n >> 1];
            if (p == 0) p = n;
        } else {  // beyond the table (never for terms of a sum the plan covers)
            p = n;
            for (uint32_t f = 3; f * f <= n; f += 2)
                if (n % f == 0) {
                    p = f;
                    break;
                }
        }
        uint32_t inv = p;  // p * p == 1 mod 8
#pragma unroll
        for (int i = 0; i < 4; ++i) inv *= 2u - p * inv;
        const uint32_t lim = 0xFFFFFFFFu / p;
        uint32_t e = 0;
        do {
            n *= inv;
            ++e;
        } while (n * inv <= lim);
        int i = 0;
        while (i < nf && pr[i] != p) ++i;
        if (i == nf) {
            pr[nf] = p;
            ++nf;
        }
        ex += (uint64_t)e << (5 * i);
    }
}

// AllButOneProductOfTermPairsOfSumOfFactorPairsHasSingleFactorPair(k, d*o)
// (segmentedSieve.C:424-477) for the split s = d*2^k + o.  The split's own
// factor pair (d*2^k, o) is always among the pairs counted and always has
// PTP(s) false -- that is statement 1, which every searched sum satisfies --
// so the reference's "exactly one" is "no other pair has PTP false": the
// enumeration stops at the first other one.  q = d*o is factored from d and
// o; divisors come from an odometer over exponent fields; the cofactor q/dv
// is an exact double division (q < 2^53).  The f = 2^k pair is PTP(2^k + q),
// already seen true by the caller for d > 1, and is the own pair for d = 1.
// Terms beyond the map wait for Miller-Rabin until the map terms are done.
__device__ __noinline__ bool ab1(const TermTables a, uint32_t k, uint32_t d, uint32_t o, uint32_t *pr)
{
    const uint64_t q = (uint64_t)d * o;
    const uint64_t own = ((uint64_t)d << k) + o - 2;  // s - 2
    const uint64_t P = q << k;
    const uint64_t L = isqrt53(P);
    const uint64_t Lk = L >> k;
    uint64_t late0 = 0, late1 = 0;
#define FFX_TERM(X)                                     \
    do {                                                \
        const uint64_t x_ = (X);                        \
        if (x_ != own) {                                \
            if (x_ <= a.M) {                            \
                FFX_COUNT(7, 1);                        \
                if (!map_bit(a.map, x_)) return false;  \
            } else if (!late0) {                        \
                late0 = x_;                             \
            } else if (!late1) {                        \
                late1 = x_;                             \
            } else if (!mr_is_prime(x_)) {              \
                return false;                           \
            }                                           \
        }                                               \
    } while (0)
    uint64_t ex = 0;
    int nf = 0;
    factor_into(a, d, pr, ex, nf);
    factor_into(a, o, pr, ex, nf);
    const double qd = (double)q;
    uint64_t dig = 0, dv = 1;
    for (;;) {
        int j = 0;
        for (; j < nf; ++j) {
            const uint32_t sh = 5 * j;
            const uint32_t e = (uint32_t)(dig >> sh) & 31, x = (uint32_t)(ex >> sh) & 31;
            const uint32_t p = pr[j];
            if (e < x) {
                dig += 1ull << sh;
                dv *= p;
                break;
            }
            uint64_t pe = 1;  // digit j wraps: divide p^x back out
            for (uint32_t t = 0; t < x; ++t) pe *= p;
            dv = (uint64_t)((double)dv / (double)pe);
            dig &= ~(31ull << sh);
        }
        if (j == nf) break;
        const uint64_t c = (uint64_t)(qd / (double)dv);
        if (dv <= L) FFX_TERM(dv + (c << k) - 2);
        if (dv <= Lk) FFX_TERM((dv << k) + c - 2);
    }
#undef FFX_TERM
    if (late0 && !mr_is_prime(late0)) return false;
    if (late1 && !mr_is_prime(late1)) return false;
    return true;
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
__device__ __forceinline__ void drain(const SearchArgs &a, const uint32_t *tab, const uint2 *queue, int from, int nb,
                                      int lane, uint32_t *pr, int &cnt, uint32_t &te, uint32_t &to)
{
    bool valid = false;
    uint32_t ee = 0, oo = 0;
    if (lane < nb) {
        const uint2 ent = queue[from + lane];
        const uint32_t d = ent.x & ((1u << kDBits) - 1), k = ent.x >> kDBits;
        oo = ent.y;
        ee = d << k;
        // Quick rejection before the full check: for p the smallest prime
        // factor of d and p' that of o (of o alone when d = 1), AllButOne
        // visits the pairs (p, (q/p) 2^k) and (p 2^k, q/p) of q = d*o -- each
        // when its smaller member is at most isqrt(q 2^k) -- and the same for
        // p'.  Their terms p + (q/p) 2^k - 2 and p 2^k + q/p - 2 are first
        // tested for a factor in 3..13, then read in three rounds (below); a
        // composite one (not the split's own term) makes AllButOne false, so
        // only splits that survive pay for factoring d and o.
        const uint32_t n1 = d > 1 ? d : oo, n2 = d > 1 && oo > 1 ? oo : 1u;
        uint32_t p = n1 < a.spfLimit ? a.spf[n1 >> 1] : 0u;
        uint32_t p2 = n2 > 1 && n2 < a.spfLimit ? a.spf[n2 >> 1] : 0u;
        if (p == 0) p = n1;
        if (p2 == 0) p2 = n2;
        bool maybe = true;
        if (p > 1) {
            const uint64_t q = (uint64_t)d * oo, P = q << k, L = isqrt53(P);
            const uint64_t own = ((uint64_t)d << k) + oo - 2;
            const bool two = p2 > 1 && p2 != p;
            const uint64_t c1 = (uint64_t)((double)q / (double)p);  // exact: p | q < 2^53
            const uint64_t c2 = two ? (uint64_t)((double)q / (double)p2) : 1;
            const uint64_t t[4] = {(uint64_t)p + (c1 << k) - 2, ((uint64_t)p << k) + c1 - 2,
                                   (uint64_t)p2 + (c2 << k) - 2, ((uint64_t)p2 << k) + c2 - 2};
            const bool u[4] = {p <= L, p <= (L >> k), two && p2 <= L, two && p2 <= (L >> k)};
            // one term of p's pair, then its other term, then p''s pair, each
            // read only while the split is still in (fewer reads: the kernel
            // is bound by them)
            // a term with a clear residue bit is composite: rejected before
            // any read (C2: 104 M of these reads per launch down to 19 M,
            // 11.96 -> 11.88 ms)
            bool r[4] = {true, true, true, true};
            pair_residues(tab, p, c1, k, r[0], r[1]);
            if (two) pair_residues(tab, p2, c2, k, r[2], r[3]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (u[i] && t[i] != own && !r[i]) maybe = false;
            uint32_t b[4];
            bool use[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) use[i] = maybe && u[i] && t[i] != own && t[i] <= a.M;
            b[0] = use[0] ? map_byte(a.map, t[0]) : 0u;
            maybe = maybe && (!use[0] || byte_bit(b[0], t[0]));
            use[1] = use[1] && maybe;
            b[1] = use[1] ? map_byte(a.map, t[1]) : 0u;
            maybe = maybe && (!use[1] || byte_bit(b[1], t[1]));
#pragma unroll
            for (int i = 2; i < 4; ++i) {
                use[i] = use[i] && maybe;
                b[i] = use[i] ? map_byte(a.map, t[i]) : 0u;
            }
#pragma unroll
            for (int i = 2; i < 4; ++i) maybe = maybe && (!use[i] || byte_bit(b[i], t[i]));
#ifdef FFX_PROBE_COUNTERS
            FFX_COUNT(7, (int)use[0] + (int)use[1] + (int)use[2] + (int)use[3]);
#endif
        }
        valid = maybe && ab1(TermTables{a.map, a.M, a.spf, a.spfLimit}, k, d, oo, pr);
    }
#ifdef FFX_PROBE_COUNTERS
    if (lane == 0) FFX_COUNT(6, nb);
#endif
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

// Map probes of queued splits (entries [from, from+nb) of the wave's probe
// queue, two per lane so their random map reads are in flight together):
// the prime-o condition of Power2Even (:619), PTP(2^k + q) and PTP(d + o*2^k)
// of DoesPeterKnow (:632-639).  Survivors go to the AllButOne queue.
__device__ __forceinline__ void probe_batch(const SearchArgs &a, const uint2 *pq, int from, int nb, int lane,
                                            uint64_t lt, uint2 *queue, int &qn)
{
    uint32_t d[2], k[2], o[2], ob[2], oo[2];
    bool st[2], needO[2];
    uint64_t x[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int idx = 64 * u + lane;
        st[u] = idx < nb;
        const uint2 e = st[u] ? pq[from + idx] : make_uint2(3u | (2u << kDBits), 1u);
        d[u] = e.x & ((1u << kDBits) - 1);
        k[u] = e.x >> kDBits;
        o[u] = e.y & ~kNeedO;
        needO[u] = st[u] && (e.y & kNeedO);
        x[u] = (1ull << k[u]) + (uint64_t)d[u] * o[u] - 2;
    }
    // o and 2^k + q - 2 read together
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        oo[u] = (needO[u] && o[u] <= a.M) ? map_byte(a.map, o[u]) : 0u;
        ob[u] = (st[u] && x[u] <= a.M) ? map_byte(a.map, x[u]) : 0u;
#ifdef FFX_PROBE_COUNTERS
        const uint64_t b1 = __ballot(needO[u] && o[u] <= a.M), b2 = __ballot(st[u] && x[u] <= a.M);
        if (lane == 0) {
            FFX_COUNT(5, __popcll(b1));
            FFX_COUNT(3, __popcll(b2));
        }
#endif
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        if (needO[u]) st[u] = o[u] <= a.M ? byte_bit(oo[u], o[u]) : mr_is_prime(o[u]);
        if (st[u]) st[u] = x[u] <= a.M ? byte_bit(ob[u], x[u]) : mr_is_prime(x[u]);
    }
    // d + o*2^k - 2 unless d == o
    bool need2[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        need2[u] = st[u] && d[u] != o[u];
        x[u] = (uint64_t)d[u] + ((uint64_t)o[u] << k[u]) - 2;
        ob[u] = (need2[u] && x[u] <= a.M) ? map_byte(a.map, x[u]) : 0u;
#ifdef FFX_PROBE_COUNTERS
        const uint64_t b3 = __ballot(need2[u] && x[u] <= a.M);
        if (lane == 0) FFX_COUNT(4, __popcll(b3));
#endif
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
        if (need2[u]) st[u] = x[u] <= a.M ? byte_bit(ob[u], x[u]) : mr_is_prime(x[u]);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const uint64_t cm = __ballot(st[u]);
        if (st[u]) queue[qn + __popcll(cm & lt)] = make_uint2(d[u] | (k[u] << kDBits), o[u]);
        qn += __popcll(cm);
    }
}

// The sieve over d of one sum.  An item is (prime q, value r): r = 0 the
// root of the linear value d + o*2^k - 2, r = 1, 2 the roots
// (sm +- sqrt(D)) / 2^(k+1) of the quadratic 2^k d^2 - s d - (2^k - 2) = 0
// (2^k + d*o - 2 = 0 mod q) with D = s^2 + 4 * 2^k (2^k - 2).  Lane 3i+r
// holds struck item (q_{kNSmallSieve+i}, r) in a SieveItem, and lanes below
// 3 kNSmallSieve also hold masked item (q_i, r) packed in one register
// (root | kind << 8 | s mod q << 16 | q << 24).  s mod q is taken once per sum, the
// roots once per k-segment; each chunk then only places them relative to
// the chunk's first d.  Fields are bytes, packed two registers to a lane
// (the search runs at its register limit).
struct SieveItem {
    uint32_t qhs = 1;  // q | (kSieveBits/2 mod q) << 8 | (s mod q) << 16
    uint32_t rk = 0;   // root | kind << 8: d = root (mod q) makes the value divisible by q;
                       // kind 0 no root, 1 root, 2 every d (4^k = 1 and s 2^k = 2 mod q)
    __device__ uint32_t q() const { return qhs & 255; }
    __device__ uint32_t half() const { return (qhs >> 8) & 255; }
    __device__ uint32_t sm() const { return qhs >> 16; }
    __device__ uint32_t root() const { return rk & 255; }
    __device__ uint32_t kind() const { return rk >> 8; }
};

__device__ __forceinline__ void sieve_item_for_sum(SieveItem &it, uint32_t &small, uint32_t &far, uint32_t s, int lane)
{
    // far item of lanes below 3 kNFarSieve: q | s mod q << 8 | root << 16 | kind << 24
    if (lane < 3 * kNFarSieve) {
        const uint32_t q = kSievePrimes[kNSmallSieve + kNLargeSieve + lane / 3];
        far = q | (s % q) << 8;
    } else {
        far = 1u;
    }
    if (lane < 3 * kNLargeSieve) {
        const uint32_t q = kSievePrimes[kNSmallSieve + lane / 3];
        it.qhs = q | ((kSieveBits / 2) % q) << 8 | (s % q) << 16;
    }
    if (lane < 3 * kNSmallSieve) {
        const uint32_t q = kSievePrimes[lane / 3];
        small = (s % q) << 16 | q << 24;
    } else {
        small = 1u << 24;  // q = 1: no item
    }
}

// (root, kind) of an item for 2^k = pk (mod q), from its table record and,
// for the quadratic's roots, the square root r of its discriminant
__device__ __forceinline__ uint2 item_root(uint32_t q, uint32_t sm, uint32_t pk, uint32_t rec, uint32_t r, int which)
{
    uint2 it = make_uint2(0u, 0u);
    const uint32_t iY = rec & 0xFFFFu, iX = rec >> 16;  // iY = 0: 1 - 4^k = 0 mod q
    if (which == 0) {
        const uint32_t rhs = (2 + q - (sm * pk) % q) % q;  // d (1 - 4^k) = 2 - s 2^k
        if (iY) it = make_uint2((rhs * iY) % q, 1u);
        else if (rhs == 0) it.y = 2;  // 4^k = 1 mod q: the value is s 2^k - 2 for every d
    } else if (r != 0xFFFFu) {
        it = make_uint2((((which == 1 ? sm + r : sm + q - r) % q) * iX) % q, 1u);
    }
    return it;
}

// Roots of both of the lane's items for 2^k: x for its struck item (prime
// q, s mod q = sm), y for its masked item (packed in small, q = 1 when the
// lane has none); root | kind << 8 each.  2^k mod q is computed, so every
// table read (record, square root) is issued before any is used: one memory
// latency per k-segment instead of four.
__device__ __noinline__ uint3 sieve_items_for_k(uint32_t q, uint32_t sm, uint32_t small, uint32_t far,
                                                const uint32_t *sieveK, const uint16_t *sieveSqrt, uint32_t k, int lane)
{
    const int which = lane % 3;
    const bool vS = lane < 3 * kNLargeSieve, vM = lane < 3 * kNSmallSieve, vF = lane < 3 * kNFarSieve;
    const int piS = kNSmallSieve + lane / 3, piM = lane / 3, piF = kNSmallSieve + kNLargeSieve + lane / 3;
    const uint32_t qM = small >> 24, smM = (small >> 16) & 255;
    const uint32_t qF = far & 255, smF = (far >> 8) & 255;
    const uint32_t pkS = (1u << k) % q, pkM = (1u << k) % qM, pkF = (1u << k) % qF;
    const uint32_t DS = (sm * sm + 4 * pkS * ((pkS + q - 2) % q)) % q;
    const uint32_t DM = (smM * smM + 4 * pkM * ((pkM + qM - 2) % qM)) % qM;
    const uint32_t DF = (smF * smF + 4 * pkF * ((pkF + qF - 2) % qF)) % qF;
    const uint32_t recS = vS ? sieveK[piS * 32 + k] : 0u, recM = vM ? sieveK[piM * 32 + k] : 0u;
    const uint32_t recF = vF ? sieveK[piF * 32 + k] : 0u;
    const uint32_t rS = vS && which ? sieveSqrt[piS * kSqrtStride + DS] : 0xFFFFu;
    const uint32_t rM = vM && which ? sieveSqrt[piM * kSqrtStride + DM] : 0xFFFFu;
    const uint32_t rF = vF && which ? sieveSqrt[piF * kSqrtStride + DF] : 0xFFFFu;
    const uint2 a = vS ? item_root(q, sm, pkS, recS, rS, which) : make_uint2(0u, 0u);
    const uint2 b = vM ? item_root(qM, smM, pkM, recM, rM, which) : make_uint2(0u, 0u);
    const uint2 c = vF ? item_root(qF, smF, pkF, recF, rF, which) : make_uint2(0u, 0u);
    return make_uint3(a.x | a.y << 8, b.x | b.y << 8, c.x | c.y << 8);
}

using lds_u32 = __attribute__((address_space(3))) uint32_t;

// Far-2 items: per sum, q and s mod q of each lane's item in every set
// (s = 0: none, every item marks nothing).
__device__ __forceinline__ void far2_for_sum(lds_u32 *far2, uint32_t s, int lane)
{
    for (int t = 0; t < kNFar2Sets; ++t) {
        uint32_t v = 0;
        if (lane < 63 && s) {
            const uint32_t q = kSievePrimes[kFar2First + 21 * t + lane / 3];
            v = q | (s % q) << 10;
        }
        far2[64 * t + lane] = v;
    }
}

// Far-2 items for 2^k: each item's root, placed as the first bit of the
// chunk starting at odd d = dLo; every table read of all sets is issued
// before any is used.
__device__ __noinline__ void far2_for_k(lds_u32 *far2, const uint32_t *sieveK, const uint16_t *sieveSqrt, uint32_t k,
                                        uint32_t dLo, int lane)
{
    if (lane >= 63) return;
    const int which = lane % 3;
    uint32_t v[kNFar2Sets], rec[kNFar2Sets], r[kNFar2Sets], pk[kNFar2Sets];
#pragma unroll
    for (int t = 0; t < kNFar2Sets; ++t) {
        v[t] = far2[64 * t + lane];
        const uint32_t q = v[t] & 1023, sm = (v[t] >> 10) & 1023, pi = kFar2First + 21 * t + lane / 3;
        pk[t] = (1u << k) % q;
        const uint32_t D = (sm * sm + 4 * pk[t] * ((pk[t] + q - 2) % q)) % q;
        rec[t] = sieveK[pi * 32 + k];
        r[t] = which ? sieveSqrt[pi * kSqrtStride + D] : 0xFFFFu;
    }
#pragma unroll
    for (int t = 0; t < kNFar2Sets; ++t) {
        const uint32_t q = v[t] & 1023, sm = (v[t] >> 10) & 1023;
        const uint2 it = item_root(q, sm, pk[t], rec[t], r[t], which);
        uint32_t b0 = 0;
        if (it.y == 1) {
            uint32_t x = it.x + q - dLo % q;
            if (x >= q) x -= q;
            if (x & 1) x += q;
            b0 = x >> 1;
        }
        far2[64 * t + lane] = (v[t] & 0xFFFFFu) | b0 << 20 | it.y << 30;
    }
}

// Bits 0, q, 2q, ... of a 64-bit word.
constexpr uint64_t every_qth_bit(uint32_t q)
{
    uint64_t m = 0;
    for (uint32_t b = 0; b < 64; b += q) m |= 1ull << b;
    return m;
}

// Marks, in bits (kSieveBits, bit b = odd d = dLo + 2b), the d of the
// k-segment's splits whose 2^k + d*o - 2 or d + o*2^k - 2 has a sieve prime
// factor (o = s - d*2^k).  Lane l first writes bits 64l..64l+63 as the
// union of the masked items' residue classes (the items' first bits taken
// from the lanes that hold them, one readlane each), then the struck items
// are OR-ed in: items of small primes mark many more bits than those of
// large ones, so lane l strikes the first half of its own item's bits and
// the second half of item 62-l's, and every lane carries a small and a
// large prime.
__device__ __noinline__ void sieve_chunk(const SieveItem it, uint32_t small, uint32_t far, lds_u32 *far2, uint32_t dLo,
                                        lds_u32 *bits, int lane)
{
    static_assert(3 * kNLargeSieve <= 63, "struck items fit lanes 0..62, paired l <-> 62-l");
    static_assert(kSieveBits == 64 * 64, "one 64-bit word of masked items per lane");
    constexpr uint32_t kHalf = kSieveBits / 2;
    {
        // masked item of this lane: its first bit b0 < q (odd d, dLo odd)
        uint32_t place = 0;  // b0 | kind << 8
        if (lane < 3 * kNSmallSieve) {
            const uint32_t q = small >> 24, kind = (small >> 8) & 3;
            uint32_t b0 = 0;
            if (kind == 1) {
                uint32_t t = (small & 255) + q - dLo % q;
                if (t >= q) t -= q;
                if (t & 1) t += q;
                b0 = t >> 1;
            }
            place = b0 | kind << 8;
        }
        uint64_t m = 0;
#pragma unroll
        for (int i = 0; i < 3 * kNSmallSieve; ++i) {
            const uint32_t q = kSievePrimes[i / 3];  // a constant once unrolled
            const uint32_t v = __builtin_amdgcn_readlane(place, i);  // uniform
            const uint32_t r = (64u * (uint32_t)lane) % q;           // first bit of this lane's word, mod q
            const uint32_t u = (v & 255) - r, ph = min(u, u + q);   // (b0 - r) mod q
            m |= ((v >> 8) == 1 ? every_qth_bit(q) : 0ull) << ph;
        }
        if (__ballot((place >> 8) == 2)) m = ~0ull;  // an item that marks every d (rare)
        bits[2 * lane] = (uint32_t)m;
        bits[2 * lane + 1] = (uint32_t)(m >> 32);
    }
    // own struck item: first bit b0 < q with dLo + 2 b0 = root (mod q), odd d
    uint32_t b0 = 0;
    const uint32_t q = it.q(), kind = it.kind();
    if (kind == 1) {
        uint32_t t = it.root() + q - dLo % q;  // d - dLo = t (mod q), t in [1, 2q)
        if (t >= q) t -= q;
        if (t & 1) t += q;  // odd d (dLo is odd): an even offset
        b0 = t >> 1;
    }
    // the partner item's second half
    const int partner = 62 - lane;
    const uint32_t qhsB = __shfl(it.qhs, partner), qB = qhsB & 255, hB = (qhsB >> 8) & 255,
                   kindB = __shfl(it.rk, partner) >> 8;
    const uint32_t b0B = __shfl(b0, partner);
    wave_sync_lds();
    if (lane < 63) {
        if (kind == 2) {
            for (uint32_t i = 0; i < kHalf / 32; ++i)
                __hip_atomic_fetch_or(&bits[i], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        } else if (kind == 1) {
            for (uint32_t b = b0; b < kHalf; b += q)
                __hip_atomic_fetch_or(&bits[b >> 5], 1u << (b & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        if (kindB == 2) {
            for (uint32_t i = kHalf / 32; i < kSieveBits / 32; ++i)
                __hip_atomic_fetch_or(&bits[i], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        } else if (kindB == 1) {
            // first b >= kHalf with b = b0B (mod qB): kHalf + ((b0B - kHalf) mod qB)
            uint32_t b = kHalf + (b0B >= hB ? b0B - hB : b0B + qB - hB);
            for (; b < kSieveBits; b += qB)
                __hip_atomic_fetch_or(&bits[b >> 5], 1u << (b & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    // far item: all its bits (at most kSieveBits / 173 + 1 = 24)
    const uint32_t qF = far & 255, kindF = far >> 24;
    if (kindF == 2) {
        for (uint32_t i = 0; i < kSieveBits / 32; ++i)
            __hip_atomic_fetch_or(&bits[i], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    } else if (kindF == 1) {
        uint32_t t = ((far >> 16) & 255) + qF - dLo % qF;
        if (t >= qF) t -= qF;
        if (t & 1) t += qF;
        for (uint32_t b = t >> 1; b < kSieveBits; b += qF)
            __hip_atomic_fetch_or(&bits[b >> 5], 1u << (b & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    // far-2 items: from the bit the previous chunk (or far2_for_k) left,
    // and the bit past this chunk kept for the next one
    if (lane < 63) {
#pragma unroll
        for (int t = 0; t < kNFar2Sets; ++t) {
            const uint32_t v = far2[64 * t + lane], q2 = v & 1023, kind2 = v >> 30;
            if (kind2 == 2) {
                for (uint32_t i = 0; i < kSieveBits / 32; ++i)
                    __hip_atomic_fetch_or(&bits[i], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            } else if (kind2 == 1) {
                uint32_t b = (v >> 20) & 1023;
                for (; b < kSieveBits; b += q2)
                    __hip_atomic_fetch_or(&bits[b >> 5], 1u << (b & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                far2[64 * t + lane] = (v & 0xC00FFFFFu) | (b - kSieveBits) << 20;
            }
        }
    }
    wave_sync_lds();
}

__device__ __forceinline__ uint32_t solve_sum_wave(const SearchArgs &a, const uint32_t *tab, uint32_t s, int lane,
                                                   uint2 *queue, uint2 *pq, uint4 *segs, uint8_t *segOffs,
                                                   uint32_t *pr, uint32_t *dbits, uint32_t *far2)
{
    // statement 1 (RunIt :806-807)
    if (is_prime_odd(a, s - 2)) return 0;
    if (s % 3 == 0 && !is_prime_odd(a, s / 3)) return 0;

    int cnt = 0, qn = 0;
    uint32_t te = 0, to = 0;
    const uint64_t lt = lanes_below(lane);

    // Power2Prime (:547-564): lane l tests k = l + 2; composite odd terms
    // become Power2Composite candidates (:591-605)
    {
        const uint32_t k = 2 + lane;
        const bool act = k < 32 && (1u << k) < s - 2;
        const bool isp = act && is_prime_odd(a, s - (1u << k));
        const uint64_t pm = __ballot(isp), cm = __ballot(act && !isp);
        cnt = __popcll(pm);
        if (cnt > 1) return 0;
        if (cnt == 1) {
            te = 1u << (2 + __ffsll((unsigned long long)pm) - 1);
            to = s - te;
        }
        if (act && !isp) queue[__popcll(cm & lt)] = make_uint2(1u | (k << kDBits), s - (1u << k));
        qn = __popcll(cm);
    }

    // Power2Odd and Power2Even splits, k descending as the reference visits
    // them (:570-572, :611-613), both families interleaved by k, flattened
    // into one index space: segment j is k = kTop - j with its d sequence.
    // Each segment lists only the residue classes of d mod 30 that can hold
    // a valid split, decided from d, o = s - d*2^k and 2^k modulo 15:
    //  * Power2Odd (k odd) needs 3 !| d and 3 !| o (its family, :578);
    //  * both probe values X = 2^k + d*o - 2 and Y = d + o*2^k - 2 must be
    //    prime, so neither may be divisible by 3 or 5 -- for s >= 4096 both
    //    exceed 5 (see tinySum), and Y is not tested for the one split with
    //    d = o, whose class is kept whole.
    // That drops ~60% of the splits before any per-split work.
    // For s >= 4096 every split has X, Y > 47: if o >= 16, d*o >= 48 and
    // o*2^k >= 64; otherwise d*2^k > 4080, so 2^k >= 64 or d >= 64.
    const bool tinySum = s < 4096;
    const uint32_t t3 = (s - 3) / 3;
    const int kTop = t3 >= 4 ? 31 - __clz(t3) : 0;  // largest k with 3*2^k <= s-3
    const int nseg = kTop >= 2 ? kTop - 1 : 0;
    uint32_t total = 0;
    {
        constexpr uint32_t kCoprime15 = 0x6996u;  // r mod 15 coprime to 15: 1,2,4,7,8,11,13,14
        constexpr uint32_t kNot3 = 0x6db6u;       // r mod 15 not divisible by 3
        uint32_t n = 0, kj = 0, c = 0, skip = 0;
        if (lane < nseg) {
            kj = (uint32_t)(kTop - lane);
            const uint32_t D = (s - 3) >> kj;  // d*2^k < s-2  <=>  d <= (s-3)>>k
            const uint32_t dmin = (kj & 1) ? 5 : 3;
            const uint32_t pk = (1u << kj) % 15, s15 = s % 15;
            int dstarT = -1;  // class of the d = o split, if any
            const uint32_t m = (1u << kj) + 1;
            if (s % m == 0) {
                const uint32_t ds = s / m;
                if ((ds & 1) && ds >= dmin && ds <= D) dstarT = (int)((ds % 30) >> 1);
            }
            // whole periods of 30 in [0, D] and the last partial one: a class
            // with offset off holds P + (off <= rem) values of d <= D
            const uint32_t P = D / 30, rem = D - 30 * P;
            uint32_t le = 0;
#pragma unroll
            for (uint32_t t = 0; t < 15; ++t) {
                const uint32_t dr = (2 * t + 1) % 15;
                const uint32_t orr = (s15 + 225 - dr * pk) % 15;
                const uint32_t xr = (pk + dr * orr + 13) % 15, yr = (dr + orr * pk + 13) % 15;
                bool ok = !(kj & 1) || (((kNot3 >> dr) & (kNot3 >> orr)) & 1);
                if (!tinySum) ok = ok && ((kCoprime15 >> xr) & 1) && (((kCoprime15 >> yr) & 1) || (int)t == dstarT);
                if (ok) {
                    // bit 7: a Power2Even split of this class has 3 !| d*o,
                    // so o must be prime (:619); d = 2t+1 and o = orr mod 3
                    const bool needO = !(kj & 1) && dr % 3 != 0 && orr % 3 != 0;
                    segOffs[lane * 16 + c] = (uint8_t)((2 * t + 1) | (needO ? 0x80u : 0u));
                    ++c;
                    skip += 2 * t + 1 < dmin;
                    le += 2 * t + 1 <= rem;
                }
            }
            n = (c && D >= dmin) ? c * P + le - skip : 0;
        }
        uint32_t incl = n;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        // index i of the segment: ii = i + skip, d = 30*(ii / c) + offs[ii % c]
        // (ii < 2^24; for 2 <= c <= 15 the quotient is one high multiply by
        // ceil(2^32 / c), exact since ii * (magic*c - 2^32) < 2^32)
        const uint32_t magic = c >= 2 ? (uint32_t)(((1ull << 32) + c - 1) / c) : 0u;
        if (lane < nseg) segs[lane] = make_uint4(incl - n, kj | (c << 8) | (skip << 16), pow2_residues(kj), magic);
        total = nseg ? __shfl(incl, nseg - 1) : 0;
    }
    wave_sync_lds();

    // Each round: 64 * kRoundSplits splits through the filters that need no
    // map read -- the family test of Power2Odd/Power2Even (:578, :619) and
    // the residue filter of both probe values (and of o where Power2Even
    // needs it prime); the few survivors are compacted into the probe queue
    // and probed 128 at a time, so the map reads run on full waves.  While
    // the sum is young, queues are emptied every round (early exit at two
    // valid splits).
    int hint = 0, pn = 0;
#ifdef FFX_PROBE_COUNTERS
    if (lane == 0) {
        FFX_COUNT(0, 1);
        FFX_COUNT(9, total);
    }
    // splits visited, by the sum's outcome: 10 one split (a solution), 11 an
    // early exit at two, 12 none
    uint32_t visitedSplits = 0;
#define FFX_COUNT_EXIT(i) \
    do {                  \
        if (lane == 0) FFX_COUNT(i, visitedSplits); \
    } while (0)
#else
#define FFX_COUNT_EXIT(i) ((void)0)
#endif
    // the d sieve's items, their roots taken for segment sj
    int sj = -1;
    const bool sieveSum = s >= kSieveMinSum;
    SieveItem item;
    uint32_t small = 0, far = 1;
    const bool far2On = s >= kFar2MinSum;  // uniform
    if (sieveSum) {
        sieve_item_for_sum(item, small, far, s, lane);
        far2_for_sum((lds_u32 *)far2, far2On ? s : 0u, lane);
    }
    // Probes queued splits until fewer than 128 wait (all of them when
    // flush), draining the AllButOne queue in between; true once the sum
    // has two valid splits.
    auto process = [&](bool flush) -> bool {
        if (pn < 128 && !(flush && pn > 0)) return false;
        wave_sync_lds();
        while (pn >= 128 || (flush && pn > 0)) {
            const int nb = pn < 128 ? pn : 128;
            probe_batch(a, pq, pn - nb, nb, lane, lt, queue, qn);
            pn -= nb;
            wave_sync_lds();
            while (qn >= 64 || (flush && qn > 0)) {
                const int na = qn < 64 ? qn : 64;
                drain(a, tab, queue, qn - na, na, lane, pr, cnt, te, to);
                qn -= na;
                if (cnt > 1) return true;
            }
            wave_sync_lds();
        }
        return false;
    };
    uint32_t base = 0;
    while (base < total) {
        // uniform: the segment of base
        int jr = hint;
        while (jr + 1 < nseg && base >= segs[jr + 1].x) ++jr;
        const uint32_t segEnd = jr + 1 < nseg ? segs[jr + 1].x : total;
        if (sieveSum && base >= a.eager && segEnd - segs[jr].x >= kSieveMinListed) {
            // Scan: the rest of a long segment, [base, segEnd), is read off
            // the sieve over d instead of split by split.  Chunk by chunk
            // (odd d in [lo, lo + 2 kSieveBits)), lane l takes words l and
            // 64 + l of the chunk: its survivors are the unmarked bits of
            // listed classes within [d(base), d(segEnd - 1)], plus the
            // d = o split if listed (the chunk does not decide it: its Y is
            // not tested).  Survivors go straight to the probe queue, each
            // lane writing its own at its place in the wave's count.
            const uint32_t ey = __builtin_amdgcn_readfirstlane(segs[jr].y);
            const uint32_t k = ey & 31, c = (ey >> 8) & 15;
            const uint32_t first = __builtin_amdgcn_readfirstlane(segs[jr].x) - (ey >> 16);
            const uint32_t w = __builtin_amdgcn_readfirstlane(segs[jr].w);
            auto d_at = [&](uint32_t g) {
                const uint32_t ii = g - first;
                const uint32_t q = c == 1 ? ii : __umulhi(ii, w);
                return __umul24(q, 30u) + (segOffs[jr * 16 + (ii - q * c)] & 0x7Fu);
            };
            const uint32_t dStart = __builtin_amdgcn_readfirstlane(d_at(base));
            const uint32_t dEnd = __builtin_amdgcn_readfirstlane(d_at(segEnd - 1));
            // listed classes t (d = 2t+1 mod 30), and those whose o must be prime
            uint32_t cm = 0, nm = 0;
            if ((uint32_t)lane < c) {
                const uint32_t v = segOffs[jr * 16 + lane];
                cm = 1u << ((v & 0x7Fu) >> 1);
                nm = (v >> 7) ? cm : 0u;
            }
            for (int off = 1; off < 16; off <<= 1) {
                cm |= __shfl_xor(cm, off);
                nm |= __shfl_xor(nm, off);
            }
            const uint32_t classes = __builtin_amdgcn_readfirstlane(cm), needOs = __builtin_amdgcn_readfirstlane(nm);
            const uint64_t rep = classes | (uint64_t)classes << 15 | (uint64_t)classes << 30 | (uint64_t)classes << 45;
            const uint32_t m2 = (1u << k) + 1, q2 = s / m2;
            const uint32_t dStar = (q2 * m2 == s && (q2 & 1)) ? q2 : 0u;  // the d = o split, odd
            if (jr != sj) {
                const uint3 rk = sieve_items_for_k(item.q(), item.sm(), small, far, a.sieveK, a.sieveSqrt, k, lane);
                item.rk = rk.x;
                small = (small & 0xFFFF0000u) | rk.y;
                far = (far & 0xFFFFu) | rk.z << 16;
                if (far2On) far2_for_k((lds_u32 *)far2, a.sieveK, a.sieveSqrt, k, dStart, lane);
                sj = jr;
            }
#ifdef FFX_PROBE_COUNTERS
            if (lane == 0) FFX_COUNT(1, segEnd - base);
            visitedSplits += segEnd - base;
#endif
            for (uint32_t lo = dStart; lo <= dEnd; lo += 2 * kSieveBits) {
                sieve_chunk(item, small, far, (lds_u32 *)far2, lo, (lds_u32 *)dbits, lane);
                const uint32_t t0 = (lo % 30) >> 1;  // class of bit 0
                const uint32_t bEnd = (dEnd - lo) >> 1;
                const uint32_t bStar = (dStar - lo) >> 1;  // huge unless the d = o split is in the chunk
#pragma unroll 1
                for (uint32_t h = 0; h < kSieveBits / 2048; ++h) {
                    // queued in windows that fit the probe queue; a word's
                    // survivors are read again after each probe pass rather
                    // than held across it (registers)
                    for (uint32_t done = 0;;) {
                        const uint32_t wi = 64 * h + lane, wb = 32 * wi, r = (t0 + 2 * wi) % 15;  // 32 = 2 mod 15
                        uint32_t sv = ~dbits[wi];
                        if (dStar >= lo && (bStar >> 5) == wi) sv |= 1u << (bStar & 31);
                        sv &= (uint32_t)(rep >> r);
                        sv &= wb > bEnd ? 0u : (bEnd - wb >= 31 ? ~0u : (2u << (bEnd - wb)) - 1u);
                        // this lane's place among the wave's survivors (counts <= 32: six ballots)
                        const uint32_t n = __popc(sv);
                        uint32_t pre = 0, tot = 0;
#pragma unroll
                        for (int j = 0; j < 6; ++j) {
                            const uint64_t bj = __ballot((n >> j) & 1u);
                            pre += (uint32_t)__popcll(bj & lt) << j;
                            tot += (uint32_t)__popcll(bj) << j;
                        }
#ifdef FFX_PROBE_COUNTERS
                        if (lane == 0 && done == 0) {
                            FFX_COUNT(2, tot);
                            FFX_COUNT(13, 1);
                            FFX_COUNT(14, 1);
                        }
#endif
                        const uint32_t take = min(tot - done, (uint32_t)(kProbeCap - pn));
                        uint32_t idx = pre;
                        while (sv) {
                            const uint32_t bit = __ffs(sv) - 1;
                            sv &= sv - 1;
                            if (idx - done < take) {
                                const uint32_t d = lo + 2 * (wb + bit), o = s - (d << k);
                                const bool needO = needOs && ((needOs >> ((r + bit) % 15)) & 1u);
                                pq[pn + (idx - done)] = make_uint2(d | (k << kDBits), o | (needO ? kNeedO : 0u));
                            }
                            ++idx;
                        }
                        pn += take;
                        done += take;
                        const bool more = done < tot;
                        if (process(false)) {
                            FFX_COUNT_EXIT(11);
                            return 0;
                        }
                        if (!more) break;
                    }
                }
            }
            hint = jr;
            base = segEnd;
            if (process(base >= total)) {
                FFX_COUNT_EXIT(11);
                return 0;
            }
            continue;
        }
#ifdef FFX_PROBE_COUNTERS
        if (lane == 0) FFX_COUNT(1, total - base < 64u * kRoundSplits ? total - base : 64u * kRoundSplits);
        visitedSplits += total - base < 64u * kRoundSplits ? total - base : 64u * kRoundSplits;
        if (lane == 0) FFX_COUNT(13, 1);
#endif
        // Split g of segment j, through the filters that need no map read;
        // survivors go to the probe queue.
        auto filter = [&](uint32_t g, int j, uint4 e) {
            const bool live = g < total;
            const uint32_t k = e.y & 31, c = (e.y >> 8) & 15, ii = live ? g - e.x + (e.y >> 16) : 0u;
            const uint32_t q = c == 1 ? ii : __umulhi(ii, e.w);
            const uint32_t cls = segOffs[j * 16 + (ii - q * c)];
            const uint32_t d = 30 * q + (cls & 0x7Fu);
            const uint32_t o = live ? s - (d << k) : 1;
            // Power2Even without 3 | d*o: o must be prime (Power2Odd lists
            // hold only d with 3 !| d*o); fixed by d's class
            const bool needO = cls >> 7;
            bool rx, ry, ro;
            probe_residues(tab, d, o, k, e.z, rx, ry, ro);
            const bool keep = live && rx && (ry || d == o) && (!needO || ro);
            const uint64_t km = __ballot(keep);
            if (keep) pq[pn + __popcll(km & lt)] = make_uint2(d | (k << kDBits), o | (needO ? kNeedO : 0u));
#ifdef FFX_PROBE_COUNTERS
            if (lane == 0) FFX_COUNT(2, __popcll(km));
#endif
            pn += __popcll(km);
        };
        // uniform: whether the whole round lies in the segment of base (then
        // no lane searches the segment table)
        const uint32_t glast = (base + 64 * kRoundSplits < total ? base + 64 * kRoundSplits : total) - 1;
        if (glast < segEnd) {
            const uint4 e = segs[jr];
#pragma unroll
            for (int u = 0; u < kRoundSplits; ++u) filter(base + 64u * u + lane, jr, e);
            hint = jr;
        } else {
            int segOf[kRoundSplits];
#pragma unroll
            for (int u = 0; u < kRoundSplits; ++u) {
                const uint32_t g = base + 64u * u + lane;
                int j = hint;
                if (g < total)
                    while (j + 1 < nseg && g >= segs[j + 1].x) ++j;
                segOf[u] = j;
                filter(g, j, segs[j]);
            }
            hint = __shfl(segOf[kRoundSplits - 1], 63);
        }
        const bool flush = base + 64 * kRoundSplits <= a.eager || base + 64 * kRoundSplits >= total;
        base += 64 * kRoundSplits;
        if (process(flush)) {
            FFX_COUNT_EXIT(11);
            return 0;
        }
    }
    wave_sync_lds();
    while (qn > 0) {
        const int nb = qn < 64 ? qn : 64;
        drain(a, tab, queue, qn - nb, nb, lane, pr, cnt, te, to);
        qn -= nb;
        if (cnt > 1) {
            FFX_COUNT_EXIT(11);
            return 0;
        }
    }
    if (cnt != 1) {
        FFX_COUNT_EXIT(12);
        return 0;
    }
    FFX_COUNT_EXIT(10);
#undef FFX_COUNT_EXIT

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

// 4 blocks of 4 waves per CU: 128 registers per lane, no spills, and the
// probe and AllButOne queues of four waves fit four blocks' LDS.
#ifndef FFX_SEARCH_OCCUPANCY
#define FFX_SEARCH_OCCUPANCY 4
#endif
__global__ __launch_bounds__(kBlock, FFX_SEARCH_OCCUPANCY) void search_sums(SearchArgs a)
{
    __shared__ uint2 queues[kWavesPerBlock][kQueueCap];
    __shared__ uint2 probes[kWavesPerBlock][kProbeCap];
    __shared__ uint4 segs[kWavesPerBlock][kMaxSegs];
    __shared__ uint8_t segOffs[kWavesPerBlock][kMaxSegs * 16];  // allowed d mod 30 per segment
    __shared__ uint32_t primes[kBlock][kMaxPrimes];
    __shared__ uint32_t tab[kTabWords];
    __shared__ uint32_t dsieve[kWavesPerBlock][kSieveBits / 32];
    __shared__ uint32_t far2s[kWavesPerBlock][64 * kNFar2Sets];
    for (uint32_t i = threadIdx.x; i < kTabWords; i += kBlock) tab[i] = a.coprime[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.counter, a.grab);
        base = __shfl(base, 0);
        if (base >= a.nsums) break;
        const uint32_t end = base + a.grab < a.nsums ? base + a.grab : a.nsums;
        for (uint32_t idx = base; idx < end; ++idx) {
            const uint32_t slotIdx = a.nsums - 1 - idx;  // largest sums first
            const uint32_t s = sum_of_slot(a, slotIdx);
            const uint32_t res =
                solve_sum_wave(a, tab, s, lane, queues[w], probes[w], segs[w], segOffs[w], primes[threadIdx.x], dsieve[w],
                               far2s[w]);
            if (lane == 0) {
                if (!a.bySum) a.slots[slotIdx] = res;
                else if (s >= a.sumBase && ((s - a.sumBase) >> 1) < a.slotCap) a.slots[(s - a.sumBase) >> 1] = res;
            }
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
    uint32_t grab, eager, blocksPerCU, cus;
    uint32_t *dCoprime;
    uint16_t *dSpf;
    uint32_t spfLimit;
    uint32_t *dCounter;
    uint32_t *dTile;      // tile counts / offsets
    uint32_t tileCap;
    uint32_t *dSlots;     // scratch slots for ffx_search_async
    uint64_t slotCap;
    uint32_t *dSieveK;    // d-sieve tables (per prime and k; square roots)
    uint16_t *dSieveSqrt;
    int grid;
    int device;
};

// The plan's read-only tables, built on the host: the smallest prime factor
// of each odd n < spfLimit (index n>>1; 0 when n is 1 or prime), the
// residue filter's coprimality bits (bit r of modulus m set when
// gcd(r, m) == 1), and the sieve over d's tables: for sieve prime q and
// k < 32, 2^k mod q | (1-4^k)^-1 << 8 | (2^(k+1))^-1 << 16 | (1-4^k != 0)
// << 24, and a square root mod q of every quadratic residue (0xFF for the
// others).
struct SearchTables {
    uint32_t spfLimit;
    std::vector<uint16_t> spf;
    std::vector<uint32_t> coprime;
    std::vector<uint32_t> sieveK;
    std::vector<uint16_t> sieveSqrt;
};

static SearchTables build_search_tables(uint32_t sumLimitMax)
{
    SearchTables t;
    t.spfLimit = sumLimitMax + 2;
    t.spf.assign((t.spfLimit >> 1) + 1, 0);
    for (uint64_t p = 3; p * p < t.spfLimit; p += 2) {
        if (t.spf[p >> 1]) continue;
        for (uint64_t m = p * p; m < t.spfLimit; m += 2 * p)
            if (!t.spf[m >> 1]) t.spf[m >> 1] = (uint16_t)p;
    }
    t.coprime.assign(kTabWords, 0);
    for (int i = 0; i < kNMods; ++i)
        for (uint32_t r = 0; r < kMods[i]; ++r) {
            uint32_t g = r, m = kMods[i];
            while (m) {
                const uint32_t tmp = g % m;
                g = m;
                m = tmp;
            }
            // coprime, or a prime factor of the modulus (so a probe value
            // that is one of those primes is not ruled out)
            bool factor = r > 1 && kMods[i] % r == 0;
            for (uint32_t f = 2; factor && f * f <= r; ++f) factor = r % f != 0;
            if (g == 1 || factor) t.coprime[tab_offset(i) + r / 32] |= 1u << (r % 32);
        }
    t.sieveK.assign(kNSievePrimes * 32, 0);
    t.sieveSqrt.assign(kSieveSqrtBytes / 2, 0xFFFF);
    for (int i = 0; i < kNSievePrimes; ++i) {
        const uint32_t q = kSievePrimes[i];
        auto inv = [q](uint32_t x) {
            for (uint32_t y = 1; y < q; ++y)
                if (x * y % q == 1) return y;
            return 0u;
        };
        uint32_t pk = 1;
        for (int k = 0; k < 32; ++k) {
            const uint32_t A = (1 + q * q - pk * pk % q) % q;
            t.sieveK[i * 32 + k] = (A ? inv(A) : 0u) | inv(2 * pk % q) << 16;  // an inverse is never 0
            pk = 2 * pk % q;
        }
        for (uint32_t r = 0; r < q; ++r)
            if (t.sieveSqrt[i * kSqrtStride + r * r % q] == 0xFFFF) t.sieveSqrt[i * kSqrtStride + r * r % q] = (uint16_t)r;
    }
    return t;
}

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

// ------------------------------------------------------ allocation-free search
//
// ffx_search_records_noalloc runs the search of ffx_search_records with no
// allocation at call time: its work counter, tables and slots live in one
// file-scope device array, `scratch`, of kNaBytes (6.6 MB: sums up to
// FFX_NOALLOC_MAX_SUM).  Each device holds its own copy at its own address;
// the host takes that address once per device at ffx_init (with the device
// current) and passes it to the kernels.  The array's name is never used as
// a device address in host code.
constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint32_t kNaSpfLimit = FFX_NOALLOC_MAX_SUM + 2;
constexpr uint64_t kNaSlots = (FFX_NOALLOC_MAX_SUM - 5) / 2 + 1;
constexpr uint64_t kNaCounterOff = 0;  // 16 u32 work counters
constexpr uint64_t kNaCoprimeOff = 256;
constexpr uint64_t kNaSieveKOff = kNaCoprimeOff + na_align(kTabWords * 4ull);
constexpr uint64_t kNaSqrtOff = kNaSieveKOff + na_align(kNSievePrimes * 32ull * 4);
constexpr uint64_t kNaSpfOff = kNaSqrtOff + na_align(kSieveSqrtBytes);
constexpr uint64_t kNaSlotsOff = kNaSpfOff + na_align(((kNaSpfLimit >> 1) + 1ull) * 2);
constexpr uint64_t kNaBytes = kNaSlotsOff + na_align(kNaSlots * 4);
static_assert(kNaBytes < (8ull << 20), "the allocation-free scratch stays below 8 MiB");

static __device__ __align__(256) uint8_t scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
static std::atomic<uint8_t *> g_naScratch[kNaMaxDevices];  // scratch's device address, per device
static uint32_t g_naCus[kNaMaxDevices];

namespace ffx {

// Once per device, with that device current: the address of `scratch` on
// it, and the search tables for sums up to FFX_NOALLOC_MAX_SUM copied in.
int noalloc_setup(int device)
{
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0 || device >= kNaMaxDevices) return fail(FFX_ENODEV, "no-alloc set-up: device %d", device);
    if (g_naScratch[device].load(std::memory_order_acquire)) return FFX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    uint8_t *base = static_cast<uint8_t *>(p);
    const SearchTables t = build_search_tables(FFX_NOALLOC_MAX_SUM);
    if (t.spfLimit != kNaSpfLimit || t.spf.size() * 2 > kNaSlotsOff - kNaSpfOff)
        return fail(FFX_EIO, "no-alloc set-up: table layout");
    FFX_HIP(hipMemcpy(base + kNaCoprimeOff, t.coprime.data(), t.coprime.size() * 4, hipMemcpyHostToDevice));
    FFX_HIP(hipMemcpy(base + kNaSieveKOff, t.sieveK.data(), t.sieveK.size() * 4, hipMemcpyHostToDevice));
    FFX_HIP(hipMemcpy(base + kNaSqrtOff, t.sieveSqrt.data(), t.sieveSqrt.size() * 2, hipMemcpyHostToDevice));
    FFX_HIP(hipMemcpy(base + kNaSpfOff, t.spf.data(), t.spf.size() * 2, hipMemcpyHostToDevice));
    int cus = 0;
    FFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    g_naCus[device] = cus > 0 ? (uint32_t)cus : 256;
    g_naScratch[device].store(base, std::memory_order_release);
    return FFX_OK;
}

}  // namespace ffx

extern "C" {

#ifdef FFX_PROBE_COUNTERS
int ffx_probe_counters(uint64_t *out, int reset)
{
    if (!out) return fail(FFX_EINVAL, "probe counters: null");
    FFX_HIP(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_ffxCounters), sizeof(uint64_t) * 16));
    if (reset) {
        const uint64_t z[16] = {};
        FFX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_ffxCounters), z, sizeof z));
    }
    return FFX_OK;
}
#endif

int ffx_search_plan_create(ffx_search_plan **out, uint32_t sumLimitMax)
{
    if (!out) return fail(FFX_EINVAL, "search plan: null out");
    *out = nullptr;
    if (sumLimitMax > FFX_MAX_SUM) return fail(FFX_ERANGE, "search plan: sumLimitMax %u > %u", sumLimitMax, FFX_MAX_SUM);
    if (sumLimitMax < 5) sumLimitMax = 5;
    const SearchTables t = build_search_tables(sumLimitMax);
    ffx_search_plan *pl = new (std::nothrow) ffx_search_plan();
    if (!pl) return fail(FFX_ENOMEM, "search plan: host alloc");
    pl->sumLimitMax = sumLimitMax;
    pl->spfLimit = t.spfLimit;
    hipError_t e = hipGetDevice(&pl->device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
    if (e == hipSuccess) e = hipMalloc(&pl->dSpf, t.spf.size() * 2);
    if (e == hipSuccess) e = hipMemcpy(pl->dSpf, t.spf.data(), t.spf.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&pl->dCounter, 64);
    if (e == hipSuccess) e = hipMalloc(&pl->dCoprime, kTabWords * 4);
    if (e == hipSuccess) e = hipMemcpy(pl->dCoprime, t.coprime.data(), kTabWords * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&pl->dSieveK, t.sieveK.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(pl->dSieveK, t.sieveK.data(), t.sieveK.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&pl->dSieveSqrt, t.sieveSqrt.size() * 2);
    if (e == hipSuccess) e = hipMemcpy(pl->dSieveSqrt, t.sieveSqrt.data(), t.sieveSqrt.size() * 2, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "search plan upload");
        ffx_search_plan_destroy(pl);
        return rc;
    }
    pl->cus = cus > 0 ? (uint32_t)cus : 256;
    pl->grab = kGrab;
    pl->eager = kEagerSplits;
    pl->blocksPerCU = kBlocksPerCU;
    pl->grid = (int)(pl->cus * pl->blocksPerCU);
    *out = pl;
    return FFX_OK;
}

int ffx_search_plan_destroy(ffx_search_plan *plan)
{
    if (!plan) return FFX_OK;
    (void)hipFree(plan->dSpf);
    (void)hipFree(plan->dCoprime);
    (void)hipFree(plan->dSieveK);
    (void)hipFree(plan->dSieveSqrt);
    (void)hipFree(plan->dCounter);
    (void)hipFree(plan->dTile);
    (void)hipFree(plan->dSlots);
    delete plan;
    return FFX_OK;
}

static int launch_search(ffx_search_plan *plan, SearchArgs &a, void *stream)
{
    a.spf = plan->dSpf;
    a.spfLimit = plan->spfLimit;
    a.counter = plan->dCounter;
    a.coprime = plan->dCoprime;
    a.sieveK = plan->dSieveK;
    a.sieveSqrt = plan->dSieveSqrt;
    a.grab = plan->grab;
    a.eager = plan->eager;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, st, plan->dCounter, 16u);
    FFX_HIP(hipGetLastError());
    const uint32_t need = (a.nsums + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t grid = need < (uint32_t)plan->grid ? need : (uint32_t)plan->grid;
    hipLaunchKernelGGL(search_sums, dim3(grid), dim3(kBlock), 0, st, a);
    FFX_HIP(hipGetLastError());
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
    SearchArgs a = {};
    a.map = static_cast<const uint8_t *>(d_map);
    a.M = maxPrimeMapValue;
    a.first = first;
    a.nsums = (sumLimit - first) / 2 + 1;
    a.slots = d_slots;
    return launch_search(plan, a, stream);
}

int ffx_search_intervals_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue,
                               const ffx_interval *d_intervals, uint32_t nintervals, uint64_t totalSlots,
                               uint32_t lastSum, uint32_t *d_slots, void *stream)
{
    if (!plan || !d_map || !d_slots || (!d_intervals && nintervals)) return fail(FFX_EINVAL, "search: null argument");
    if (nintervals == 0 || totalSlots == 0) return FFX_OK;
    if (lastSum > plan->sumLimitMax)
        return fail(FFX_ERANGE, "search: interval sum %u above the plan's %u", lastSum, plan->sumLimitMax);
    if (totalSlots > 0xFFFFFFFFull) return fail(FFX_ERANGE, "search: more than 2^32 slots");
    if (maxPrimeMapValue < 15) return fail(FFX_EINVAL, "search: map too small");
    SearchArgs a = {};
    a.map = static_cast<const uint8_t *>(d_map);
    a.M = maxPrimeMapValue;
    a.nsums = (uint32_t)totalSlots;
    a.slots = d_slots;
    a.iv = d_intervals;
    a.niv = nintervals;
    return launch_search(plan, a, stream);
}

int ffx_search_intervals_by_sum_async(ffx_search_plan *plan, const void *d_map, uint64_t maxPrimeMapValue,
                                      const ffx_interval *d_intervals, uint32_t nintervals, uint64_t totalSlots,
                                      uint32_t lastSum, uint32_t slotSumBase, uint32_t *d_slots, uint64_t slotCapacity,
                                      void *stream)
{
    if (!plan || !d_map || !d_slots || (!d_intervals && nintervals)) return fail(FFX_EINVAL, "search: null argument");
    if (nintervals == 0 || totalSlots == 0) return FFX_OK;
    if (!(slotSumBase & 1)) return fail(FFX_EINVAL, "search: slotSumBase must be odd");
    if (lastSum > plan->sumLimitMax)
        return fail(FFX_ERANGE, "search: interval sum %u above the plan's %u", lastSum, plan->sumLimitMax);
    if (totalSlots > 0xFFFFFFFFull) return fail(FFX_ERANGE, "search: more than 2^32 slots");
    if (maxPrimeMapValue < 15) return fail(FFX_EINVAL, "search: map too small");
    SearchArgs a = {};
    a.map = static_cast<const uint8_t *>(d_map);
    a.M = maxPrimeMapValue;
    a.nsums = (uint32_t)totalSlots;
    a.slots = d_slots;
    a.iv = d_intervals;
    a.niv = nintervals;
    a.bySum = 1;
    a.sumBase = slotSumBase;
    a.slotCap = slotCapacity;
    return launch_search(plan, a, stream);
}

int ffx_search_plan_set(ffx_search_plan *plan, int param, uint32_t value)
{
    if (!plan) return fail(FFX_EINVAL, "search plan: null");
    switch (param) {
    case FFX_SEARCH_GRAB:
        if (value < 1 || value > 64) return fail(FFX_ERANGE, "grab %u outside 1..64", value);
        plan->grab = value;
        return FFX_OK;
    case FFX_SEARCH_EAGER_SPLITS:
        plan->eager = value;
        return FFX_OK;
    case FFX_SEARCH_BLOCKS_PER_CU:
        if (value < 1 || value > 16) return fail(FFX_ERANGE, "blocks per CU %u outside 1..16", value);
        plan->blocksPerCU = value;
        plan->grid = (int)(plan->cus * value);
        return FFX_OK;
    default:
        return fail(FFX_EINVAL, "search plan: unknown parameter %d", param);
    }
}

int ffx_search_plan_get(const ffx_search_plan *plan, int param, uint32_t *value)
{
    if (!plan || !value) return fail(FFX_EINVAL, "search plan: null");
    switch (param) {
    case FFX_SEARCH_GRAB: *value = plan->grab; return FFX_OK;
    case FFX_SEARCH_EAGER_SPLITS: *value = plan->eager; return FFX_OK;
    case FFX_SEARCH_BLOCKS_PER_CU: *value = plan->blocksPerCU; return FFX_OK;
    default: return fail(FFX_EINVAL, "search plan: unknown parameter %d", param);
    }
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

int ffx_search_records_noalloc(int device, const uint8_t *d_map, uint64_t maxPrimeMapValue, uint32_t sumStart,
                               uint32_t sumLimit, uint32_t *d_count, ffx_gpu_record *d_records, void *stream)
{
    if (!d_map || !d_count || !d_records) return fail(FFX_EINVAL, "search records (no-alloc): null argument");
    if (device < 0 || device >= kNaMaxDevices) return fail(FFX_ENODEV, "search records (no-alloc): device %d", device);
    uint8_t *p = g_naScratch[device].load(std::memory_order_acquire);
    if (!p) return fail(FFX_EINVAL, "search records (no-alloc): call ffx_init(%d) first", device);
    if ((sumStart & 1) == 0 || sumStart < 5)
        return fail(FFX_EINVAL, "search records (no-alloc): sumStart must be odd and >= 5");
    if (sumLimit > FFX_NOALLOC_MAX_SUM)
        return fail(FFX_ERANGE, "search records (no-alloc): sumLimit %u above the maximum %u of its scratch array",
                    sumLimit, FFX_NOALLOC_MAX_SUM);
    if (maxPrimeMapValue < 15) return fail(FFX_EINVAL, "search records (no-alloc): map too small");
    if (sumLimit < sumStart) return FFX_OK;
    FFX_HIP(hipSetDevice(device));
    SearchArgs a = {};
    a.map = d_map;
    a.M = maxPrimeMapValue;
    a.first = sumStart;
    a.nsums = (sumLimit - sumStart) / 2 + 1;
    a.slots = reinterpret_cast<uint32_t *>(p + kNaSlotsOff);
    a.spf = reinterpret_cast<const uint16_t *>(p + kNaSpfOff);
    a.spfLimit = kNaSpfLimit;
    a.counter = reinterpret_cast<uint32_t *>(p + kNaCounterOff);
    a.coprime = reinterpret_cast<const uint32_t *>(p + kNaCoprimeOff);
    a.sieveK = reinterpret_cast<const uint32_t *>(p + kNaSieveKOff);
    a.sieveSqrt = reinterpret_cast<const uint16_t *>(p + kNaSqrtOff);
    a.grab = kGrab;
    a.eager = kEagerSplits;
    hipStream_t st = as_stream(stream);
    // the counter is cleared by a kernel (the array keeps the last call's
    // contents); every slot in [0, nsums) is written by the search
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, st, a.counter, 16u);
    FFX_HIP(hipGetLastError());
    const uint32_t maxGrid = g_naCus[device] * kBlocksPerCU;
    const uint32_t need = (a.nsums + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(search_sums, dim3(need < maxGrid ? need : maxGrid), dim3(kBlock), 0, st, a);
    FFX_HIP(hipGetLastError());
    hipLaunchKernelGGL(slots_to_records, dim3((a.nsums + 255) / 256), dim3(256), 0, st, a.slots, (uint64_t)a.nsums,
                       sumStart, d_records, d_count);
    FFX_HIP(hipGetLastError());
    return FFX_OK;
}

}  // extern "C"
