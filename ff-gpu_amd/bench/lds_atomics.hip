// ff_lds_atomics: the device's rate of LDS atomic ORs (ds_or_b32), the
// strike instruction of the sieve, for the address patterns the sieve
// issues and for the strided patterns that show the bank structure.  Each
// 1024-thread workgroup owns an LDS array of `words` words: by default
// 28672 (112 KiB, one workgroup per CU, as the large sieve kernel runs),
// or 16384 (64 KiB, two per CU); every lane keeps 16 word addresses in
// registers and ORs into them round after round, so the timed loop is
// LDS-bound, not address-arithmetic-bound.  Patterns (lane l of a wave; a "prime" is an odd
// p from 67 up with progression stride s = 3p, the mod-6 wheel's, and a
// first element f in [0, s) from a hash):
//   stride_k         lane l on word l*k (k = 1, 2, 3, 4, 8, 16, 32, 64)
//   scattered        16 hashed words per lane
//   coop             the cooperative strikes: a half-wave per prime, lane l
//                    on element l + 32t (one fixed bit, word advancing by s)
//   coop_distinct    the coop words with each half-wave's 32 lanes moved to
//                    32 distinct banks (not a sieve pattern: the ceiling)
//   coop_mul         coop with lane l on residue (mul * l) mod 32, mul the
//                    odd multiplier that minimises the busiest bank of the
//                    two 16-lane passes (chosen per stride on the host)
//   coop_wave        one prime per wave, lane l on element l + 64t
//   coop_quarter     one prime per 16 lanes, lane l on element l + 16t
//   coop_interleave  two primes per half-wave, alternating lanes
//   coop_bitrev      coop with lane l on residue bitreverse5(l)
//   coop_even64      a half-wave per prime on the even residues of 64
//   lane             one prime per lane (p from 4779 up, the large kernel's
//                    lane primes), hashed first strike
// Prints one JSON line: per pattern, lane-ORs per second for the whole
// device, and wave instructions per CU per clock.
//   ff_lds_atomics [rounds [words]]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kThreads = 1024, kAddrs = 16;

enum Pattern : int {
    kStride = 0,  // param = k
    kScattered,
    kCoop,
    kCoopDistinct,
    kCoopMul,
    kCoopWave,
    kCoopQuarter,
    kCoopInterleave,
    kCoopBitrev,
    kCoopEven64,
    kLane,
};

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kThreads) void lds_or(int pattern, uint32_t param, const uint8_t *mulTab, uint32_t rounds,
                                                   uint32_t kWords, uint32_t *out)
{
    extern __shared__ uint32_t seg[];  // kWords words
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t i = tid; i < kWords; i += kThreads) seg[i] = 0;
    // the prime slot of this lane and its element residue / lanes per prime
    uint32_t slot = lane >> 5, res = lane & 31, per = 32;
    switch (pattern) {
    case kCoopMul: break;  // residue set below, from the stride
    case kCoopWave: slot = 0; res = lane; per = 64; break;
    case kCoopQuarter: slot = lane >> 4; res = lane & 15; per = 16; break;
    case kCoopInterleave: slot = lane & 1; res = (lane >> 1) & 31; per = 32; break;
    case kCoopBitrev: res = __builtin_bitreverse32(lane & 31) >> 27; break;
    case kCoopEven64: res = 2 * (lane & 31); per = 64; break;
    default: break;
    }
    const uint32_t pidx = (wave * 4 + slot) * 37 + blockIdx.x * 13;
    const uint32_t p = 67 + 2 * (pidx % 1160);  // odd, 67..2385
    const uint32_t s = 3 * p;
    const uint32_t f = (uint32_t)(mix64(pidx) % s);
    if (pattern == kCoopMul) res = (mulTab[(s & 1023) >> 1] * (lane & 31)) & 31;
    uint32_t addr[kAddrs], mask[kAddrs];
#pragma unroll
    for (int k = 0; k < kAddrs; ++k) {
        uint32_t w = 0, b = lane & 31;
        if (pattern == kStride) {
            w = lane * param + 64u * param * (wave * kAddrs + k);
        } else if (pattern == kScattered) {
            const uint64_t h = mix64(((uint64_t)blockIdx.x * kThreads + tid) * kAddrs + k);
            w = (uint32_t)(h >> 32);
            b = (uint32_t)h & 31;
        } else if (pattern == kLane) {
            const uint32_t pl = 4779 + 2 * (tid + kThreads * (blockIdx.x & 15));
            const uint32_t j = (uint32_t)(mix64((uint64_t)blockIdx.x * kThreads + tid) % (kWords * 32ull)) + k * pl;
            w = j >> 5;
            b = j & 31;
        } else {
            const uint32_t j = f + (res + per * k) * s;
            w = j >> 5;
            b = j & 31;
            if (pattern == kCoopDistinct) w = (w & ~31u) | (2 * (lane & 15) + ((lane >> 4) & 1));
        }
        addr[k] = w % kWords;
        mask[k] = 1u << b;
    }
    __syncthreads();
    for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
        for (int k = 0; k < kAddrs; ++k) atomicOr(&seg[addr[k]], mask[k]);
    }
    __syncthreads();
    uint32_t acc = 0;
    for (uint32_t i = tid; i < kWords; i += kThreads) acc ^= seg[i];
    out[blockIdx.x * kThreads + tid] = acc;
}

// The odd multiplier mul < 32 for the stride s whose two 16-lane passes
// (lanes 0..15 and 16..31 on residues mul * l mod 32) have the smallest
// summed busiest-bank count over the 32 offsets of the first element in its
// word (element r's word is (f0 + r * s) >> 5 past a common base).
uint8_t lane_multiplier(uint32_t s)
{
    uint32_t best = 1, bestCost = ~0u;
    for (uint32_t mul = 1; mul < 32; mul += 2) {
        uint32_t cost = 0;
        for (uint32_t f0 = 0; f0 < 32; ++f0)
            for (uint32_t pass = 0; pass < 2; ++pass) {
                uint8_t cnt[32] = {};
                uint32_t busiest = 0;
                for (uint32_t l = 16 * pass; l < 16 * pass + 16; ++l) {
                    const uint32_t bank = ((f0 + ((mul * l) & 31) * s) >> 5) & 31;
                    busiest = std::max<uint32_t>(busiest, ++cnt[bank]);
                }
                cost += busiest;
            }
        if (cost < bestCost) {
            bestCost = cost;
            best = mul;
        }
    }
    return (uint8_t)best;
}

int check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    std::fprintf(stderr, "ff_lds_atomics: %s: %s\n", what, hipGetErrorString(e));
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const uint32_t rounds = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 256;
    const uint32_t words = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 28672;
    if (words != 28672 && words != 16384) {
        std::fprintf(stderr, "ff_lds_atomics: words must be 28672 (one workgroup per CU) or 16384 (two)\n");
        return 2;
    }
    int dev = 0, cus = 0, clockKHz = 0;
    if (check(hipGetDevice(&dev), "device") ||
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attr") ||
        check(hipDeviceGetAttribute(&clockKHz, hipDeviceAttributeClockRate, dev), "attr"))
        return 1;
    const int blocks = words == 16384 ? 2 * cus : cus;
    uint8_t mulHost[512];
    for (uint32_t i = 0; i < 512; ++i) mulHost[i] = lane_multiplier(2 * i + 1);
    uint32_t *out = nullptr;
    uint8_t *mulTab = nullptr;
    hipEvent_t e0, e1;
    if (check(hipMalloc(&out, (size_t)blocks * kThreads * 4), "hipMalloc") || check(hipMalloc(&mulTab, 512), "hipMalloc") ||
        check(hipMemcpy(mulTab, mulHost, 512, hipMemcpyHostToDevice), "copy") || check(hipEventCreate(&e0), "event") ||
        check(hipEventCreate(&e1), "event"))
        return 1;
    struct Case {
        int pattern;
        uint32_t param;
        const char *name;
    };
    const Case cases[] = {
        {kStride, 1, "stride_1"},       {kStride, 2, "stride_2"},       {kStride, 3, "stride_3"},
        {kStride, 4, "stride_4"},       {kStride, 8, "stride_8"},       {kStride, 16, "stride_16"},
        {kStride, 32, "stride_32"},     {kStride, 64, "stride_64"},     {kScattered, 0, "scattered"},
        {kCoop, 0, "coop"},             {kCoopDistinct, 0, "coop_distinct"}, {kCoopMul, 0, "coop_mul"},
        {kCoopWave, 0, "coop_wave"},    {kCoopQuarter, 0, "coop_quarter"}, {kCoopInterleave, 0, "coop_interleave"},
        {kCoopBitrev, 0, "coop_bitrev"}, {kCoopEven64, 0, "coop_even64"}, {kLane, 0, "lane"},
    };
    std::printf("{\"cus\": %d, \"clock_mhz\": %.0f, \"rounds\": %u, \"words\": %u, \"workgroups_per_cu\": %d", cus,
                clockKHz / 1000.0, rounds, words, blocks / cus);
    for (const Case &c : cases) {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
            if (check(hipEventRecord(e0), "record")) return 1;
            hipLaunchKernelGGL(lds_or, dim3(blocks), dim3(kThreads), words * 4, 0, c.pattern, c.param, mulTab, rounds, words,
                               out);
            if (check(hipGetLastError(), "launch") || check(hipEventRecord(e1), "record") ||
                check(hipEventSynchronize(e1), "sync"))
                return 1;
            float ms = 0;
            if (check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
            if (rep > 0 && ms < best) best = ms;
        }
        const double laneOps = (double)blocks * kThreads * kAddrs * rounds;
        const double waveOps = laneOps / 64.0;
        const double perCuClock = waveOps / cus / (best * 1e-3 * clockKHz * 1e3);
        std::printf(", \"%s\": {\"ms\": %.4f, \"lane_ors_per_s\": %.4e, \"wave_ors_per_cu_clock\": %.4f}", c.name, best,
                    laneOps / (best * 1e-3), perCuClock);
        std::fflush(stdout);
    }
    std::printf("}\n");
    (void)hipFree(out);
    (void)hipFree(mulTab);
    return 0;
}
