// ff_lds_atomics: the device's rate of LDS atomic ORs (ds_or_b32), the
// strike instruction of the sieve, for the address patterns the sieve
// issues.  Each 1024-thread workgroup owns a 64 KiB LDS array (two per CU,
// as the sieve's 64 KiB-segment kernel runs); every lane keeps 16 word
// addresses in registers and ORs into them round after round, so the timed
// loop is LDS-bound, not address-arithmetic-bound.  Patterns:
//   conflict_free  lane l of a wave hits word 64*k + l: 64 distinct banks
//   random         16 uniform words per lane (LCG)
//   coop           the cooperative strikes: each half-wave one odd prime p,
//                  lane l on the word of multiple (l + 32t)p (one prime per
//                  half-wave, p from 67 up)
//   lane           one prime per lane, p from 8191 up, consecutive multiples
// Prints one JSON line: per pattern, lane-ORs per second for the whole
// device, and wave instructions per CU per clock.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kThreads = 1024, kWords = 16384, kAddrs = 16;

__global__ __launch_bounds__(kThreads) void lds_or(int mode, uint32_t rounds, uint32_t *out)
{
    __shared__ uint32_t seg[kWords];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t i = tid; i < kWords; i += kThreads) seg[i] = 0;
    uint32_t addr[kAddrs], mask[kAddrs];
    uint64_t x = 0x9E3779B97F4A7C15ull * (blockIdx.x * kThreads + tid + 1);
    const uint32_t half = lane >> 5, hl = lane & 31;
    const uint32_t pCoop = 67 + 2 * (2 * (wave + 16 * (blockIdx.x & 15)) + half) * 3;  // odd, spread over 67..~6200
    const uint32_t pLane = 8191 + 2 * (tid + kThreads * (blockIdx.x & 7));           // odd
#pragma unroll
    for (int k = 0; k < kAddrs; ++k) {
        uint32_t w, b;
        if (mode == 0) {
            w = (64u * (wave * kAddrs + k) + lane) & (kWords - 1);
            b = lane & 31;
        } else if (mode == 1) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            w = (uint32_t)(x >> 40) & (kWords - 1);
            b = (uint32_t)(x >> 33) & 31;
        } else if (mode == 2) {
            const uint32_t j = (hl + 32u * k) * pCoop + 7u * wave;
            w = (j >> 5) & (kWords - 1);
            b = j & 31;
        } else {
            const uint32_t j = k * pLane + 13u * tid;
            w = (j >> 5) & (kWords - 1);
            b = j & 31;
        }
        addr[k] = w;
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

int check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    std::fprintf(stderr, "ff_lds_atomics: %s: %s\n", what, hipGetErrorString(e));
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const uint32_t rounds = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 512;
    int dev = 0, cus = 0, clockKHz = 0;
    if (check(hipGetDevice(&dev), "device") ||
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attr") ||
        check(hipDeviceGetAttribute(&clockKHz, hipDeviceAttributeClockRate, dev), "attr"))
        return 1;
    const int blocks = 2 * cus;
    uint32_t *out = nullptr;
    hipEvent_t e0, e1;
    if (check(hipMalloc(&out, (size_t)blocks * kThreads * 4), "hipMalloc") || check(hipEventCreate(&e0), "event") ||
        check(hipEventCreate(&e1), "event"))
        return 1;
    const char *names[4] = {"conflict_free", "random", "coop", "lane"};
    std::printf("{\"cus\": %d, \"clock_mhz\": %.0f, \"rounds\": %u", cus, clockKHz / 1000.0, rounds);
    for (int mode = 0; mode < 4; ++mode) {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
            if (check(hipEventRecord(e0), "record")) return 1;
            hipLaunchKernelGGL(lds_or, dim3(blocks), dim3(kThreads), 0, 0, mode, rounds, out);
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
        std::printf(", \"%s\": {\"ms\": %.4f, \"lane_ors_per_s\": %.4e, \"wave_ors_per_cu_clock\": %.4f}", names[mode],
                    best, laneOps / (best * 1e-3), perCuClock);
    }
    std::printf("}\n");
    (void)hipFree(out);
    return 0;
}
