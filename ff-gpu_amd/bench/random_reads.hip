// ff_random_reads: the device's ceiling for random single-byte reads, the
// access pattern of the search's map probes.  Each of 4096 x 256 threads
// reads 256 bytes at addresses from its own 64-bit LCG over a buffer of
// `bytes` (default 4 GiB, the 2^20 workload's map size); prints one JSON
// line with the best of three launches: reads per second, and HBM bytes per
// read when the caller supplies FETCH_SIZE (rocprofv3) separately.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kThreads = 256, kBlocks = 4096, kReads = 256;

__global__ __launch_bounds__(kThreads) void random_reads(const uint8_t *buf, uint64_t n, uint32_t *out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x = 0x9E3779B97F4A7C15ull * (t + 1);
    uint32_t acc = 0;
    for (int r = 0; r < kReads; r += 4) {
        uint64_t a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            a[u] = (x >> 16) % n;  // always inside the buffer
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += buf[a[u]];
    }
    out[t] = acc;
}

int check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    std::fprintf(stderr, "ff_random_reads: %s: %s\n", what, hipGetErrorString(e));
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const uint64_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (1ull << 32);
    if (n < 1024) return std::fprintf(stderr, "ff_random_reads: buffer too small\n"), 2;
    uint8_t *buf = nullptr;
    uint32_t *out = nullptr;
    hipEvent_t e0, e1;
    if (check(hipMalloc(&buf, n), "hipMalloc") || check(hipMalloc(&out, (size_t)kThreads * kBlocks * 4), "hipMalloc") ||
        check(hipMemset(buf, 1, n), "hipMemset") || check(hipEventCreate(&e0), "event") ||
        check(hipEventCreate(&e1), "event"))
        return 1;
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
        if (check(hipEventRecord(e0), "record")) return 1;
        hipLaunchKernelGGL(random_reads, dim3(kBlocks), dim3(kThreads), 0, 0, buf, n, out);
        if (check(hipGetLastError(), "launch") || check(hipEventRecord(e1), "record") ||
            check(hipEventSynchronize(e1), "sync"))
            return 1;
        float ms = 0;
        if (check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
        if (rep > 0 && ms < best) best = ms;
    }
    const double reads = (double)kReads * kThreads * kBlocks;
    std::printf("{\"buffer_bytes\": %llu, \"reads_per_launch\": %.0f, \"best_ms\": %.4f, \"reads_per_s\": %.4e}\n",
                (unsigned long long)n, reads, best, reads / (best * 1e-3));
    (void)hipFree(buf);
    (void)hipFree(out);
    return 0;
}
