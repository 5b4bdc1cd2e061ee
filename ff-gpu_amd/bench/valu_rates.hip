// ff_valu_rates: issue rate of single VALU instructions on one device, in
// wave64 instructions per CU per clock.  Each kernel runs 8 independent
// register-only chains per lane (16 instructions per loop step, inline asm so
// the instruction is the one named), 8 workgroups of 256 per CU; the best of
// two timed launches after a warm-up counts.  Prints one JSON line.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>

#define CHAIN8(ASM) \
    asm volatile(ASM : "+v"(a0) : "v"(b)); asm volatile(ASM : "+v"(a1) : "v"(b)); \
    asm volatile(ASM : "+v"(a2) : "v"(b)); asm volatile(ASM : "+v"(a3) : "v"(b)); \
    asm volatile(ASM : "+v"(a4) : "v"(b)); asm volatile(ASM : "+v"(a5) : "v"(b)); \
    asm volatile(ASM : "+v"(a6) : "v"(b)); asm volatile(ASM : "+v"(a7) : "v"(b));

#define KERNEL(NAME, ASM) \
__global__ __launch_bounds__(256) void NAME(uint32_t *out, int iters, uint32_t seed) { \
    uint32_t a0 = threadIdx.x ^ seed, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7; \
    uint32_t b = seed | 1; \
    for (int i = 0; i < iters; ++i) { CHAIN8(ASM) CHAIN8(ASM) } \
    out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7; }

KERNEL(k_add, "v_add_u32 %0, %0, %1")
KERNEL(k_mul_lo, "v_mul_lo_u32 %0, %0, %1")
KERNEL(k_mul_hi, "v_mul_hi_u32 %0, %0, %1")
KERNEL(k_mul_u24, "v_mul_u32_u24 %0, %0, %1")
KERNEL(k_mul_hi_u24, "v_mul_hi_u32_u24 %0, %0, %1")
KERNEL(k_mad_u24, "v_mad_u32_u24 %0, %0, %1, %0")
KERNEL(k_mad_i24, "v_mad_i32_i24 %0, %0, %1, %0")
KERNEL(k_cvt_f32_u32, "v_cvt_f32_u32 %0, %0")
KERNEL(k_cvt_u32_f32, "v_cvt_u32_f32 %0, %0")
KERNEL(k_mul_f32, "v_mul_f32 %0, %0, %1")
KERNEL(k_lshr, "v_lshrrev_b32 %0, %1, %0")
KERNEL(k_bfe, "v_bfe_u32 %0, %0, %1, 5")
KERNEL(k_min, "v_min_u32 %0, %0, %1")
KERNEL(k_add3, "v_add3_u32 %0, %0, %1, %0")
KERNEL(k_lshl_add, "v_lshl_add_u32 %0, %0, 2, %1")
KERNEL(k_rcp_f32, "v_rcp_f32 %0, %0")

__global__ __launch_bounds__(256) void k_mad64(uint32_t *out, int iters, uint32_t seed)
{
    uint64_t a[8];
    for (int j = 0; j < 8; ++j) a[j] = threadIdx.x ^ seed ^ j;
    uint32_t b = seed | 1;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int j = 0; j < 8; ++j) asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(a[j]) : "v"(b), "v"(b) : "vcc");
    uint32_t x = 0;
    for (int j = 0; j < 8; ++j) x ^= (uint32_t)a[j];
    out[blockIdx.x * 256 + threadIdx.x] = x;
}

int main()
{
    int dev = 0, cus = 0, clk = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev) != hipSuccess)
        return 1;
    const int blocks = cus * 8, iters = 4096;
    uint32_t *out;
    if (hipMalloc(&out, (size_t)blocks * 256 * 4) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    struct { const char *n; void (*k)(uint32_t *, int, uint32_t); } ks[] = {
        {"v_add_u32", k_add}, {"v_mul_lo_u32", k_mul_lo}, {"v_mul_hi_u32", k_mul_hi}, {"v_mul_u32_u24", k_mul_u24},
        {"v_mul_hi_u32_u24", k_mul_hi_u24}, {"v_mad_u32_u24", k_mad_u24}, {"v_mad_i32_i24", k_mad_i24},
        {"v_cvt_f32_u32", k_cvt_f32_u32}, {"v_cvt_u32_f32", k_cvt_u32_f32}, {"v_mul_f32", k_mul_f32},
        {"v_lshrrev_b32", k_lshr}, {"v_bfe_u32", k_bfe}, {"v_min_u32", k_min}, {"v_add3_u32", k_add3},
        {"v_lshl_add_u32", k_lshl_add}, {"v_rcp_f32", k_rcp_f32}, {"v_mad_u64_u32", k_mad64}};
    printf("{\"cus\": %d, \"clock_mhz\": %d", cus, clk / 1000);
    for (auto &k : ks) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(e0) != hipSuccess) return 2;
            hipLaunchKernelGGL(k.k, dim3(blocks), dim3(256), 0, 0, out, iters, 12345u + rep);
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return 2;
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 2;
            if (rep && ms < best) best = ms;
        }
        const double waveInsts = (double)blocks * 4 * iters * 16;
        printf(", \"%s\": %.3f", k.n, waveInsts / cus / (best * 1e-3 * clk * 1e3));
    }
    printf("}\n");
    (void)hipFree(out);
    return 0;
}
