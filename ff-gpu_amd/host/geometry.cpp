#include "geometry.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace ffg {

uint64_t isqrt64(uint64_t n)
{
    uint64_t r = (uint64_t)std::sqrt((double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}

LegGeometry computeLegGeometry(uint64_t sumStart, uint64_t sumLimit)
{
    LegGeometry g;
    g.sumStart = sumStart | 1;
    if (g.sumStart < 5) g.sumStart = 5;
    g.sumLimit = sumLimit;
    if (g.sumStart > g.sumLimit) g.sumLimit = g.sumStart;
    g.maxGeneratedPrime = (g.sumLimit + 1) >> 1;
    g.productLimit = g.maxGeneratedPrime * g.maxGeneratedPrime;
    g.primeLimit = (g.productLimit >> 2) + 2;
    const uint64_t lim = g.primeLimit > 1 ? g.primeLimit : 2;
    g.mapBytes = (lim + 15) >> 4;
    g.maxPrimeMapValue = (g.mapBytes << 4) - 1;
    return g;
}

uint64_t searchWorkspaceBytes(uint64_t sumLimit)
{
    const uint64_t mgp = (sumLimit + 1) >> 1;
    const uint64_t productLimit = mgp * mgp;
    // greatestNumberPrimeFactors = log(productLimit)/log(2) + .5, truncated
    uint64_t g = (uint64_t)(std::log((double)productLimit) / std::log(2.0) + 0.5);
    if (g < 2) g = 2;
    return ((g * (g - 1)) << 1) * sizeof(uint64_t);
}

// double -> uint32 as g++/x86-64 converts it in ff_seg.
static uint32_t dbl_to_u32(double x) { return (uint32_t)(int64_t)x; }

static std::string fmt(const char *f, unsigned long long a, unsigned long long b, unsigned long long c)
{
    char buf[200];
    std::snprintf(buf, sizeof buf, f, a, b, c);
    return buf;
}

SegGeometry computeSegGeometry(uint32_t sumStart, uint32_t sumLimit, uint64_t primeMem)
{
    SegGeometry g;
    g.sumStart = sumStart;
    g.sumLimit = sumLimit;
    const uint32_t mgp = (uint32_t)(sumLimit + 1u) >> 1;
    g.productLimit = (uint64_t)mgp * mgp;
    g.primeLimit = (g.productLimit >> 2) + 2;
    const double xlogx = (double)mgp / std::log((double)mgp);
    const uint32_t nPrimes = dbl_to_u32(xlogx + 1.2762 * xlogx / std::log((double)mgp));
    uint64_t R = nPrimes > 1 ? nPrimes : 1;
    uint64_t mapBytes = ((g.primeLimit > 1 ? g.primeLimit : 2) + 15) >> 4;
    uint64_t M = (mapBytes << 4) - 1;
    uint64_t est = (uint64_t)((double)R * (std::log((double)R) + std::log(std::log((double)R))));
    if (M < est) M = est;
    mapBytes = (M + 15) >> 4;
    uint64_t req = 4 * R, total = req + mapBytes;
    g.headers.push_back(fmt("maxPrimeMapValue %llu numPrimesRequested %llu totalBytes %llu", M, R, total));
    if (total > primeMem) {
        if (16 * primeMem > 3 * est) {
            g.capped = true;
            const uint64_t R2 = 2 * R / 5, req2 = 4 * R2;
            est = (uint64_t)((double)R2 * (std::log((double)R2) + std::log(std::log((double)R2))));
            if (req2 + (est + 1) / 2 > primeMem) {
                R = ((req * primeMem + total - 1) / total + 3) & ~(uint64_t)3;
                req = 4 * R;
                g.headers.push_back("You probably should consider how badly this run is going to perform, just sayin'.");
            } else {
                R = R2;
                req = req2;
            }
            mapBytes = primeMem - req;
            total = req + mapBytes;
            M = (mapBytes << 4) - 1;
            g.headers.push_back(fmt("numPrimesRequested %llu maxPrimeMapValue %llu totalBytes %llu", R, M, total));
        } else {
            g.fatal = true;
            return g;
        }
    }
    g.numPrimesRequested = R;
    g.maxPrimeMapValue = M;
    g.mapBytes = mapBytes;
    g.totalBytes = total;
    g.sqrtLimit = isqrt64(M);
    g.headers.push_back(fmt("maxPrimeMapValue %llu numPrimesRequested %llu sqrtLimit %llu", M, R, g.sqrtLimit));
    return g;
}

void segPositionals(const std::vector<std::string> &pos, uint32_t *sumStart, uint32_t *sumLimit)
{
    uint32_t a, b;
    if (pos.size() > 1) {
        a = (uint32_t)std::atoi(pos[0].c_str()) | 1u;
        b = (uint32_t)std::atoi(pos[1].c_str());
    } else {
        a = 5;
        b = pos.size() == 1 ? (uint32_t)std::atoi(pos[0].c_str()) : 2627u;
    }
    if (a < 5) a = 5;
    if (a > b) b = a;
    *sumStart = a;
    *sumLimit = b;
}

uint64_t pmapSumLimit(uint64_t M) { return isqrt64((M - 2) << 2) << 1; }

}  // namespace ffg
