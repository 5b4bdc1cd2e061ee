#include "workload.h"

#include <cmath>

namespace ffg {

static uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

Rng::Rng(uint64_t seed)
{
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(seed);
}

uint64_t Rng::next()
{
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

double Rng::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

std::vector<SumRange> batchRanges(uint64_t seed, uint32_t count, uint32_t startMax, double wLo, double wHi)
{
    Rng r(seed);
    std::vector<SumRange> out;
    out.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t start = 5 + (uint32_t)(r.uniform() * (double)(startMax - 5));
        const uint32_t width = (uint32_t)std::floor(std::exp2(wLo + (wHi - wLo) * r.uniform()));
        out.push_back({start, start + width - 1});
    }
    return out;
}

std::vector<Request> servedRequests(uint64_t seed, uint32_t count, uint32_t limLo, uint32_t limHi, double wLo,
                                    double wHi, double rate)
{
    Rng r(seed);
    std::vector<Request> out;
    out.reserve(count);
    double t = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t lim = limLo + (uint32_t)(r.uniform() * (double)(limHi - limLo + 1));
        const uint32_t width = (uint32_t)std::floor(std::exp2(wLo + (wHi - wLo) * r.uniform()));
        const uint32_t start = lim + 1 > width + 5 ? lim - width + 1 : 5;
        t += -std::log(1.0 - r.uniform()) / rate;
        out.push_back({start, lim, t});
    }
    return out;
}

}  // namespace ffg
