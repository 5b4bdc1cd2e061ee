#include "sha256.h"

#include <cstring>

namespace ffg {

namespace {

using u128 = unsigned __int128;

// The first 64 primes.
struct Primes64 {
    uint32_t p[64];
    Primes64()
    {
        int n = 0;
        for (uint32_t c = 2; n < 64; ++c) {
            bool prime = true;
            for (uint32_t d = 2; d * d <= c; ++d)
                if (c % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) p[n++] = c;
        }
    }
};

// floor(r-th root of v) for r = 2, 3 by bisection (v < 2^120).
uint64_t iroot(u128 v, int r)
{
    uint64_t lo = 0, hi = 1ull << (r == 2 ? 61 : 41);
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        u128 x = mid, pw = x * x;
        if (r == 3) pw *= x;
        if (pw <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// FIPS 180-4 section 4.2.2 and 5.3.3: the round constants are the first 32
// bits of the fractional parts of the cube roots of the first 64 primes,
// the initial hash value those of the square roots of the first 8.  Both
// are derived here exactly in integer arithmetic: floor(cbrt(p * 2^96)) and
// floor(sqrt(p * 2^64)) keep 32 fraction bits, and the low 32 bits of each
// are the constant.
struct Constants {
    uint32_t k[64], h0[8];
    Constants()
    {
        const Primes64 P;
        for (int i = 0; i < 64; ++i) k[i] = (uint32_t)iroot((u128)P.p[i] << 96, 3);
        for (int i = 0; i < 8; ++i) h0[i] = (uint32_t)iroot((u128)P.p[i] << 64, 2);
    }
};

const Constants &constants()
{
    static const Constants c;
    return c;
}

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

}  // namespace

Sha256::Sha256() { std::memcpy(h_, constants().h0, sizeof h_); }

void Sha256::block(const uint8_t *p)
{
    const uint32_t *K = constants().k;
    uint32_t w[64];
    for (int t = 0; t < 16; ++t)
        w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 | p[4 * t + 3];
    for (int t = 16; t < 64; ++t) {
        const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
        const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
    for (int t = 0; t < 64; ++t) {
        const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K[t] + w[t];
        const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    h_[0] += a;
    h_[1] += b;
    h_[2] += c;
    h_[3] += d;
    h_[4] += e;
    h_[5] += f;
    h_[6] += g;
    h_[7] += h;
}

void Sha256::update(const void *data, size_t n)
{
    const uint8_t *p = static_cast<const uint8_t *>(data);
    bytes_ += n;
    if (fill_) {
        const size_t take = n < 64 - fill_ ? n : 64 - fill_;
        std::memcpy(buf_ + fill_, p, take);
        fill_ += take;
        p += take;
        n -= take;
        if (fill_ < 64) return;
        block(buf_);
        fill_ = 0;
    }
    for (; n >= 64; p += 64, n -= 64) block(p);
    std::memcpy(buf_, p, n);
    fill_ = n;
}

void Sha256::final(uint8_t out[32])
{
    // padding: 0x80, zeros, then the message length in bits, big-endian
    const uint64_t bits = bytes_ * 8;
    const uint8_t one = 0x80, zero = 0;
    update(&one, 1);
    while (fill_ != 56) update(&zero, 1);
    uint8_t len[8];
    for (int i = 0; i < 8; ++i) len[i] = (uint8_t)(bits >> (56 - 8 * i));
    update(len, 8);
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h_[i] >> (24 - 8 * j));
}

std::string Sha256::hex()
{
    uint8_t d[32];
    final(d);
    static const char digits[] = "0123456789abcdef";
    std::string s(64, '0');
    for (int i = 0; i < 32; ++i) {
        s[2 * i] = digits[d[i] >> 4];
        s[2 * i + 1] = digits[d[i] & 15];
    }
    return s;
}

std::string sha256Hex(const std::string &s)
{
    Sha256 h;
    h.update(s);
    return h.hex();
}

}  // namespace ffg
