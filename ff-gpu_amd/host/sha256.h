// SHA-256 (FIPS 180-4) for per-range output digests: the batched and served
// modes hash each range's stdout so a run of thousands of ranges can be
// compared with its expected text without writing the text out.
#ifndef FFG_SHA256_H
#define FFG_SHA256_H

#include <cstddef>
#include <cstdint>
#include <string>

namespace ffg {

class Sha256 {
public:
    Sha256();
    void update(const void *data, size_t n);
    void update(const std::string &s) { update(s.data(), s.size()); }
    // 32-byte digest; the object is spent afterwards.
    void final(uint8_t out[32]);
    std::string hex();

private:
    void block(const uint8_t *p);
    uint32_t h_[8];
    uint8_t buf_[64];
    size_t fill_ = 0;
    uint64_t bytes_ = 0;
};

std::string sha256Hex(const std::string &s);

}  // namespace ffg

#endif
