// host_md5.hpp -- RFC 1321 MD5 on the host, for the sequential hashes that
// stay on the CPU: hash_full / hash16k (src/source.rs:223-228), file ids
// (src/source.rs:303-309) and the small critical packets (src/format.rs:40-44).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>

namespace p2g {

class HostMd5 {
  public:
    HostMd5() { reset(); }

    void reset() {
        h_[0] = 0x67452301u; h_[1] = 0xefcdab89u; h_[2] = 0x98badcfeu; h_[3] = 0x10325476u;
        total_ = 0;
        fill_ = 0;
    }

    void update(const void *data, size_t len) {
        const uint8_t *p = static_cast<const uint8_t *>(data);
        total_ += len;
        if (fill_) {
            size_t take = 64 - fill_;
            if (take > len) take = len;
            memcpy(buf_ + fill_, p, take);
            fill_ += take;
            p += take;
            len -= take;
            if (fill_ == 64) { compress(buf_); fill_ = 0; }
        }
        while (len >= 64) { compress(p); p += 64; len -= 64; }
        if (len) { memcpy(buf_, p, len); fill_ = len; }
    }

    void finish(uint8_t out[16]) {
        const uint64_t bits = total_ * 8;
        uint8_t pad[72] = {0x80};
        update(pad, fill_ < 56 ? 56 - fill_ : 120 - fill_);
        uint8_t l[8];
        for (int i = 0; i < 8; i++) l[i] = (uint8_t)(bits >> (8 * i));
        update(l, 8);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(h_[i] >> (8 * j));
    }

    static void digest(const void *data, size_t len, uint8_t out[16]) {
        HostMd5 m;
        m.update(data, len);
        m.finish(out);
    }

  private:
    static inline uint32_t rotl(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

    void compress(const uint8_t *p) {
        uint32_t x[16];
        memcpy(x, p, 64);  // little-endian host
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3];
#define P2G_R(f, a, b, c, d, k, s, t) a = b + rotl(a + f(b, c, d) + x[k] + (t), s)
#define P2G_F(b, c, d) ((d) ^ ((b) & ((c) ^ (d))))
#define P2G_G(b, c, d) ((c) ^ ((d) & ((b) ^ (c))))
#define P2G_H(b, c, d) ((b) ^ (c) ^ (d))
#define P2G_I(b, c, d) ((c) ^ ((b) | ~(d)))
        // round 1
        P2G_R(P2G_F, a, b, c, d, 0, 7, 0xd76aa478u);  P2G_R(P2G_F, d, a, b, c, 1, 12, 0xe8c7b756u);
        P2G_R(P2G_F, c, d, a, b, 2, 17, 0x242070dbu); P2G_R(P2G_F, b, c, d, a, 3, 22, 0xc1bdceeeu);
        P2G_R(P2G_F, a, b, c, d, 4, 7, 0xf57c0fafu);  P2G_R(P2G_F, d, a, b, c, 5, 12, 0x4787c62au);
        P2G_R(P2G_F, c, d, a, b, 6, 17, 0xa8304613u); P2G_R(P2G_F, b, c, d, a, 7, 22, 0xfd469501u);
        P2G_R(P2G_F, a, b, c, d, 8, 7, 0x698098d8u);  P2G_R(P2G_F, d, a, b, c, 9, 12, 0x8b44f7afu);
        P2G_R(P2G_F, c, d, a, b, 10, 17, 0xffff5bb1u); P2G_R(P2G_F, b, c, d, a, 11, 22, 0x895cd7beu);
        P2G_R(P2G_F, a, b, c, d, 12, 7, 0x6b901122u); P2G_R(P2G_F, d, a, b, c, 13, 12, 0xfd987193u);
        P2G_R(P2G_F, c, d, a, b, 14, 17, 0xa679438eu); P2G_R(P2G_F, b, c, d, a, 15, 22, 0x49b40821u);
        // round 2
        P2G_R(P2G_G, a, b, c, d, 1, 5, 0xf61e2562u);  P2G_R(P2G_G, d, a, b, c, 6, 9, 0xc040b340u);
        P2G_R(P2G_G, c, d, a, b, 11, 14, 0x265e5a51u); P2G_R(P2G_G, b, c, d, a, 0, 20, 0xe9b6c7aau);
        P2G_R(P2G_G, a, b, c, d, 5, 5, 0xd62f105du);  P2G_R(P2G_G, d, a, b, c, 10, 9, 0x02441453u);
        P2G_R(P2G_G, c, d, a, b, 15, 14, 0xd8a1e681u); P2G_R(P2G_G, b, c, d, a, 4, 20, 0xe7d3fbc8u);
        P2G_R(P2G_G, a, b, c, d, 9, 5, 0x21e1cde6u);  P2G_R(P2G_G, d, a, b, c, 14, 9, 0xc33707d6u);
        P2G_R(P2G_G, c, d, a, b, 3, 14, 0xf4d50d87u); P2G_R(P2G_G, b, c, d, a, 8, 20, 0x455a14edu);
        P2G_R(P2G_G, a, b, c, d, 13, 5, 0xa9e3e905u); P2G_R(P2G_G, d, a, b, c, 2, 9, 0xfcefa3f8u);
        P2G_R(P2G_G, c, d, a, b, 7, 14, 0x676f02d9u); P2G_R(P2G_G, b, c, d, a, 12, 20, 0x8d2a4c8au);
        // round 3
        P2G_R(P2G_H, a, b, c, d, 5, 4, 0xfffa3942u);  P2G_R(P2G_H, d, a, b, c, 8, 11, 0x8771f681u);
        P2G_R(P2G_H, c, d, a, b, 11, 16, 0x6d9d6122u); P2G_R(P2G_H, b, c, d, a, 14, 23, 0xfde5380cu);
        P2G_R(P2G_H, a, b, c, d, 1, 4, 0xa4beea44u);  P2G_R(P2G_H, d, a, b, c, 4, 11, 0x4bdecfa9u);
        P2G_R(P2G_H, c, d, a, b, 7, 16, 0xf6bb4b60u); P2G_R(P2G_H, b, c, d, a, 10, 23, 0xbebfbc70u);
        P2G_R(P2G_H, a, b, c, d, 13, 4, 0x289b7ec6u); P2G_R(P2G_H, d, a, b, c, 0, 11, 0xeaa127fau);
        P2G_R(P2G_H, c, d, a, b, 3, 16, 0xd4ef3085u); P2G_R(P2G_H, b, c, d, a, 6, 23, 0x04881d05u);
        P2G_R(P2G_H, a, b, c, d, 9, 4, 0xd9d4d039u);  P2G_R(P2G_H, d, a, b, c, 12, 11, 0xe6db99e5u);
        P2G_R(P2G_H, c, d, a, b, 15, 16, 0x1fa27cf8u); P2G_R(P2G_H, b, c, d, a, 2, 23, 0xc4ac5665u);
        // round 4
        P2G_R(P2G_I, a, b, c, d, 0, 6, 0xf4292244u);  P2G_R(P2G_I, d, a, b, c, 7, 10, 0x432aff97u);
        P2G_R(P2G_I, c, d, a, b, 14, 15, 0xab9423a7u); P2G_R(P2G_I, b, c, d, a, 5, 21, 0xfc93a039u);
        P2G_R(P2G_I, a, b, c, d, 12, 6, 0x655b59c3u); P2G_R(P2G_I, d, a, b, c, 3, 10, 0x8f0ccc92u);
        P2G_R(P2G_I, c, d, a, b, 10, 15, 0xffeff47du); P2G_R(P2G_I, b, c, d, a, 1, 21, 0x85845dd1u);
        P2G_R(P2G_I, a, b, c, d, 8, 6, 0x6fa87e4fu);  P2G_R(P2G_I, d, a, b, c, 15, 10, 0xfe2ce6e0u);
        P2G_R(P2G_I, c, d, a, b, 6, 15, 0xa3014314u); P2G_R(P2G_I, b, c, d, a, 13, 21, 0x4e0811a1u);
        P2G_R(P2G_I, a, b, c, d, 4, 6, 0xf7537e82u);  P2G_R(P2G_I, d, a, b, c, 11, 10, 0xbd3af235u);
        P2G_R(P2G_I, c, d, a, b, 2, 15, 0x2ad7d2bbu); P2G_R(P2G_I, b, c, d, a, 9, 21, 0xeb86d391u);
#undef P2G_R
#undef P2G_F
#undef P2G_G
#undef P2G_H
#undef P2G_I
        h_[0] += a; h_[1] += b; h_[2] += c; h_[3] += d;
    }

    uint32_t h_[4];
    uint64_t total_;
    uint8_t buf_[64];
    size_t fill_;
};

}  // namespace p2g
