// Hash primitives of the ShockHash host side: 128-bit fingerprints, the mixing and range-reduction
// functions the reference takes from sux / ByteHamster util, and SpookyHash V2 for string keys.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace shockhash {

// 128-bit fingerprint (sux::function::hash128_t in the reference).
struct hash128_t {
    uint64_t first, second;
    hash128_t() = default;
    constexpr hash128_t(uint64_t f, uint64_t s) : first(f), second(s) {}
    bool operator==(const hash128_t &o) const { return first == o.first && second == o.second; }
};

// Stafford Mix13 (include/shockhash/SimdUtils.h:42-46).
inline uint64_t remix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint32_t fastrange32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }
inline uint64_t fastrange64(uint64_t x, uint64_t n) { return (uint64_t)(((unsigned __int128)x * n) >> 64); }
// sux::remap16: 48-bit fraction scaled to [0, n).
inline uint64_t remap16(uint64_t x, uint64_t n) { return ((x & ((1ull << 48) - 1)) * n) >> 48; }
// sux::remap128: bucket of a fingerprint's first half.
inline uint64_t remap128(uint64_t x, uint64_t n) { return (uint64_t)(((unsigned __int128)x * n) >> 64); }
// MurmurHash3 fmix32 and the SIMDShockHash split hash (include/SIMDShockHash.hpp:76-83,762-766).
inline uint32_t remix32(uint32_t z) {
    z ^= z >> 16; z *= 0x85ebca6bu; z ^= z >> 13; z *= 0xc2b2ae35u; z ^= z >> 16;
    return z;
}
inline uint32_t remixAndRemap(uint64_t x, uint32_t n) {
    return ((remix32((uint32_t)(x >> 32) ^ (uint32_t)x) & 0xFFFFu) * n) >> 16;
}

// Per-depth start seeds (include/ShockHash.h:47-49).
static constexpr uint64_t start_seed[20] = {
    0x106393c187cae21a, 0x6453cec3f7376937, 0x643e521ddbd2be98, 0x3740c6412f6572cb, 0x717d47562f1ce470,
    0x4cd6eb4c63befb7c, 0x9bfd8c5e18c8da73, 0x082f20e10092a9a3, 0x2ada2ce68d21defc, 0xe33cb4f3e7c6466b,
    0x3980be458c509c59, 0xc466fd9584828e8c, 0x45f0aabe1a61ede6, 0xf6e7b8b33ad9b98d, 0x4ef95e25f4b4983d,
    0x81175195173b92d3, 0x4e50927d8dd15978, 0x1ea2099d1fafae7f, 0x425c8a06fbaaa815, 0xcd4216006c74052a};
static constexpr int NUM_START_SEEDS = 20;

// Seed-pair codes (include/shockhash/PairingFunction.h:47-96).
inline uint64_t pairTriangular(uint64_t x, uint64_t y) { return (x - 1) * x / 2 + y; }
inline void unpairTriangular(uint64_t z, uint64_t &x, uint64_t &y) {
    x = (uint64_t)__builtin_floor(0.5 + __builtin_sqrt(0.25 + 2.0 * (double)z));
    y = z - ((x - 1) * x) / 2;
}
inline uint64_t pairElegant(uint64_t x, uint64_t y) { return x < y ? y * y + x : x * x + x + y; }
inline void unpairElegant(uint64_t z, uint64_t &x, uint64_t &y) {
    const uint64_t f = (uint64_t)__builtin_floor(__builtin_sqrt((double)z));
    const uint64_t r = f * f;
    if (z - r < f) { x = z - r; y = f; } else { x = f; y = z - r - f; }
}

// ---- SpookyHash V2 (Bob Jenkins, public domain), 128-bit hash of messages shorter than 192 bytes
// plus the long-message path for completeness. sux's spooky(k, len, seed) seeds both state words
// with `seed` and returns {h1, h0}; ShockHash's first_hash uses seed 0 (include/ShockHash.h:129).
namespace spooky_detail {
inline uint64_t rot64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
inline uint64_t ld64(const uint8_t *p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
inline uint64_t ld32(const uint8_t *p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
constexpr uint64_t SC_CONST = 0xdeadbeefdeadbeefull;

inline void short_mix(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h2 = rot64(h2, 50); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 52); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 30); h0 += h1; h2 ^= h0;  h1 = rot64(h1, 41); h1 += h2; h3 ^= h1;
    h2 = rot64(h2, 54); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 48); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 38); h0 += h1; h2 ^= h0;  h1 = rot64(h1, 37); h1 += h2; h3 ^= h1;
    h2 = rot64(h2, 62); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 34); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 5);  h0 += h1; h2 ^= h0;  h1 = rot64(h1, 36); h1 += h2; h3 ^= h1;
}
inline void short_end(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h3 ^= h2; h2 = rot64(h2, 15); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 52); h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 26); h1 += h0;  h2 ^= h1; h1 = rot64(h1, 51); h2 += h1;
    h3 ^= h2; h2 = rot64(h2, 28); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 9);  h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 47); h1 += h0;  h2 ^= h1; h1 = rot64(h1, 54); h2 += h1;
    h3 ^= h2; h2 = rot64(h2, 32); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 25); h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 63); h1 += h0;
}
inline void short_hash(const uint8_t *p, size_t len, uint64_t &o1, uint64_t &o2) {
    size_t rem = len % 32;
    uint64_t a = o1, b = o2, c = SC_CONST, d = SC_CONST;
    if (len > 15) {
        const uint8_t *end = p + (len / 32) * 32;
        for (; p < end; p += 32) { c += ld64(p); d += ld64(p + 8); short_mix(a, b, c, d); a += ld64(p + 16); b += ld64(p + 24); }
        if (rem >= 16) { c += ld64(p); d += ld64(p + 8); short_mix(a, b, c, d); p += 16; rem -= 16; }
    }
    d += ((uint64_t)len) << 56;
    switch (rem) {
        case 15: d += ((uint64_t)p[14]) << 48; [[fallthrough]];
        case 14: d += ((uint64_t)p[13]) << 40; [[fallthrough]];
        case 13: d += ((uint64_t)p[12]) << 32; [[fallthrough]];
        case 12: d += ld32(p + 8); c += ld64(p); break;
        case 11: d += ((uint64_t)p[10]) << 16; [[fallthrough]];
        case 10: d += ((uint64_t)p[9]) << 8; [[fallthrough]];
        case 9: d += (uint64_t)p[8]; [[fallthrough]];
        case 8: c += ld64(p); break;
        case 7: c += ((uint64_t)p[6]) << 48; [[fallthrough]];
        case 6: c += ((uint64_t)p[5]) << 40; [[fallthrough]];
        case 5: c += ((uint64_t)p[4]) << 32; [[fallthrough]];
        case 4: c += ld32(p); break;
        case 3: c += ((uint64_t)p[2]) << 16; [[fallthrough]];
        case 2: c += ((uint64_t)p[1]) << 8; [[fallthrough]];
        case 1: c += (uint64_t)p[0]; break;
        case 0: c += SC_CONST; d += SC_CONST;
    }
    short_end(a, b, c, d);
    o1 = a; o2 = b;
}
constexpr size_t NUM_VARS = 12, BLOCK_SIZE = NUM_VARS * 8, BUF_SIZE = 2 * BLOCK_SIZE;
inline void mix(const uint64_t *data, uint64_t *s) {
    static const int R[12] = {11, 32, 43, 31, 17, 28, 39, 57, 55, 54, 22, 46};
    for (int i = 0; i < 12; i++) {
        s[i] += data[i];
        s[(i + 2) % 12] ^= s[(i + 10) % 12];
        s[(i + 11) % 12] ^= s[i];
        s[i] = rot64(s[i], R[i]);
        s[(i + 11) % 12] += s[(i + 1) % 12];
    }
}
inline void end_partial(uint64_t *h) {
    static const int R[12] = {44, 15, 34, 21, 38, 33, 10, 13, 38, 53, 42, 54};
    for (int i = 0; i < 12; i++) {
        h[(i + 11) % 12] += h[(i + 1) % 12];
        h[(i + 2) % 12] ^= h[(i + 11) % 12];
        h[(i + 1) % 12] = rot64(h[(i + 1) % 12], R[i]);
    }
}
inline void end(const uint64_t *data, uint64_t *h) {
    for (int i = 0; i < 12; i++) h[i] += data[i];
    end_partial(h); end_partial(h); end_partial(h);
}
}  // namespace spooky_detail

// SpookyHash::Hash128(message, length, &hash1, &hash2).
inline void spookyHash128(const void *message, size_t length, uint64_t *hash1, uint64_t *hash2) {
    using namespace spooky_detail;
    const uint8_t *p = (const uint8_t *)message;
    if (length < BUF_SIZE) { short_hash(p, length, *hash1, *hash2); return; }
    uint64_t h[12], buf[NUM_VARS];
    h[0] = h[3] = h[6] = h[9] = *hash1;
    h[1] = h[4] = h[7] = h[10] = *hash2;
    h[2] = h[5] = h[8] = h[11] = SC_CONST;
    const uint8_t *end = p + (length / BLOCK_SIZE) * BLOCK_SIZE;
    for (; p < end; p += BLOCK_SIZE) { std::memcpy(buf, p, BLOCK_SIZE); mix(buf, h); }
    const size_t rem = length - (size_t)(end - (const uint8_t *)message);
    uint8_t last[BLOCK_SIZE];
    std::memset(last, 0, BLOCK_SIZE);
    std::memcpy(last, end, rem);
    last[BLOCK_SIZE - 1] = (uint8_t)rem;
    std::memcpy(buf, last, BLOCK_SIZE);
    spooky_detail::end(buf, h);
    *hash1 = h[0];
    *hash2 = h[1];
}

// sux::function::spooky(data, length, seed) -> {h1, h0}.
inline hash128_t spooky(const void *data, size_t length, uint64_t seed) {
    uint64_t h0 = seed, h1 = seed;
    spookyHash128(data, length, &h0, &h1);
    return hash128_t(h1, h0);
}

}  // namespace shockhash
