// String keys to 128-bit fingerprints on the device: SpookyHash V2 (Bob Jenkins, public domain) 128-bit,
// seed 0, the fingerprint ShockHash's string constructors take (first_hash, include/ShockHash.h:129,159-185;
// sux::function::spooky(data, length, seed) returns {hash2, hash1} as {first, second}). One thread per key:
// messages under 192 bytes take SpookyHash's short path (32-byte rounds of ShortMix, ShortEnd), longer ones
// the 96-byte-block path (Mix per block, End over the zero-padded last block that ends in its length).
// Keys are a byte buffer and n + 1 offsets; bytes are read one by one (keys are not aligned).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/shockhash_amd.h"
#include "strings.h"

namespace {

constexpr uint64_t SC_CONST = 0xdeadbeefdeadbeefull;
constexpr uint64_t SHORT_LIMIT = 192;  // 2 * 12 * 8 bytes
constexpr uint32_t BLOCK_BYTES = 96;   // 12 state words

__device__ __forceinline__ uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Little-endian word of `len` (<= 8) bytes at p.
__device__ __forceinline__ uint64_t load_le(const uint8_t *p, uint32_t len) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < len; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

__device__ __forceinline__ void short_mix(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h2 = rotl(h2, 50); h2 += h3; h0 ^= h2;
    h3 = rotl(h3, 52); h3 += h0; h1 ^= h3;
    h0 = rotl(h0, 30); h0 += h1; h2 ^= h0;
    h1 = rotl(h1, 41); h1 += h2; h3 ^= h1;
    h2 = rotl(h2, 54); h2 += h3; h0 ^= h2;
    h3 = rotl(h3, 48); h3 += h0; h1 ^= h3;
    h0 = rotl(h0, 38); h0 += h1; h2 ^= h0;
    h1 = rotl(h1, 37); h1 += h2; h3 ^= h1;
    h2 = rotl(h2, 62); h2 += h3; h0 ^= h2;
    h3 = rotl(h3, 34); h3 += h0; h1 ^= h3;
    h0 = rotl(h0, 5);  h0 += h1; h2 ^= h0;
    h1 = rotl(h1, 36); h1 += h2; h3 ^= h1;
}

__device__ __forceinline__ void short_end(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h3 ^= h2; h2 = rotl(h2, 15); h3 += h2;
    h0 ^= h3; h3 = rotl(h3, 52); h0 += h3;
    h1 ^= h0; h0 = rotl(h0, 26); h1 += h0;
    h2 ^= h1; h1 = rotl(h1, 51); h2 += h1;
    h3 ^= h2; h2 = rotl(h2, 28); h3 += h2;
    h0 ^= h3; h3 = rotl(h3, 9);  h0 += h3;
    h1 ^= h0; h0 = rotl(h0, 47); h1 += h0;
    h2 ^= h1; h1 = rotl(h1, 54); h2 += h1;
    h3 ^= h2; h2 = rotl(h2, 32); h3 += h2;
    h0 ^= h3; h3 = rotl(h3, 25); h0 += h3;
    h1 ^= h0; h0 = rotl(h0, 63); h1 += h0;
}

// SpookyHash::Short with both seeds in (a, b).
__device__ void short_hash(const uint8_t *p, uint64_t len, uint64_t &a, uint64_t &b) {
    uint64_t c = SC_CONST, d = SC_CONST;
    uint64_t rem = len % 32;
    if (len > 15) {
        for (uint64_t r = 0; r < len / 32; r++, p += 32) {
            c += load_le(p, 8);
            d += load_le(p + 8, 8);
            short_mix(a, b, c, d);
            a += load_le(p + 16, 8);
            b += load_le(p + 24, 8);
        }
        if (rem >= 16) {
            c += load_le(p, 8);
            d += load_le(p + 8, 8);
            short_mix(a, b, c, d);
            p += 16;
            rem -= 16;
        }
    }
    // the last 0..15 bytes: bytes 0..7 go to c, bytes 8..14 to d (little-endian), length in d's top byte
    d += len << 56;
    if (rem == 0) {
        c += SC_CONST;
        d += SC_CONST;
    } else if (rem >= 8) {
        c += load_le(p, 8);
        d += load_le(p + 8, (uint32_t)rem - 8);
    } else {
        c += load_le(p, (uint32_t)rem);
    }
    short_end(a, b, c, d);
}

__device__ __forceinline__ void mix(const uint64_t (&data)[12], uint64_t (&s)[12]) {
    constexpr int R[12] = {11, 32, 43, 31, 17, 28, 39, 57, 55, 54, 22, 46};
#pragma unroll
    for (int i = 0; i < 12; i++) {
        s[i] += data[i];
        s[(i + 2) % 12] ^= s[(i + 10) % 12];
        s[(i + 11) % 12] ^= s[i];
        s[i] = rotl(s[i], R[i]);
        s[(i + 11) % 12] += s[(i + 1) % 12];
    }
}

__device__ __forceinline__ void end_partial(uint64_t (&h)[12]) {
    constexpr int R[12] = {44, 15, 34, 21, 38, 33, 10, 13, 38, 53, 42, 54};
#pragma unroll
    for (int i = 0; i < 12; i++) {
        h[(i + 11) % 12] += h[(i + 1) % 12];
        h[(i + 2) % 12] ^= h[(i + 11) % 12];
        h[(i + 1) % 12] = rotl(h[(i + 1) % 12], R[i]);
    }
}

// SpookyHash::Hash128 with both seeds in (a, b) for messages of at least 192 bytes.
__device__ void long_hash(const uint8_t *p, uint64_t len, uint64_t &a, uint64_t &b) {
    uint64_t h[12], data[12];
#pragma unroll
    for (int i = 0; i < 12; i += 3) {
        h[i] = a;
        h[i + 1] = b;
        h[i + 2] = SC_CONST;
    }
    const uint64_t blocks = len / BLOCK_BYTES;
    for (uint64_t k = 0; k < blocks; k++, p += BLOCK_BYTES) {
#pragma unroll
        for (int i = 0; i < 12; i++) data[i] = load_le(p + 8 * i, 8);
        mix(data, h);
    }
    const uint32_t rem = (uint32_t)(len - blocks * BLOCK_BYTES);
    // last block: the remaining bytes, zeros, and the remainder length in its last byte
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const uint32_t lo = 8 * i;
        data[i] = lo >= rem ? 0 : load_le(p + lo, rem - lo < 8 ? rem - lo : 8);
    }
    data[11] |= (uint64_t)rem << 56;
#pragma unroll
    for (int i = 0; i < 12; i++) h[i] += data[i];
    end_partial(h);
    end_partial(h);
    end_partial(h);
    a = h[0];
    b = h[1];
}

__global__ void k_hash_strings(const uint8_t *__restrict__ data, const uint64_t *__restrict__ offsets, uint64_t n,
                               uint64_t *__restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = offsets[i], len = offsets[i + 1] - o;
        uint64_t h1 = 0, h2 = 0;  // seed 0 in both state words
        if (len < SHORT_LIMIT) short_hash(data + o, len, h1, h2);
        else long_hash(data + o, len, h1, h2);
        out[2 * i] = h2;      // .first
        out[2 * i + 1] = h1;  // .second
    }
}

}  // namespace

int shgpu_hash_strings_impl(hipStream_t s, const uint8_t *d_data, const uint64_t *d_offsets, uint64_t n, uint64_t *d_hashes,
                            std::string &err) {
    if (n == 0) return SHGPU_OK;
    k_hash_strings<<<2048, 256, 0, s>>>(d_data, d_offsets, n, d_hashes);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("k_hash_strings: ") + hipGetErrorString(e);
        return SHGPU_ERR_HIP;
    }
    return SHGPU_OK;
}
