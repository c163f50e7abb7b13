// One-bit retrieval of the leaf orientation: stores, for every key (fingerprint .second), the bit
// that says which of its two candidate cells the key occupies. It plays the role of the reference's
// SimpleRibbon<1, ...> (include/ShockHash.h:152-153,657): query answers depend only on the bits
// returned for keys of the set, which this structure returns exactly.
//
// Construction: a random 3-hypergraph XOR system (one cell in each third of ~1.23 n cells),
// solved by peeling; a failed peeling (rare) retries with the next seed. The class templates build it on the
// device (shgpu_retrieval_build, same layout) and load the result with the (seed, seg, words) constructor;
// the host construction remains for the tests that compare the two.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "Hashing.h"

namespace shockhash {

class BitRetrieval {
    uint64_t seed = 0;
    uint64_t seg = 1;  // cells per third
    std::vector<uint64_t> bits;

    void cells(uint64_t key, uint64_t c[3]) const {
        const uint64_t h0 = remix(key + seed * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull);
        const uint64_t h1 = remix(h0 ^ 0xd6e8feb86659fd93ull);
        c[0] = fastrange64(h0, seg);
        c[1] = seg + fastrange64(h1, seg);
        c[2] = 2 * seg + fastrange64(remix(h1 + 0x8cb92ba72f3d8dd7ull), seg);
    }
    bool get(uint64_t c) const { return (bits[c / 64] >> (c % 64)) & 1; }
    void flip(uint64_t c) { bits[c / 64] ^= uint64_t(1) << (c % 64); }

    bool try_build(const std::vector<std::pair<uint64_t, uint8_t>> &in) {
        const uint64_t ncell = 3 * seg;
        std::vector<uint32_t> deg(ncell, 0);
        std::vector<uint64_t> xr(ncell, 0);  // xor of key indices per cell
        uint64_t c[3];
        for (uint64_t i = 0; i < in.size(); i++) {
            cells(in[i].first, c);
            for (int k = 0; k < 3; k++) { deg[c[k]]++; xr[c[k]] ^= i; }
        }
        std::vector<uint64_t> stack;  // (key index, cell) in peeling order
        stack.reserve(2 * in.size());
        std::vector<uint64_t> q;
        for (uint64_t v = 0; v < ncell; v++) if (deg[v] == 1) q.push_back(v);
        while (!q.empty()) {
            const uint64_t v = q.back();
            q.pop_back();
            if (deg[v] != 1) continue;
            const uint64_t i = xr[v];
            stack.push_back(i);
            stack.push_back(v);
            cells(in[i].first, c);
            for (int k = 0; k < 3; k++) {
                deg[c[k]]--;
                xr[c[k]] ^= i;
                if (deg[c[k]] == 1) q.push_back(c[k]);
            }
        }
        if (stack.size() != 2 * in.size()) return false;
        bits.assign((ncell + 63) / 64, 0);
        for (size_t s = stack.size(); s > 0; s -= 2) {
            const uint64_t i = stack[s - 2], v = stack[s - 1];
            cells(in[i].first, c);
            const bool cur = get(c[0]) ^ get(c[1]) ^ get(c[2]);
            if (cur != (bool)(in[i].second & 1)) flip(v);
        }
        return true;
    }

  public:
    BitRetrieval() = default;
    // A structure built elsewhere (shgpu_retrieval_build): its seed, cells per third and packed cell bits.
    BitRetrieval(uint64_t seed_, uint64_t seg_, std::vector<uint64_t> words) : seed(seed_), seg(seg_), bits(std::move(words)) {
        if (bits.size() != (3 * seg + 63) / 64) throw std::invalid_argument("BitRetrieval: word count does not match seg");
    }
    explicit BitRetrieval(const std::vector<std::pair<uint64_t, uint8_t>> &in) {
        seg = (uint64_t)(in.size() * 1.23 / 3.0) + 8;
        for (seed = 0; seed < 64; seed++)
            if (try_build(in)) return;
        throw std::runtime_error("BitRetrieval: peeling failed (duplicate keys?)");
    }
    uint64_t retrieve(uint64_t key) const {
        uint64_t c[3];
        cells(key, c);
        return (uint64_t)(get(c[0]) ^ get(c[1]) ^ get(c[2]));
    }
    size_t sizeBytes() const { return bits.size() * 8 + 16; }
};

}  // namespace shockhash
