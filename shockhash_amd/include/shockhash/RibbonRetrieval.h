// One-bit retrieval of the leaf orientation: returns, for every key (fingerprint .second) of the set, the bit that
// says which of its two candidate cells the key occupies. It plays the role of the reference's SimpleRibbon<1, ...>
// (include/ShockHash.h:152-153,657): query answers depend only on the bits returned for keys of the set, which this
// structure returns exactly.
//
// The structure is a sharded ribbon built on the device (shgpu_ribbon_build, layout in include/shockhash_amd.h):
// keys are split into shards of ~2048 by a hash; in its shard a key has a start slot and a 64-bit coefficient row,
// and its bit is the parity of the row AND the 64 slot bits from the start. About 1.05 slots per key plus 40 bits
// per shard (slot offset and seed byte).
#pragma once
#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../../../shockhash_amd/csrc/shgpu_types.h"
#include "Hashing.h"

namespace shockhash {

class RibbonRetrieval {
    std::vector<uint32_t> offsets = {0, 0};  // nshards + 1 slot offsets
    std::vector<uint8_t> seeds = {0};        // seed byte per shard
    std::vector<uint64_t> words = {0};       // slot bits (slot c = bit c % 64 of word c / 64), one word of padding

  public:
    RibbonRetrieval() = default;
    // A structure built by shgpu_ribbon_build: its slot offsets, seed bytes and packed slot bits.
    RibbonRetrieval(std::vector<uint32_t> offs, std::vector<uint8_t> sds, std::vector<uint64_t> w)
        : offsets(std::move(offs)), seeds(std::move(sds)), words(std::move(w)) {
        if (offsets.size() != seeds.size() + 1 || seeds.empty()) throw std::invalid_argument("RibbonRetrieval: offsets/seeds");
        if (words.size() < ((uint64_t)offsets.back() + 63) / 64 + 1) throw std::invalid_argument("RibbonRetrieval: word count");
    }
    uint64_t retrieve(uint64_t key) const {
        const uint64_t h = remix(key + SHGPU_RIBBON_SALT_SHARD);
        const uint64_t s = fastrange64(h, seeds.size());
        const uint64_t base = offsets[s], m = offsets[s + 1] - base;
        if (m < SHGPU_RIBBON_WIDTH) return 0;  // an empty shard: no key of the set maps here
        const uint64_t g = remix(h + ((uint64_t)seeds[s] + 1) * SHGPU_RIBBON_SEED_STEP);
        const uint64_t p = base + fastrange64(g, m - SHGPU_RIBBON_WIDTH + 1);
        const uint64_t row = remix(g ^ SHGPU_RIBBON_SALT_ROW) | 1;
        const uint64_t w = p >> 6, b = p & 63;
        const uint64_t window = b ? (words[w] >> b) | (words[w + 1] << (64 - b)) : words[w];
        return (uint64_t)__builtin_popcountll(row & window) & 1;
    }
    size_t sizeBytes() const { return words.size() * 8 + offsets.size() * 4 + seeds.size(); }
    uint64_t shards() const { return seeds.size(); }

    // Serialized form: shard count, slot offsets, seed bytes, word count, words (little-endian u64 / u32 / u8).
    void save(std::ostream &os) const {
        const uint64_t ns = seeds.size(), nw = words.size();
        os.write((const char *)&ns, 8);
        os.write((const char *)offsets.data(), (std::streamsize)(offsets.size() * 4));
        os.write((const char *)seeds.data(), (std::streamsize)seeds.size());
        os.write((const char *)&nw, 8);
        os.write((const char *)words.data(), (std::streamsize)(nw * 8));
    }
    void load(std::istream &is) {
        uint64_t ns = 0, nw = 0;
        is.read((char *)&ns, 8);
        if (!is || ns == 0 || ns > (1ull << 32)) throw std::runtime_error("RibbonRetrieval: bad shard count");
        std::vector<uint32_t> offs(ns + 1);
        std::vector<uint8_t> sds(ns);
        is.read((char *)offs.data(), (std::streamsize)(offs.size() * 4));
        is.read((char *)sds.data(), (std::streamsize)ns);
        is.read((char *)&nw, 8);
        if (!is || nw > (1ull << 40)) throw std::runtime_error("RibbonRetrieval: bad word count");
        std::vector<uint64_t> w(nw);
        is.read((char *)w.data(), (std::streamsize)(nw * 8));
        if (!is) throw std::runtime_error("RibbonRetrieval: truncated stream");
        *this = RibbonRetrieval(std::move(offs), std::move(sds), std::move(w));
    }
};

}  // namespace shockhash
