// Common body of the three drop-in class templates: construction through the MI355X engine, host
// encoding, and the RecSplit query walk.
#pragma once
#include <chrono>
#include <cstdint>
#include <iostream>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "RecSplitHost.h"

namespace shockhash {
namespace detail {

// ROT: ROTATION_FITTING of ShockHash / SIMDShockHash (full leaves carry x * LEAF_SIZE + r; include/ShockHash.h:
// 251-265,316-403, include/SIMDShockHash.hpp:419-506,826-840).
template <int VARIANT, size_t LEAF_SIZE, bool ROT = false>
class GpuRecSplit {
  protected:
    static constexpr bool BIPARTITE = VARIANT == SHGPU_SHOCKHASH2;
    static_assert(!(ROT && BIPARTITE), "rotation fitting applies to ShockHash and SIMDShockHash leaves");
    static constexpr uint64_t GROUP_A_HF_MASK = LEAF_SIZE > 32 ? 7u : 0u;  // include/ShockHash.h:141
    static constexpr size_t _leaf = LEAF_SIZE;
    static constexpr size_t lower_aggr = Splitting<LEAF_SIZE, BIPARTITE>::lower_aggr;
    static constexpr size_t upper_aggr = Splitting<LEAF_SIZE, BIPARTITE>::upper_aggr;
    static constexpr std::array<uint64_t, MAX_BUCKET_SIZE> memo = golomb_memo<LEAF_SIZE, BIPARTITE>();
    static int golomb_param(size_t m) { return (int)(memo[m] >> 27); }
    static uint64_t skip_bits(size_t m) { return memo[m] & 0xFFFF; }
    static uint64_t skip_nodes(size_t m) { return (memo[m] >> 16) & 0x7FF; }

    size_t bucket_size = 0;
    size_t nbuckets = 0;
    size_t keys_count = 0;
    RiceBitVector descriptors;
    DoubleEF ef;
    RibbonRetrieval ribbon;
    std::vector<uint64_t> ribbon_keys;  // retrieval input, kept for verification when gpu::keep_retrieval_input
    std::vector<uint8_t> ribbon_bits;
    gpu::Forest forest_meta;                                  // codes and nodes, kept for verification
    double host_ms = 0;
    double phase_ms[3] = {0, 0, 0};  // Golomb-Rice + Elias-Fano, leaf orientation (device), retrieval (device)

    void construct(const hash128_t *h, size_t n, size_t bsize, size_t num_threads) {
        construct_forest(gpu::Engine::instance().build(VARIANT, (uint32_t)LEAF_SIZE, h, n, bsize, ROT), n, bsize, num_threads);
    }

    // String keys: fingerprinted on the device (first_hash, include/ShockHash.h:129,159-185).
    template <class Strings>
    void construct_strings(const Strings &keys, size_t bsize, size_t num_threads) {
        construct_forest(gpu::Engine::instance().build_strings(VARIANT, (uint32_t)LEAF_SIZE, keys, bsize, ROT), keys.size(),
                         bsize, num_threads);
    }

    // num_threads (the reference's host parallelism) has no host phase left to split: the per-key work
    // (sort, seed search, orientation, retrieval) runs on the device; the output never depended on it.
    void construct_forest(gpu::Forest &&f, size_t n, size_t bsize, size_t /*num_threads*/) {
        bucket_size = bsize;
        keys_count = n;
        const auto t0 = std::chrono::steady_clock::now();
        nbuckets = f.nbuckets;
        std::vector<uint64_t> words, pos_acc;
        encodeForest(f, memo, words, pos_acc);
        descriptors = RiceBitVector(std::move(words));
        ef = DoubleEF(f.size_acc, pos_acc);
        phase_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<uint32_t> roffsets;
        std::vector<uint8_t> rseeds;
        std::vector<uint64_t> rwords;
        gpu::Engine::instance().orientation_retrieval(VARIANT, f, gpu::keep_retrieval_input, ribbon_keys, ribbon_bits,
                                                      roffsets, rseeds, rwords, phase_ms + 1, ROT ? (uint32_t)LEAF_SIZE : 0u);
        ribbon = RibbonRetrieval(std::move(roffsets), std::move(rseeds), std::move(rwords));
        f.device.reset();  // frees the build's device keys
        forest_meta = std::move(f);
        host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    uint64_t split_hash(uint64_t x, size_t m) const {
        if (VARIANT == SHGPU_SIMD_SHOCKHASH) return remixAndRemap(x, (uint32_t)m);
        return remap16(remix(x), m);
    }

  public:
    size_t query(const hash128_t &hash) const {
        const size_t bucket = remap128(hash.first, nbuckets);
        uint64_t cum_keys, cum_keys_next, bit_pos;
        ef.get(bucket, cum_keys, cum_keys_next, bit_pos);
        size_t m = cum_keys_next - cum_keys;
        if (m <= 1) return cum_keys;
        auto reader = descriptors.reader();
        reader.readReset(bit_pos, skip_bits(m));
        int level = 0;
        while (m > upper_aggr) {  // fanout 2
            const uint64_t d = reader.readNext(golomb_param(m));
            const uint64_t hmod = split_hash(hash.second + d + start_seed[level], m);
            const uint32_t split = ((uint16_t(m / 2 + upper_aggr - 1) / upper_aggr)) * upper_aggr;
            if (hmod < split) {
                m = split;
            } else {
                reader.skipSubtree(skip_nodes(split), skip_bits(split));
                m -= split;
                cum_keys += split;
            }
            level++;
        }
        if (m > lower_aggr) {
            const uint64_t d = reader.readNext(golomb_param(m));
            const uint64_t s = VARIANT == SHGPU_SIMD_SHOCKHASH ? start_seed[NUM_START_SEEDS - 3] : start_seed[level];
            const uint64_t hmod = split_hash(hash.second + d + s, m);
            const uint64_t part = uint16_t(hmod) / lower_aggr;
            m = std::min<size_t>(lower_aggr, m - part * lower_aggr);
            cum_keys += lower_aggr * part;
            if (part) reader.skipSubtree(skip_nodes(lower_aggr) * part, skip_bits(lower_aggr) * part);
            if (VARIANT != SHGPU_SIMD_SHOCKHASH) level++;
        }
        if (m > _leaf) {
            const uint64_t d = reader.readNext(golomb_param(m));
            const uint64_t s = VARIANT == SHGPU_SIMD_SHOCKHASH ? start_seed[NUM_START_SEEDS - 2] : start_seed[level];
            const uint64_t hmod = split_hash(hash.second + d + s, m);
            const uint64_t part = uint16_t(hmod) / _leaf;
            m = std::min<size_t>(_leaf, m - part * _leaf);
            cum_keys += _leaf * part;
            if (part) reader.skipSubtree(part, skip_bits(_leaf) * part);
            if (VARIANT != SHGPU_SIMD_SHOCKHASH) level++;
        }
        const uint64_t b = reader.readNext(golomb_param(m));
        const uint64_t retrieved = ribbon.retrieve(hash.second);
        if (BIPARTITE) return cum_keys + sh2Cell((uint32_t)m, b, hash.second, retrieved);
        if (m <= 1) return cum_keys;
        const uint64_t leaf_seed = VARIANT == SHGPU_SIMD_SHOCKHASH ? start_seed[NUM_START_SEEDS - 1] : start_seed[level];
        if (ROT && m == LEAF_SIZE) {  // include/ShockHash.h:251-265, include/SIMDShockHash.hpp:826-840
            const uint64_t r = b % LEAF_SIZE, x = b / LEAF_SIZE + leaf_seed;
            if ((hash.second & 1) == 0) {
                const CandidateCells c = sh1CandidateCells(hash.second, x & ~GROUP_A_HF_MASK, (uint32_t)m);
                return cum_keys + (retrieved == 0 ? c.cell1 : c.cell2);
            }
            const CandidateCells c = sh1CandidateCells(hash.second, x, (uint32_t)m);
            return cum_keys + ((retrieved == 0 ? c.cell1 : c.cell2) + r) % LEAF_SIZE;
        }
        const uint64_t seed = b + leaf_seed;
        const CandidateCells c = sh1CandidateCells(hash.second, seed, (uint32_t)m);
        return cum_keys + (retrieved == 0 ? c.cell1 : c.cell2);
    }

    size_t size() const { return keys_count; }
    size_t getBits() const {
        return ef.bitCountCumKeys() + ef.bitCountPosition() + descriptors.getBits() + 8 * ribbon.sizeBytes() +
               8 * sizeof(*this);
    }
    void printBits() const {
        std::cout << "EF 1:   " << (double)ef.bitCountCumKeys() / keys_count << std::endl;
        std::cout << "EF 2:   " << (double)ef.bitCountPosition() / keys_count << std::endl;
        std::cout << "trees:  " << (double)descriptors.getBits() / keys_count << std::endl;
        std::cout << "ribbon: " << (double)(8 * ribbon.sizeBytes()) / keys_count << std::endl;
    }

    // Serialized form of the whole function (not in the reference interface, whose only stream I/O is the
    // descriptor bit vector's, include/shockhash/RiceBitVector.h:133-141): a header (magic, format version,
    // variant (+ 0x100 for rotation fitting), leaf size, bucket size, bucket count, key count), the descriptor words, the two Elias-Fano
    // sequences and the retrieval structure, all little-endian. load() restores a function that answers every
    // query as the saved one; it throws std::runtime_error on a stream of another class or leaf size.
    static constexpr uint64_t FORMAT_MAGIC = 0x3146485048534853ull;  // "SHSHPHF1"
    void save(std::ostream &os) const {
        const uint64_t hdr[7] = {FORMAT_MAGIC, 1, (uint64_t)VARIANT | (ROT ? 0x100u : 0u), (uint64_t)LEAF_SIZE, bucket_size,
                                 nbuckets, keys_count};
        os.write((const char *)hdr, sizeof(hdr));
        descriptors.save(os);
        ef.save(os);
        ribbon.save(os);
        if (!os) throw std::runtime_error("ShockHash save: stream error");
    }
    void load(std::istream &is) {
        uint64_t hdr[7] = {0, 0, 0, 0, 0, 0, 0};
        is.read((char *)hdr, sizeof(hdr));
        if (!is || hdr[0] != FORMAT_MAGIC || hdr[1] != 1) throw std::runtime_error("ShockHash load: not a saved function");
        if (hdr[2] != ((uint64_t)VARIANT | (ROT ? 0x100u : 0u)) || hdr[3] != (uint64_t)LEAF_SIZE)
            throw std::runtime_error("ShockHash load: the stream holds another class or leaf size");
        bucket_size = hdr[4];
        nbuckets = hdr[5];
        keys_count = hdr[6];
        descriptors.load(is);
        ef.load(is);
        ribbon.load(is);
        ribbon_keys.clear();
        ribbon_bits.clear();
        forest_meta = gpu::Forest{};
    }

    // Verification accessors (not in the reference interface).
    const std::vector<uint64_t> &descriptorWords() const { return descriptors.words(); }
    const std::vector<uint64_t> &ribbonKeys() const { return ribbon_keys; }
    const std::vector<uint8_t> &ribbonBits() const { return ribbon_bits; }
    const gpu::Forest &forest() const { return forest_meta; }
    size_t bucketCount() const { return nbuckets; }
    double hostEncodeMs() const { return host_ms; }
    const double *hostPhaseMs() const { return phase_ms; }
};

}  // namespace detail
}  // namespace shockhash
