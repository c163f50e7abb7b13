// shockhash::ShockHash2Flat<k> backed by the MI355X engine: the k-perfect "flat" ShockHash2 variant
// (include/ShockHash2Flat.h:37-302). Keys are hashed to 64-bit values (mhc), bucketed into N / k buckets of
// exactly k keys by two threshold layers that bump overflowing keys, the bumped keys placed through a ShockHash2<k>
// fallback over their decimal strings and a free-position map, and each bucket solved by ShockHash2's leaf search
// (BijectionsShockHash2<k, true, QuadSplitCandidateFinderList>). The per-bucket searches, their orientation and
// the retrieval structure run on the GPU (one batched leaf search for all buckets); the bucketing runs on the host.
#pragma once
#include <array>
#include <charconv>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <map>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "ShockHash2.h"
#include "shockhash/Parallel.h"

namespace shockhash {

namespace flat_detail {

// MurmurHash64A (Austin Appleby, public domain). The reference hashes string keys with the un-vendored
// bytehamster::util::MurmurHash64 (ShockHash2Flat.h:75, 258-260); this is the published 64-bit MurmurHash2 with
// seed 0, so string keys are placed as by that function only if it is this one.
inline uint64_t murmurHash64A(const void *key, size_t len, uint64_t seed = 0) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = seed ^ (len * m);
    const unsigned char *p = (const unsigned char *)key;
    const size_t blocks = len / 8;
    for (size_t i = 0; i < blocks; i++) {
        uint64_t k;
        std::memcpy(&k, p + 8 * i, 8);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    const unsigned char *t = p + 8 * blocks;
    const size_t rest = len & 7;
    if (rest) {
        for (size_t i = rest; i-- > 0;) h ^= (uint64_t)t[i] << (8 * i);
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

constexpr size_t floorLog2(size_t x) { return x <= 1 ? 0 : 1 + floorLog2(x >> 1); }

// THRESHOLD_MAPPING (_fill_mapping, ShockHash2Flat.h:17-35): compacted threshold index -> 32-bit threshold.
template <size_t RANGE>
constexpr std::array<uint32_t, RANGE> fillMapping() {
    constexpr uint64_t ONE_THIRD = 0xFFFFFFFFull / 3;
    std::array<uint32_t, RANGE> a{};
    if constexpr (RANGE == 2) {
        a[0] = 0;
        a[1] = 0xFFFFFFFFu;
    } else if constexpr (RANGE > 2) {
        a[0] = 0;
        a[1] = (uint32_t)ONE_THIRD;
        constexpr uint64_t ir = RANGE - 3;
        for (uint64_t i = 0; i < ir; i++) a[2 + i] = (uint32_t)(2 * ONE_THIRD + ONE_THIRD * i / ir);
        a[RANGE - 1] = 0xFFFFFFFFu;
    }
    return a;
}

// Unsigned integers of a fixed bit width, packed (bytehamster::util::IntVector's role).
class PackedInts {
    std::vector<uint64_t> w;
    size_t width = 1, n = 0;

  public:
    void reset(size_t count, size_t bits) {
        n = count;
        width = bits;
        w.assign((count * bits + 63) / 64 + 1, 0);
    }
    uint64_t get(size_t i) const {
        const size_t b = i * width, q = b >> 6, o = b & 63;
        uint64_t v = w[q] >> o;
        if (o + width > 64) v |= w[q + 1] << (64 - o);
        return width == 64 ? v : v & ((1ull << width) - 1);
    }
    void set(size_t i, uint64_t v) {
        const size_t b = i * width, q = b >> 6, o = b & 63;
        const uint64_t mask = width == 64 ? ~0ull : (1ull << width) - 1;
        v &= mask;
        w[q] = (w[q] & ~(mask << o)) | (v << o);
        if (o + width > 64) {
            const size_t hi = o + width - 64;
            w[q + 1] = (w[q + 1] & ~((1ull << hi) - 1)) | (v >> (64 - o));
        }
    }
    size_t dataSizeBytes() const { return w.size() * 8; }
};

// freePositions as a bit vector with a one at i + freePositions[i] and select1 by sampled positions (the reference
// uses pasta::FlatRankSelect on the same bit vector, ShockHash2Flat.h:139-146, 290-291): select1(j + 1) - j is
// freePositions[j].
class FreePositions {
    std::vector<uint64_t> bits;
    std::vector<uint64_t> sampleWord, sampleRank;  // for ones j = 0, SAMPLE, 2 SAMPLE, ...: its word, ones before it
    size_t nbits = 0, ones = 0;
    static constexpr size_t SAMPLE = 256;

  public:
    void build(const std::vector<size_t> &fp) {
        ones = fp.size();
        nbits = fp.empty() ? 0 : fp.size() + fp.back() + 1;
        bits.assign((nbits + 63) / 64 + 1, 0);
        for (size_t i = 0; i < fp.size(); i++) {
            const size_t p = i + fp[i];
            bits[p >> 6] |= 1ull << (p & 63);
        }
        sampleWord.clear();
        sampleRank.clear();
        size_t rank = 0;
        for (size_t q = 0; q < bits.size(); q++) {
            const size_t c = (size_t)__builtin_popcountll(bits[q]);
            while (sampleWord.size() * SAMPLE < rank + c) {  // the next sampled one lies in word q
                sampleWord.push_back(q);
                sampleRank.push_back(rank);
            }
            rank += c;
        }
    }
    // select1(j + 1) - j, that is freePositions[j].
    size_t at(size_t j) const {
        if (j >= ones) throw std::out_of_range("ShockHash2Flat: free position index");
        size_t q = sampleWord[j / SAMPLE], rank = sampleRank[j / SAMPLE];
        for (;; q++) {
            uint64_t w = bits[q];
            const size_t c = (size_t)__builtin_popcountll(w);
            if (rank + c > j) {
                for (size_t t = rank; t < j; t++) w &= w - 1;
                return q * 64 + (size_t)__builtin_ctzll(w) - j;
            }
            rank += c;
        }
    }
    size_t sizeBits() const { return nbits; }
    size_t selectBytes() const { return sampleWord.size() * 16; }
};

}  // namespace flat_detail

template <size_t k>
class ShockHash2Flat {
    static_assert(k >= 10 && k <= (size_t)MAX_LEAF_SIZE2,
                  "the base case runs the QuadSplit candidate search, k >= 10 (leaf sizes up to 128)");
    static constexpr double OVERLOAD_FACTOR = 0.9;
    static constexpr size_t THRESHOLD_BITS = flat_detail::floorLog2(k) - 1;
    static constexpr size_t THRESHOLD_RANGE = 1ul << THRESHOLD_BITS;
    static constexpr std::array<uint32_t, THRESHOLD_RANGE> THRESHOLD_MAPPING = flat_detail::fillMapping<THRESHOLD_RANGE>();
    static size_t seedBits() { return (size_t)std::ceil(0.442 * k - 0.2); }  // SEED_BITS, evaluated as the reference
    static constexpr size_t SEED_FALLBACK_INDICATOR = 0;


    flat_detail::PackedInts thresholdsAndSeeds;
    std::map<size_t, size_t> seedsFallback;
    std::vector<size_t> layerBases;
    ShockHash2<k> fallbackPhf;
    bool haveFallback = false;
    size_t N = 0, nbuckets = 0, layers = 2, maxSeed = 0;
    std::vector<size_t> freePositionsList;
    flat_detail::FreePositions freePositions;
    RibbonRetrieval ribbon;
    std::vector<uint64_t> retrievalKeys;  // retrieval input, kept when gpu::keep_retrieval_input
    std::vector<uint8_t> retrievalBits;
    std::vector<uint64_t> bucketCodes;  // base-case seed per bucket, for verification
    // construction phases (ms): threshold layers and bumping, fallback function, placement of every key (free
    // positions, bucket of each key), base-case searches + orientation + retrieval on the device, and of that the
    // batched base-case search alone
    double phaseMs[5] = {0, 0, 0, 0, 0};

    void setThreshold(size_t b, size_t v) {
        uint64_t x = thresholdsAndSeeds.get(b);
        x &= ~((1ull << THRESHOLD_BITS) - 1);
        thresholdsAndSeeds.set(b, x | v);
    }
    void setSeed(size_t b, size_t v) {
        uint64_t x = thresholdsAndSeeds.get(b);
        x &= (1ull << THRESHOLD_BITS) - 1;
        thresholdsAndSeeds.set(b, x | ((uint64_t)v << THRESHOLD_BITS));
    }
    std::pair<size_t, size_t> getThresholdAndSeed(size_t b) const {
        const uint64_t x = thresholdsAndSeeds.get(b);
        return {x & (THRESHOLD_RANGE - 1), x >> THRESHOLD_BITS};
    }
    // The fallback function of a key's decimal form (std::to_string's digits), built in a buffer the thread keeps.
    size_t fallbackQuery(uint64_t mhc) const {
        thread_local std::string s;
        char buf[20];
        const auto r = std::to_chars(buf, buf + sizeof buf, mhc);
        s.assign(buf, r.ptr);
        return fallbackPhf(s);
    }

    // The reference's construction (ShockHash2Flat.h:66-171) with its per-key loops on the device: the threshold
    // layers (shgpu_flat_layers), the placement of every key and the bucket key lists (shgpu_flat_bucket_keys), the
    // base-case searches, orientation and retrieval. The host keeps what is per bucket or per bumped key: thresholds
    // and free positions, the fallback function over the bumped keys' decimal strings, and their placement.
    void construct(std::span<const uint64_t> mhcs) {
        using clk = std::chrono::steady_clock;
        const auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        auto t0 = clk::now();
        N = mhcs.size();
        nbuckets = N / k;
        maxSeed = 1ull << seedBits();
        const size_t keysInEndBucket = N - nbuckets * k;
        thresholdsAndSeeds.reset(nbuckets, THRESHOLD_BITS + seedBits());
        gpu::Engine &eng = gpu::Engine::instance();
        std::vector<uint8_t> thresholds;
        std::vector<uint32_t> freeCounts;
        std::vector<uint64_t> bases, bumpedMhc, bumpedIndex;
        eng.flat_layers(mhcs.data(), N, (uint32_t)k, THRESHOLD_MAPPING.data(), (uint32_t)THRESHOLD_RANGE, thresholds, freeCounts,
                        bases, bumpedMhc, bumpedIndex);
        layers = bases.size() - 1;
        layerBases.assign(bases.begin(), bases.end());
        for (size_t b = 0; b < layerBases.back(); b++) {  // flushBucket's thresholds and free positions, in bucket order
            setThreshold(b, thresholds[b]);
            for (uint32_t q = 0; q < freeCounts[b]; q++) freePositionsList.push_back(b);
        }
        auto t1 = clk::now();
        phaseMs[0] = ms(t0, t1);
        if (!bumpedMhc.empty()) {
            // the bumped keys' decimal forms (std::to_string's digits) side by side in one buffer: a std::string per
            // key would be a heap block (20 digits exceed the short-string buffer), allocated on the worker threads
            // and freed on this one, which took longer than the fallback's construction
            std::vector<char> digits(bumpedMhc.size() * 20);
            std::vector<std::string_view> content(bumpedMhc.size());
            parallelFor(bumpedMhc.size(), [&](size_t j) {
                char *p = digits.data() + 20 * j;
                const auto r = std::to_chars(p, p + 20, bumpedMhc[j]);
                content[j] = std::string_view(p, (size_t)(r.ptr - p));
            });
            fallbackPhf = ShockHash2<k>(std::span<const std::string_view>(content.data(), content.size()), 2000, 1);
            haveFallback = true;
        }
        auto t2 = clk::now();
        phaseMs[1] = ms(t1, t2);
        const size_t additional = bumpedMhc.size() - freePositionsList.size(), handled = layerBases.back();
        {
            size_t i = 0;
            for (; i < additional - keysInEndBucket; i++) freePositionsList.push_back(handled + i / k);
            for (; i < additional; i++) freePositionsList.push_back(nbuckets + i);
        }
        freePositions.build(freePositionsList);
        std::vector<size_t>().swap(freePositionsList);
        // the bumped keys' buckets (evaluateKPerfect's fallback branch, ShockHash2Flat.h:293-301)
        std::vector<uint64_t> fbBucket(bumpedMhc.size());
        parallelFor(bumpedMhc.size(), [&](size_t j) {
            const size_t bucket = freePositions.at(fallbackQuery(bumpedMhc[j]));
            fbBucket[j] = bucket >= nbuckets ? bucket - nbuckets + k * nbuckets : bucket;
        });
        std::vector<uint32_t> roffsets;
        std::vector<uint8_t> rseeds;
        std::vector<uint64_t> rwords;
        auto t3 = clk::now();
        phaseMs[2] = ms(t2, t3);
        eng.flat_leaves_retrieval(N, (uint32_t)k, THRESHOLD_MAPPING.data(), (uint32_t)THRESHOLD_RANGE, layerBases, thresholds,
                                  bumpedIndex, fbBucket, bucketCodes, gpu::keep_retrieval_input, retrievalKeys, retrievalBits,
                                  roffsets, rseeds, rwords, &phaseMs[4]);
        ribbon = RibbonRetrieval(std::move(roffsets), std::move(rseeds), std::move(rwords));
        phaseMs[3] = ms(t3, clk::now());
        for (size_t b = 0; b < nbuckets; b++) {
            size_t seed = bucketCodes[b];
            if (seed >= maxSeed || seed == SEED_FALLBACK_INDICATOR) {
                seedsFallback.insert({b, seed});
                seed = SEED_FALLBACK_INDICATOR;
            }
            setSeed(b, seed);
        }
    }

  public:
    ShockHash2Flat() = default;
    ShockHash2Flat(std::span<const std::string> keys, size_t ignore, size_t ignore2) : ShockHash2Flat(keys) {
        (void)ignore;
        (void)ignore2;
    }
    explicit ShockHash2Flat(std::span<const std::string> keys) {
        std::vector<uint64_t> mhcs(keys.size());
        parallelFor(keys.size(), [&](size_t i) { mhcs[i] = flat_detail::murmurHash64A(keys[i].data(), keys[i].size()); });
        construct(mhcs);
    }
    // Keys given by their 64-bit hashes (not in the reference interface, which hashes string keys itself).
    explicit ShockHash2Flat(std::span<const uint64_t> mhcs) { construct(mhcs); }

    size_t operator()(const std::string &key) const { return operator()(flat_detail::murmurHash64A(key.data(), key.size())); }
    size_t operator()(const hash128_t &hash) const { return operator()(hash.second); }
    // ShockHash2Flat.h:258-278
    size_t operator()(uint64_t hash) const {
        auto [bucket, seed] = evaluateKPerfect(hash);
        if (bucket >= nbuckets) return bucket;  // N that are not multiples of k
        if (seed == SEED_FALLBACK_INDICATOR) seed = seedsFallback.at(bucket);
        const uint64_t retrieved = ribbon.retrieve(hash);
        return bucket * k + sh2Cell((uint32_t)k, seed, hash, retrieved);
    }

    // (bucket, stored seed) of a key (ShockHash2Flat.h:280-302).
    std::pair<size_t, size_t> evaluateKPerfect(uint64_t mhc) const {
        for (size_t layer = 0; layer < layers; layer++) {
            if (layer != 0) mhc = remix(mhc);
            const size_t base = layerBases[layer], layerSize = layerBases[layer + 1] - base;
            const uint32_t bucket = fastrange32((uint32_t)mhc, (uint32_t)layerSize);
            const uint32_t threshold = (uint32_t)(mhc >> 32);
            auto [storedThreshold, storedSeed] = getThresholdAndSeed(base + bucket);
            if (threshold <= THRESHOLD_MAPPING[storedThreshold]) return {base + bucket, storedSeed};
        }
        if (!haveFallback) throw std::out_of_range("ShockHash2Flat: key not in the set");
        const size_t phf = fallbackQuery(mhc);
        const size_t bucket = freePositions.at(phf);
        if (bucket >= nbuckets) return {bucket - nbuckets + k * nbuckets, 1};  // last, partial bucket
        return {bucket, getThresholdAndSeed(bucket).second};
    }

    size_t size() const { return N; }
    [[nodiscard]] size_t getBits() const {
        return 8 * sizeof(*this) + (haveFallback ? fallbackPhf.getBits() : 0) + freePositions.sizeBits() +
               8 * freePositions.selectBytes() + 8 * ribbon.sizeBytes() + 8 * thresholdsAndSeeds.dataSizeBytes() +
               64 * seedsFallback.size();
    }
    void printBits() const {
        std::cout << "Thresholds: " << 1.0f * THRESHOLD_BITS / k << std::endl;
        std::cout << "Fallback PHF keys: " << (freePositions.sizeBits() > 0 ? freePositions.sizeBits() - N / k : 0) << std::endl;
        std::cout << "PHF: " << (haveFallback ? 1.0f * fallbackPhf.getBits() / N : 0.0f) << std::endl;
        std::cout << "Base case seeds: " << 1.0f * seedBits() / k << std::endl;
        std::cout << "Base case seeds overflow: " << 1.0f * seedsFallback.size() * 64 / N << std::endl;
        std::cout << "Ribbon: " << 8.0f * ribbon.sizeBytes() / N << std::endl;
    }

    // Verification accessors (not in the reference interface).
    const std::vector<uint64_t> &baseCaseCodes() const { return bucketCodes; }
    // construction phases in ms: threshold layers, fallback function, placement, device searches and retrieval, and
    // the batched base-case search alone
    const double *phaseTimings() const { return phaseMs; }
    const std::vector<uint64_t> &ribbonKeys() const { return retrievalKeys; }
    const std::vector<uint8_t> &ribbonBits() const { return retrievalBits; }
    size_t bucketCount() const { return nbuckets; }
    size_t layerCount() const { return layers; }
    size_t fallbackSeedCount() const { return seedsFallback.size(); }
};

}  // namespace shockhash
