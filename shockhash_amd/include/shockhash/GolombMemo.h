// Compile-time Golomb-Rice parameter tables of the RecSplit encoding. They fix the bit layout of the
// descriptor stream, so their values must equal the reference's: the leaf tables are copied values
// (bij_memo, include/ShockHash.h:53-61; bij_memo2, include/ShockHash2.h:33-47) and the split entries
// follow the same closed form (_fill_golomb_rice, include/ShockHash.h:79-127; _fill_golomb_rice2,
// include/ShockHash2.h:63-115), evaluated by g++ at compile time. Entry m packs
//   bits 27..31: log2 Golomb modulus of a node of m keys,
//   bits 16..26: nodes in the subtree of m keys,
//   bits  0..15: fixed-part bits of the subtree.
#pragma once
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace shockhash {

static const int MAX_BUCKET_SIZE = 3000;
static const int MAX_FANOUT = 32;
static const int MAX_LEAF_SIZE = 60;
static const int MAX_LEAF_SIZE2 = 128;

static constexpr uint8_t bij_memo[MAX_LEAF_SIZE + 1] = {
    0,  0,  0,  0,  0,  0,  1,  1,  2,  2,  2,  3,  3,  4,  4,  4,  5,  5,  6,  6,  7,
    7,  8,  8,  9,  9,  9,  10, 10, 10, 11, 11, 12, 12, 13, 13, 13, 14, 15, 15, 15,
    16, 16, 16, 17, 17, 18, 18, 19, 19, 19, 20, 20, 20, 21, 21, 22, 22, 23, 23, 23};

static constexpr uint8_t bij_memo2[MAX_LEAF_SIZE2 + 1] = {
    0,  0,  0,  0,  0,  0,  1,  3,  2,  4,  4,  6,  5,  7,  6,  8,  7,  8,  7,  9,  8,  10, 9,  11, 10, 11,
    11, 12, 11, 13, 12, 14, 13, 14, 14, 15, 14, 16, 15, 17, 16, 18, 17, 18, 17, 19, 18, 20, 19, 21, 20, 22,
    21, 22, 22, 23, 22, 24, 23, 25, 24, 26, 25, 27, 26, 27, 27, 28, 27, 29, 28, 30, 29, 31, 30, 32, 31, 33,
    32, 33, 33, 34, 33, 35, 34, 36, 35, 37, 36, 38, 37, 39, 38, 40, 38, 40, 40, 41, 40, 42, 41, 43, 42, 44,
    43, 45, 44, 45, 45, 47, 46, 47, 46, 48, 47, 49, 48, 50, 49, 51, 50, 52, 51, 53, 52, 54, 53, 54, 54};

// Aggregation thresholds (SplittingStrategy, include/ShockHash.h:63-72; SplittingStrategy2,
// include/ShockHash2.h:49-56).
template <size_t LEAF_SIZE, bool BIPARTITE>
struct Splitting {
    static constexpr size_t leaf = LEAF_SIZE;
    static constexpr size_t lower_aggr = BIPARTITE ? LEAF_SIZE * 4 : LEAF_SIZE * (LEAF_SIZE > 24 ? 4 : 2);
    static constexpr size_t upper_aggr = BIPARTITE ? lower_aggr * 3 : lower_aggr * (LEAF_SIZE > 24 ? 3 : 2);
};

template <size_t LEAF_SIZE, bool BIPARTITE>
constexpr std::array<uint64_t, MAX_BUCKET_SIZE> golomb_memo() {
    using S = Splitting<LEAF_SIZE, BIPARTITE>;
    std::array<uint64_t, MAX_BUCKET_SIZE> memo{};
    size_t s = 0;
    for (; s <= LEAF_SIZE; ++s) {
        const uint64_t b = BIPARTITE ? bij_memo2[s] : bij_memo[s];
        memo[s] = b << 27 | uint64_t(s > 1) << 16 | b;
    }
    for (; s < MAX_BUCKET_SIZE; ++s) {
        const size_t m = s;
        size_t fanout = 0, unit = 0;
        if (m > S::upper_aggr) {
            unit = S::upper_aggr * (uint16_t(m / 2 + S::upper_aggr - 1) / S::upper_aggr);
            fanout = 2;
        } else if (m > S::lower_aggr) {
            unit = S::lower_aggr;
            fanout = uint16_t(m + S::lower_aggr - 1) / S::lower_aggr;
        } else {
            unit = LEAF_SIZE;
            fanout = uint16_t(m + LEAF_SIZE - 1) / LEAF_SIZE;
        }
        long part[MAX_FANOUT] = {};
        part[fanout - 1] = (long)m;
        for (size_t i = 0; i + 1 < fanout; ++i) {
            part[i] = (long)unit;
            part[fanout - 1] -= part[i];
        }
        double sqrt_prod = 1;
        for (size_t i = 0; i < fanout; ++i) sqrt_prod *= std::sqrt((double)part[i]);
        const double p = std::sqrt((double)m) / (std::pow(2 * M_PI, (fanout - 1.) / 2) * sqrt_prod);
        uint64_t len = (uint64_t)std::ceil(std::log2(-std::log((std::sqrt(5) + 1) / 2) / std::log1p(-p)));
        memo[m] = len << 27;
        for (size_t i = 0; i < fanout; ++i) len += memo[part[i]] & 0xFFFF;
        memo[m] |= len;
        uint64_t nodes = 1;
        for (size_t i = 0; i < fanout; ++i) nodes += (memo[part[i]] >> 16) & 0x7FF;
        memo[m] |= nodes << 16;
    }
    return memo;
}

}  // namespace shockhash
