// Host half of a GPU-built RecSplit/ShockHash forest: Golomb-Rice encoding of the seed codes in the
// reference's order and replay of the leaf cuckoo insertions that fix the retrieval bits.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "RibbonRetrieval.h"
#include "EliasFano.h"
#include "GolombMemo.h"
#include "GpuEngine.h"
#include "Hashing.h"
#include "RiceBitVector.h"

namespace shockhash {

// Leaf orientation of a ShockHash1 leaf (TinyBinaryCuckooHashTable::insert,
// include/shockhash/TinyBinaryCuckooHashTable.h:169-198), cells c1 < m/2 <= c2.
struct CandidateCells {
    uint32_t cell1, cell2;
};
inline CandidateCells sh1CandidateCells(uint64_t key, uint64_t seed, uint32_t m) {
    const uint64_t r = remix(key + seed);
    return {fastrange32((uint32_t)r, m / 2), fastrange32((uint32_t)(r >> 32), (m + 1) / 2) + m / 2};
}

// ShockHash2 leaf cell of a key for one candidate seed (ShockHash2-internal.h:100-102,473-477).
inline uint64_t sh2HalfHash(uint32_t m, uint64_t key, uint64_t seed) {
    if (m < 10) return fastrange64(remix(key + seed), (m + 1) / 2);
    uint64_t a, b;
    unpairElegant(seed, a, b);
    return fastrange64(remix(key + (((key & 1) == 0) ? a : b)), (m + 1) / 2);
}
// shockhash2query (src/ShockHash2-precompiled.cpp:6-22 -> ShockHash2-internal.h:574-584).
inline uint64_t sh2Cell(uint32_t m, uint64_t code, uint64_t key, uint64_t retrieved) {
    if (m <= 1) return 0;
    uint64_t s1, s2;
    unpairTriangular(code, s1, s2);
    return retrieved == 0 ? sh2HalfHash(m, key, s1) + m / 2 : sh2HalfHash(m, key, s2);
}

// TinyBinaryCuckooHashTable::insert (include/shockhash/TinyBinaryCuckooHashTable.h:169-198) over a table of
// key indices (cells[c] = key in cell c, -1 = empty; xr[k] = c1 ^ c2 of key k): places key k in c1 or c2,
// evicting along the alternative cells; false when the walk returns to its start with the same entry.
inline bool cuckooInsert(std::vector<int32_t> &cells, std::vector<uint32_t> &xr, int32_t k, uint32_t c1, uint32_t c2) {
    int32_t entry = k;
    xr[k] = c1 ^ c2;
    if (cells[c1] < 0) { cells[c1] = entry; return true; }
    if (cells[c2] < 0) { cells[c2] = entry; return true; }
    uint32_t cur = c1;
    do {
        const uint32_t alt = xr[entry] ^ cur;
        std::swap(entry, cells[alt]);
        if (entry < 0) return true;
        cur = alt;
    } while (cur != c1 || entry != k);
    return false;
}

// Retrieval input of one ShockHash2 leaf of m keys with pair code `code` (constructRetrieval,
// include/shockhash/ShockHash2-internal.h:548-572): the keys are inserted in leaf order with cells
// hash(key, seed1) + m/2 and hash(key, seed2); appends (key, 1 iff the key sits in its second cell) in key order.
// False when the keys cannot be placed (the code does not belong to these keys).
inline bool sh2LeafRetrieval(const uint64_t *keys, uint32_t m, uint64_t code,
                             std::vector<std::pair<uint64_t, uint8_t>> &out) {
    uint64_t s1, s2;
    unpairTriangular(code, s1, s2);
    std::vector<int32_t> cells(m, -1);
    std::vector<uint32_t> xr(m, 0), second(m, 0);
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t c1 = (uint32_t)sh2HalfHash(m, keys[k], s1) + m / 2;
        second[k] = (uint32_t)sh2HalfHash(m, keys[k], s2);
        if (!cuckooInsert(cells, xr, (int32_t)k, c1, second[k])) return false;
    }
    for (uint32_t k = 0; k < m; k++) out.emplace_back(keys[k], cells[second[k]] == (int32_t)k ? 1 : 0);
    return true;
}

// Encodes a GPU forest: descriptors (fixed parts in DFS order then unary parts per bucket, then the
// sentinel appendFixed(1, 1)) and bucket bit positions. Mirrors compute_thread + hash_gen
// (include/ShockHash.h:557-658, include/ShockHash2.h:396-497); the retrieval input (leaf orientations) is
// computed on the device (gpu::Engine::orientation_retrieval).
template <class MEMO>
void encodeForest(const gpu::Forest &f, const MEMO &memo, std::vector<uint64_t> &descriptor_words,
                  std::vector<uint64_t> &pos_acc) {
    auto golomb = [&](size_t m) { return (int)(memo[m] >> 27); };
    RiceBitVector::Builder builder;
    pos_acc.assign(f.nbuckets + 1, 0);
    std::vector<uint32_t> unary;
    size_t i = 0;
    for (uint64_t b = 0; b < f.nbuckets; b++) {
        const uint64_t m = f.size_acc[b + 1] - f.size_acc[b];
        if (m > 1) {
            if (i >= f.nodes.size() || f.nodes[i].offset != f.size_acc[b] || f.nodes[i].m != m)
                throw std::logic_error("GPU forest does not match the bucket sizes");
            const size_t end = i + f.nodes[i].subtree;
            unary.clear();
            for (; i < end; i++) {
                const shgpu_node &nd = f.nodes[i];
                const int k = golomb(nd.m);
                builder.appendFixed(f.codes[i], k);
                unary.push_back((uint32_t)(f.codes[i] >> k));
            }
            builder.appendUnaryAll(unary);
        }
        pos_acc[b + 1] = builder.getBits();
    }
    builder.appendFixed(1, 1);  // sentinel, include/ShockHash.h:652
    descriptor_words = builder.build();
}


}  // namespace shockhash
