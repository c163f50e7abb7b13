// Host half of a GPU-built RecSplit/ShockHash forest: Golomb-Rice encoding of the seed codes in the
// reference's order and replay of the leaf cuckoo insertions that fix the retrieval bits.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "BitRetrieval.h"
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

// Cuckoo insertion with entries compared by identity; returns false when a key cannot be placed.
struct TinyCuckoo {
    int32_t cell[128];
    uint32_t xr[128];
    uint32_t m;
    explicit TinyCuckoo(uint32_t m_) : m(m_) { std::fill(cell, cell + m, -1); }
    bool insert(int32_t key, uint32_t c1, uint32_t c2) {
        int32_t entry = key;
        xr[key] = c1 ^ c2;
        if (cell[c1] < 0) { cell[c1] = entry; return true; }
        if (cell[c2] < 0) { cell[c2] = entry; return true; }
        uint32_t cur = c1;
        do {
            const uint32_t alt = xr[entry] ^ cur;
            std::swap(entry, cell[alt]);
            if (entry < 0) return true;
            cur = alt;
        } while (cur != c1 || entry != key);
        return false;
    }
};

enum class LeafKind { ShockHash1, ShockHash2 };

// Encodes a GPU forest: descriptors (fixed parts in DFS order then unary parts per bucket, then the
// sentinel appendFixed(1, 1)), bucket bit positions and the retrieval input in the reference's order
// (leaf by leaf in DFS order; ShockHash1 leaves cell by cell, ShockHash2 leaves key by key).
// Mirrors compute_thread + hash_gen (include/ShockHash.h:557-658, include/ShockHash2.h:396-497).
template <class MEMO>
void encodeForest(const gpu::Forest &f, const MEMO &memo, LeafKind leaf_kind, size_t num_threads,
                  std::vector<uint64_t> &descriptor_words, std::vector<uint64_t> &pos_acc,
                  std::vector<std::pair<uint64_t, uint8_t>> &ribbon_input) {
    auto golomb = [&](size_t m) { return (int)(memo[m] >> 27); };
    RiceBitVector::Builder builder;
    pos_acc.assign(f.nbuckets + 1, 0);
    std::vector<uint32_t> unary;
    std::vector<size_t> leaves;        // node index of every leaf, DFS order
    std::vector<uint64_t> leaf_out;    // ribbon offset of every leaf
    uint64_t ribbon_n = 0;
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
                if (nd.kind == SHGPU_NODE_LEAF) {
                    leaves.push_back(i);
                    leaf_out.push_back(ribbon_n);
                    ribbon_n += nd.m;
                }
            }
            builder.appendUnaryAll(unary);
        }
        pos_acc[b + 1] = builder.getBits();
    }
    builder.appendFixed(1, 1);  // sentinel, include/ShockHash.h:652
    descriptor_words = builder.build();

    ribbon_input.assign(ribbon_n, {0, 0});
    const size_t T = std::max<size_t>(1, num_threads);
    std::vector<std::thread> th;
    std::vector<int> failed(T, 0);
    auto run = [&](size_t t) {
        for (size_t l = t; l < leaves.size(); l += T) {
            const shgpu_node &nd = f.nodes[leaves[l]];
            const uint64_t *keys = f.keys.data() + nd.offset;
            const uint32_t m = nd.m;
            auto *out = ribbon_input.data() + leaf_out[l];
            TinyCuckoo tc(m);
            if (leaf_kind == LeafKind::ShockHash1) {
                const uint64_t x = f.codes[leaves[l]] + nd.start_seed;
                for (uint32_t k = 0; k < m; k++) {
                    const CandidateCells c = sh1CandidateCells(keys[k], x, m);
                    if (!tc.insert((int32_t)k, c.cell1, c.cell2)) { failed[t] = 1; break; }
                }
                if (failed[t]) return;
                for (uint32_t c = 0; c < m; c++) out[c] = {keys[tc.cell[c]], (uint8_t)(c < m / 2 ? 0 : 1)};
            } else {
                uint64_t s1, s2;
                unpairTriangular(f.codes[leaves[l]], s1, s2);
                uint32_t c2[128];
                for (uint32_t k = 0; k < m; k++) {
                    const uint32_t c1 = (uint32_t)sh2HalfHash(m, keys[k], s1) + m / 2;
                    c2[k] = (uint32_t)sh2HalfHash(m, keys[k], s2);
                    if (!tc.insert((int32_t)k, c1, c2[k])) { failed[t] = 1; break; }
                }
                if (failed[t]) return;
                for (uint32_t k = 0; k < m; k++) out[k] = {keys[k], (uint8_t)(tc.cell[c2[k]] == (int32_t)k ? 1 : 0)};
            }
        }
    };
    if (T == 1) run(0);
    else {
        for (size_t t = 0; t < T; t++) th.emplace_back(run, t);
        for (auto &x : th) x.join();
    }
    for (int fl : failed)
        if (fl) throw std::logic_error("GPU leaf seed does not admit a cuckoo placement");
}


}  // namespace shockhash
