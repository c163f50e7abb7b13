// ShockHash seed-search ORACLE — test infrastructure only.
//
// A plain CPU restatement of the reference's construction path (ByteHamster/ShockHash),
// written to check the GPU engine. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker / the reported CPU baseline.
// The product (libshockhash_amd.so + the host class templates) never links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates. Pins (checked in
// tests/test_oracle_*.py):
//   remix            SplitMix64 published output vectors (Vigna's splitmix64 / Rosetta Code)
//   remix32          MurmurHash3_x86_32 published test vectors (fmix32 is its finaliser)
//   fastrange*/remap closed forms floor(x*n/2^k) checked with exact integer arithmetic
//   pairing          the value tables drawn in PairingFunction.h's comments + round trips
//   cuckoo predicates three independent predicates agree with exhaustive matching search
//   leaf searches    brute force over seeds with an exhaustive orientation check (tiny m)
//   rotation fitting brute force over (x, r) with perfect-matching search (m up to 12, and 33)
//   split searches   brute force over seeds with plain per-part counts
//   Rice encoding    decode round trip + hand-built streams
// Parity unpinned (no vector source in the reference or the image): sux::remap16's 48-bit
// mask width, SpookyHash V2 (string keys only), sux DoubleEF / SimpleRibbon layouts (not
// restated: queries use an exact bit map instead of a ribbon).
#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

typedef unsigned __int128 u128;

namespace orc {

// ---------------------------------------------------------------- primitives

// Stafford Mix13: in-tree text SimdUtils.h:42-46 (remixV), asserted equal to the scalar
// bytehamster::util::remix at TinyBinaryCuckooHashTable.h:162-163.
static inline uint64_t remix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// fastrange32(x, n): SimdUtils.h:48-51 (remap32V), call sites TinyBinaryCuckooHashTable.h:140-141.
static inline uint32_t fastrange32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }
// fastrange64(x, n) = floor(x*n / 2^64): ShockHash2-internal.h:84,101,423,442,476.
static inline uint64_t fastrange64(uint64_t x, uint64_t n) { return (uint64_t)(((u128)x * n) >> 64); }
// sux::remap16 (48-bit mask; call sites ShockHash.h:212,458; ShockHash2.h:297).
static inline uint64_t remap16(uint64_t x, uint64_t n) { return ((x & ((1ull << 48) - 1)) * n) >> 48; }
// sux::remap128 (bucket mapping, Sorter.hpp:29, ShockHash.h:297).
static inline uint64_t remap128(uint64_t x, uint64_t n) { return (uint64_t)(((u128)x * n) >> 64); }
// MurmurHash3 fmix32: SIMDShockHash.hpp:76-83.
static inline uint32_t remix32(uint32_t z) {
    z ^= z >> 16; z *= 0x85ebca6bu; z ^= z >> 13; z *= 0xc2b2ae35u; z ^= z >> 16; return z;
}
// SIMDShockHash.hpp:762-766 (scalar) == remap() :58-71 (vector: lo32 ^ hi32 of key+seed).
static inline uint32_t remix_and_remap(uint64_t x, uint32_t n) {
    return ((remix32((uint32_t)(x >> 32) ^ (uint32_t)x) & 0xFFFFu) * n) >> 16;
}

// ShockHash.h:47-49
static const uint64_t START_SEED[20] = {
    0x106393c187cae21a, 0x6453cec3f7376937, 0x643e521ddbd2be98, 0x3740c6412f6572cb, 0x717d47562f1ce470,
    0x4cd6eb4c63befb7c, 0x9bfd8c5e18c8da73, 0x082f20e10092a9a3, 0x2ada2ce68d21defc, 0xe33cb4f3e7c6466b,
    0x3980be458c509c59, 0xc466fd9584828e8c, 0x45f0aabe1a61ede6, 0xf6e7b8b33ad9b98d, 0x4ef95e25f4b4983d,
    0x81175195173b92d3, 0x4e50927d8dd15978, 0x1ea2099d1fafae7f, 0x425c8a06fbaaa815, 0xcd4216006c74052a};

// ---------------------------------------------------------------- pairing (PairingFunction.h)

static inline uint64_t pair_triangular(uint64_t x, uint64_t y) { return (x - 1) * x / 2 + y; }  // :47-50
static inline void unpair_triangular(uint64_t z, uint64_t &x, uint64_t &y) {                  // :52-56
    x = (uint64_t)std::floor(0.5 + std::sqrt(0.25 + 2.0 * (double)z));
    y = z - ((x - 1) * x) / 2;
}
static inline bool float_accurate_triangular(uint64_t x, uint64_t y) {                          // :58-61
    uint64_t a, b; unpair_triangular(pair_triangular(x, y), a, b); return a == x && b == y;
}
static inline uint64_t pair_elegant(uint64_t x, uint64_t y) { return x < y ? y * y + x : x * x + x + y; }  // :75-81
static inline void unpair_elegant(uint64_t z, uint64_t &x, uint64_t &y) {                       // :83-91
    uint64_t f = (uint64_t)std::floor(std::sqrt((double)z));
    uint64_t r = f * f;
    if (z - r < f) { x = z - r; y = f; } else { x = f; y = z - r - f; }
}
static inline bool float_accurate_elegant(uint64_t x, uint64_t y) {                              // :93-96
    uint64_t a, b; unpair_elegant(pair_elegant(x, y), a, b); return a == x && b == y;
}

// ---------------------------------------------------------------- cuckoo predicates

// CuckooUnionFind::unionIsStillPseudoforest, CuckooUnionFind.h:19-41 (entries are xor values).
struct XorCuckoo {
    static constexpr uint32_t EMPTY = ~0u;
    std::vector<uint32_t> cells;
    explicit XorCuckoo(size_t n) : cells(n, EMPTY) {}
    void clear() { std::fill(cells.begin(), cells.end(), EMPTY); }
    bool add(uint32_t c1, uint32_t c2) {
        uint32_t entry = c1 ^ c2, orig_entry = entry;
        if (cells[c1] == EMPTY) { cells[c1] = entry; return true; }
        if (cells[c2] == EMPTY) { cells[c2] = entry; return true; }
        uint32_t cur = c1, orig_cell = cur;
        do {
            uint32_t alt = entry ^ cur;
            std::swap(entry, cells[alt]);
            if (entry == EMPTY) return true;
            cur = alt;
        } while (cur != orig_cell || entry != orig_entry);
        return false;
    }
};

// UnionFind::unionIsStillPseudoforest, UnionFind.h:58-67 (no path compression, tree flags).
struct TreeUF {
    std::vector<size_t> parent; std::vector<uint8_t> tree;
    explicit TreeUF(size_t n) : parent(n), tree(n, 1) { for (size_t i = 0; i < n; i++) parent[i] = i; }
    size_t find(size_t x) { while (parent[x] != x) x = parent[x]; return x; }
    bool add(size_t x, size_t y) {
        size_t r1 = find(x), r2 = find(y);
        parent[r1] = r2;
        bool ok = tree[r1] || tree[r2];
        tree[r2] = tree[r1] && tree[r2] && (r1 != r2);
        return ok;
    }
};

// TinyBinaryCuckooHashTable::insert, TinyBinaryCuckooHashTable.h:169-198. Entries are compared by
// identity (key index); cells[] holds key indices, -1 = empty.
struct IndexCuckoo {
    std::vector<int32_t> cells; std::vector<uint32_t> xr;
    IndexCuckoo(size_t n_cells, size_t n_keys) : cells(n_cells, -1), xr(n_keys, 0) {}
    bool insert(int32_t key, uint32_t c1, uint32_t c2) {
        int32_t entry = key, orig_entry = key;
        xr[key] = c1 ^ c2;
        if (cells[c1] < 0) { cells[c1] = entry; return true; }
        if (cells[c2] < 0) { cells[c2] = entry; return true; }
        uint32_t cur = c1, orig_cell = cur;
        do {
            uint32_t alt = xr[entry] ^ cur;
            std::swap(entry, cells[alt]);
            if (entry < 0) return true;
            cur = alt;
        } while (cur != orig_cell || entry != orig_entry);
        return false;
    }
};

// ---------------------------------------------------------------- ShockHash1 leaf

struct Cells { uint32_t c1, c2; };
// TinyBinaryCuckooHashTable::getCandidateCells (runtime range), TinyBinaryCuckooHashTable.h:138-149.
static inline Cells sh1_cells(uint64_t key, uint64_t seed, uint32_t m) {
    uint64_t r = remix(key + seed);
    return {fastrange32((uint32_t)r, m / 2), fastrange32((uint32_t)(r >> 32), (m + 1) / 2) + m / 2};
}

// TinyBinaryCuckooHashTable::construct, :80-106: insert every key in leaf order.
static bool sh1_construct(const uint64_t *keys, uint32_t m, uint64_t seed, IndexCuckoo *out) {
    IndexCuckoo t(m, m);
    for (uint32_t i = 0; i < m; i++) {
        Cells c = sh1_cells(keys[i], seed, m);
        if (!t.insert((int32_t)i, c.c1, c.c2)) return false;
    }
    if (out) *out = t;
    return true;
}

// Leaf search of ShockHash::recSplit (ShockHash.h:404-434), SIMDShockHash::leafLevel
// non-rotation (SIMDShockHash.hpp:507-542, same predicate, lanes scanned in seed order) and
// BijectionsShockHash1::calculateBijection (benchmark/bijections/ShockHash1.h:13-35, x from 0).
// Returns the absolute seed. `bits_by_cell`/`key_by_cell` receive the orientation in cell order
// (ShockHash.h:425-427): cell i holds key key_by_cell[i], bit = i < m/2 ? 0 : 1.
static uint64_t sh1_leaf(const uint64_t *keys, uint32_t m, uint64_t x, int32_t *key_by_cell) {
    const uint64_t all = (m == 64) ? ~0ull : ((1ull << m) - 1);
    for (;;) {
        for (;;) {
            uint64_t mask = 0;
            for (uint32_t i = 0; i < m; i++) {
                Cells c = sh1_cells(keys[i], x, m);
                mask |= (1ull << c.c1) | (1ull << c.c2);
            }
            if (mask == all) break;
            x++;
        }
        IndexCuckoo t(m, m);
        if (sh1_construct(keys, m, x, &t)) {
            if (key_by_cell) for (uint32_t i = 0; i < m; i++) key_by_cell[i] = t.cells[i];
            return x;
        }
        x++;
    }
}

// Rotation fitting of a full leaf (m == LEAF_SIZE, ROTATION_FITTING): ShockHash.h:316-403 and
// SIMDShockHash.hpp:419-506 (same search, vectorised over r). Keys with even mhc (group A) are hashed with seed
// x & ~mask (mask 7 for leaves above 32 keys, GROUP_A_HF_MASK at :141), odd ones (group B) with seed x; for each
// x in order and each rotation r in 0..m-1 in order, a candidate needs A's cells OR B's cells rotated by r to
// cover all m cells, then the cuckoo insertions of every key in heap order (A in leaf order, then B in reverse
// leaf order) with B's cells rotated by r must succeed. Returns x (absolute) and r; key_by_cell receives the
// leaf-index of the key in each cell, and the caller's bit for cell i is (i' < m/2) with i' = i for A and
// (i - r + m) % m for B (:380-388).
static inline uint64_t rot_mask(uint32_t m) { return m > 32 ? 7ull : 0ull; }
static inline uint64_t rotl_bits(uint64_t v, uint32_t r, uint32_t m) {  // ShockHash.h:297-299
    const uint64_t all = m == 64 ? ~0ull : ((1ull << m) - 1);
    return r == 0 ? v & all : ((v << r) | (v >> (m - r))) & all;
}
static void sh1_leaf_rot(const uint64_t *keys, uint32_t m, uint64_t x, uint64_t *x_out, uint32_t *r_out, int32_t *key_by_cell) {
    const uint64_t all = m == 64 ? ~0ull : ((1ull << m) - 1), mask = rot_mask(m);
    std::vector<int32_t> heap(m);  // heap position -> leaf index
    uint32_t na = 0, ib = m - 1;
    for (uint32_t i = 0; i < m; i++) {
        if ((keys[i] & 1) == 0) heap[na++] = (int32_t)i;
        else heap[ib--] = (int32_t)i;
    }
    std::vector<Cells> cache(m);
    uint64_t a = 0, b = 0;
    for (;; x++) {
        if (a == 0 || (x & mask) == 0) {
            a = 0;
            for (uint32_t i = 0; i < na; i++) {
                cache[i] = sh1_cells(keys[heap[i]], x & ~mask, m);
                a |= (1ull << cache[i].c1) | (1ull << cache[i].c2);
            }
        }
        b = 0;
        for (uint32_t i = na; i < m; i++) {
            cache[i] = sh1_cells(keys[heap[i]], x, m);
            b |= (1ull << cache[i].c1) | (1ull << cache[i].c2);
        }
        for (uint32_t r = 0; r < m; r++) {
            if ((a | rotl_bits(b, r, m)) != all) continue;
            IndexCuckoo t(m, m);
            uint32_t i = 0;
            for (; i < m; i++) {
                Cells c = cache[i];
                if (keys[heap[i]] & 1) { c.c1 = (c.c1 + r) % m; c.c2 = (c.c2 + r) % m; }
                if (!t.insert((int32_t)i, c.c1, c.c2)) break;
            }
            if (i == m) {
                *x_out = x;
                *r_out = r;
                if (key_by_cell) for (uint32_t q = 0; q < m; q++) key_by_cell[q] = heap[t.cells[q]];
                return;
            }
        }
    }
}
// Cell of a key in a rotation-fitted leaf (ShockHash.h:251-265, SIMDShockHash.hpp:826-840).
static uint64_t sh1_rot_query_cell(uint32_t m, uint64_t code, uint64_t start, uint64_t key, uint64_t bit) {
    const uint32_t r = (uint32_t)(code % m);
    const uint64_t x = code / m + start;
    if ((key & 1) == 0) {
        Cells c = sh1_cells(key, x & ~rot_mask(m), m);
        return bit == 0 ? c.c1 : c.c2;
    }
    Cells c = sh1_cells(key, x, m);
    return ((bit == 0 ? c.c1 : c.c2) + r) % m;
}

// ---------------------------------------------------------------- ShockHash2 leaf
// BijectionsShockHash2<m, true, QuadSplitCandidateFinderList> for m >= 10 and
// <m, true, BasicSeedCandidateFinder> for m < 10 (src/ShockHash2-precompiled.cpp:24-47).

struct Cand {            // SeedCache<leafSize, true>, ShockHash2-internal.h:26-32
    u128 isolated;
    uint64_t seed;
    uint8_t hashes[128];
};

// calculateIsolatedVertices, ShockHash2-internal.h:34-45
static void isolated_vertices(Cand &c, uint32_t m) {
    uint32_t hit[128] = {0};
    for (uint32_t i = 0; i < m; i++) hit[c.hashes[i]]++;
    c.isolated = 0;
    for (uint32_t i = 0; i < m; i++) if (hit[c.hashes[i]] == 1) c.isolated |= (u128)1 << i;
}

static inline uint64_t mask_half(uint32_t m) {  // MASK_HALF, ShockHash2-internal.h:23-24
    uint32_t h = (m + 1) / 2; return h >= 64 ? ~0ull : ((1ull << h) - 1);
}

// BasicSeedCandidateFinder::next, ShockHash2-internal.h:78-98
struct BasicFinder {
    const uint64_t *keys; uint32_t m; uint64_t cur = 0;
    Cand next() {
        Cand c;
        const uint32_t H = (m + 1) / 2;
        for (;;) {
            uint64_t taken = 0;
            for (uint32_t i = 0; i < m; i++) {
                c.hashes[i] = (uint8_t)fastrange64(remix(keys[i] + cur), H);
                taken |= 1ull << c.hashes[i];
            }
            if (taken == mask_half(m)) { c.seed = cur; isolated_vertices(c, m); cur++; return c; }
            cur++;
        }
    }
    static uint64_t hash(uint32_t m, uint64_t key, uint64_t seed) {  // :100-102
        return fastrange64(remix(key + seed), (m + 1) / 2);
    }
};

// QuadSplitCandidateFinder with CandidateList, ShockHash2-internal.h:386-483 (list :188-278).
struct QuadFinder {
    uint32_t m; uint32_t sizeA = 0; uint64_t keys[128]; uint64_t cur = ~0ull;
    std::vector<uint64_t> listA, listB;   // mask per half-seed, index = half-seed
    std::vector<Cand> extracted;          // smallest code at the back
    QuadFinder(const uint64_t *in, uint32_t m_) : m(m_) {
        for (uint32_t i = 0; i < m; i++) {  // :399-406: even keys forward, odd keys backward
            if ((in[i] & 1) == 0) keys[sizeA++] = in[i];
            else keys[m - i + sizeA - 1] = in[i];
        }
    }
    Cand make(uint64_t a, uint64_t b) {   // makeCache :418-433
        Cand c; c.seed = pair_elegant(a, b);
        const uint32_t H = (m + 1) / 2;
        for (uint32_t i = 0; i < sizeA; i++) c.hashes[i] = (uint8_t)fastrange64(remix(keys[i] + a), H);
        for (uint32_t i = sizeA; i < m; i++) c.hashes[i] = (uint8_t)fastrange64(remix(keys[i] + b), H);
        isolated_vertices(c, m);
        return c;
    }
    void prepare_next() {                  // prepareNextSeed :435-471
        cur++;
        const uint32_t H = (m + 1) / 2;
        const uint64_t full = mask_half(m);
        uint64_t ta = 0, tb = 0;
        for (uint32_t i = 0; i < sizeA; i++) ta |= 1ull << fastrange64(remix(keys[i] + cur), H);
        for (uint32_t i = sizeA; i < m; i++) tb |= 1ull << fastrange64(remix(keys[i] + cur), H);
        for (uint64_t a = 0; a < listA.size(); a++)
            if ((listA[a] | tb) == full && float_accurate_elegant(a, cur)) extracted.push_back(make(a, cur));
        listA.push_back(ta);
        listB.push_back(tb);
        for (uint64_t b = 0; b < listB.size(); b++)
            if ((listB[b] | ta) == full && float_accurate_elegant(cur, b)) extracted.push_back(make(cur, b));
        std::sort(extracted.begin(), extracted.end(), [](const Cand &x, const Cand &y) { return x.seed > y.seed; });
    }
    Cand next() {
        while (extracted.empty()) prepare_next();
        Cand c = extracted.back(); extracted.pop_back(); return c;
    }
    static uint64_t hash(uint32_t m, uint64_t key, uint64_t seed) {  // :473-477
        uint64_t a, b; unpair_elegant(seed, a, b);
        return fastrange64(remix(key + (((key & 1) == 0) ? a : b)), (m + 1) / 2);
    }
};

static inline uint64_t sh2_hash(uint32_t m, uint64_t key, uint64_t seed) {
    return m >= 10 ? QuadFinder::hash(m, key, seed) : BasicFinder::hash(m, key, seed);
}

// BijectionsShockHash2::findSeed, ShockHash2-internal.h:501-546. Also reports the number of
// half-seeds / seeds the finder hashed (for evaluation counts).
template <class Finder>
static uint64_t sh2_find_seed_t(Finder &f, uint32_t m) {
    std::vector<Cand> seen;
    seen.push_back(f.next());
    XorCuckoo uf(m);
    for (;;) {
        Cand nc = f.next();
        Cand sh = nc;
        for (uint32_t i = 0; i < m; i++) sh.hashes[i] = (uint8_t)(sh.hashes[i] + m / 2);
        for (const Cand &o : seen) {
            if ((o.isolated & nc.isolated) != 0) continue;
            uf.clear();
            uint32_t i = 0;
            for (; i < m; i++) if (!uf.add(sh.hashes[i], o.hashes[i])) break;
            if (i != m) continue;
            if (!float_accurate_triangular(nc.seed, o.seed)) continue;
            return pair_triangular(nc.seed, o.seed);
        }
        seen.push_back(nc);
    }
}

static uint64_t sh2_find_seed(const uint64_t *keys, uint32_t m) {
    if (m < 2) return 0;  // dispatchLeafSize<I<=1> returns 0 (src/ShockHash2-precompiled.cpp:27-28)
    if (m >= 10) { QuadFinder f(keys, m); return sh2_find_seed_t(f, m); }
    BasicFinder f{keys, m}; return sh2_find_seed_t(f, m);
}

// BijectionsShockHash2::constructRetrieval, ShockHash2-internal.h:548-572: bits in leaf key order.
static void sh2_retrieval(const uint64_t *keys, uint32_t m, uint64_t code, uint8_t *bits) {
    uint64_t s1, s2; unpair_triangular(code, s1, s2);
    IndexCuckoo t(m, m);
    std::vector<uint32_t> c2(m);
    for (uint32_t k = 0; k < m; k++) {
        uint32_t a = (uint32_t)sh2_hash(m, keys[k], s1) + m / 2;
        c2[k] = (uint32_t)sh2_hash(m, keys[k], s2);
        if (!t.insert((int32_t)k, a, c2[k])) throw std::logic_error("Should be possible to construct");
    }
    for (uint32_t k = 0; k < m; k++) bits[k] = (t.cells[c2[k]] == (int32_t)k) ? 1 : 0;
}

// BijectionsShockHash2::hash, ShockHash2-internal.h:574-584 (via shockhash2query).
static uint64_t sh2_query_cell(uint32_t m, uint64_t code, uint64_t key, uint64_t retrieved) {
    if (m <= 1) return 0;
    uint64_t s1, s2; unpair_triangular(code, s1, s2);
    return retrieved == 0 ? sh2_hash(m, key, s1) + m / 2 : sh2_hash(m, key, s2);
}

// ---------------------------------------------------------------- splitting strategies & memo

enum Variant { SH1 = 0, SIMD = 1, SH2 = 2 };

// ShockHash.h:53-61
static const uint8_t BIJ_MEMO[61] = {0, 0, 0, 0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7,
                                     7, 8, 8, 9, 9, 9, 10, 10, 10, 11, 11, 12, 12, 13, 13, 13, 14, 15, 15, 15,
                                     16, 16, 16, 17, 17, 18, 18, 19, 19, 19, 20, 20, 20, 21, 21, 22, 22, 23, 23, 23};
// ShockHash2.h:33-47
static const uint8_t BIJ_MEMO2[129] = {
    0, 0, 0, 0, 0, 0, 1, 3, 2, 4, 4, 6, 5, 7, 6, 8, 7, 8, 7, 9, 8, 10, 9, 11, 10, 11, 11, 12, 11, 13,
    12, 14, 13, 14, 14, 15, 14, 16, 15, 17, 16, 18, 17, 18, 17, 19, 18, 20, 19, 21, 20, 22, 21, 22, 22, 23,
    22, 24, 23, 25, 24, 26, 25, 27, 26, 27, 27, 28, 27, 29, 28, 30, 29, 31, 30, 32, 31, 33, 32, 33, 33, 34,
    33, 35, 34, 36, 35, 37, 36, 38, 37, 39, 38, 40, 38, 40, 40, 41, 40, 42, 41, 43, 42, 44, 43, 45, 44, 45,
    45, 47, 46, 47, 46, 48, 47, 49, 48, 50, 49, 51, 50, 52, 51, 53, 52, 54, 53, 54, 54};

static const size_t MAX_BUCKET = 3000;  // ShockHash.h:33

struct Strategy {
    Variant v; uint32_t leaf, lower, upper;
    Strategy(Variant v_, uint32_t L) : v(v_), leaf(L) {
        if (v == SH2) { lower = 4 * L; upper = 3 * lower; }           // ShockHash2.h:49-56
        else { lower = L * (L > 24 ? 4 : 2); upper = lower * (L > 24 ? 3 : 2); }  // ShockHash.h:63-72
    }
};

// _fill_golomb_rice / _fill_golomb_rice2 (ShockHash.h:79-127, ShockHash2.h:63-115), evaluated at
// run time with libm (the product evaluates the same formula at compile time with g++).
static std::vector<uint64_t> golomb_memo(const Strategy &st) {
    std::vector<uint64_t> memo(MAX_BUCKET, 0);
    const uint8_t *bij = st.v == SH2 ? BIJ_MEMO2 : BIJ_MEMO;
    size_t s = 0;
    for (; s <= st.leaf; ++s) memo[s] = (uint64_t)bij[s] << 27 | (uint64_t)(s > 1) << 16 | bij[s];
    for (; s < MAX_BUCKET; ++s) {
        const size_t m = s;
        size_t fanout, unit;
        if (m > st.upper) { unit = st.upper * (uint16_t(m / 2 + st.upper - 1) / st.upper); fanout = 2; }
        else if (m > st.lower) { unit = st.lower; fanout = uint16_t(m + st.lower - 1) / st.lower; }
        else { unit = st.leaf; fanout = uint16_t(m + st.leaf - 1) / st.leaf; }
        long k[32] = {0};
        k[fanout - 1] = (long)m;
        for (size_t i = 0; i < fanout - 1; ++i) { k[i] = (long)unit; k[fanout - 1] -= k[i]; }
        double sqrt_prod = 1;
        for (size_t i = 0; i < fanout; ++i) sqrt_prod *= std::sqrt((double)k[i]);
        const double p = std::sqrt((double)m) / (std::pow(2 * M_PI, (fanout - 1.) / 2) * sqrt_prod);
        uint64_t len = (uint64_t)std::ceil(std::log2(-std::log((std::sqrt(5) + 1) / 2) / std::log1p(-p)));
        memo[m] = len << 27;
        for (size_t i = 0; i < fanout; ++i) len += memo[k[i]] & 0xFFFF;
        memo[m] |= len;
        uint64_t nodes = 1;
        for (size_t i = 0; i < fanout; ++i) nodes += (memo[k[i]] >> 16) & 0x7FF;
        memo[m] |= nodes << 16;
    }
    return memo;
}

// ---------------------------------------------------------------- split searches

// Split-node geometry, ShockHash.h:451-548 / SIMDShockHash.hpp:563-617 / ShockHash2.h:290-387.
struct SplitShape { int kind; uint32_t fanout, unit; uint64_t start; };  // kind 0: fanout-2 (unit=split)
static SplitShape split_shape(const Strategy &st, uint32_t m, int level) {
    SplitShape s;
    if (m > st.upper) {
        s.kind = 0; s.fanout = 2; s.unit = (uint16_t(m / 2 + st.upper - 1) / st.upper) * st.upper;
        s.start = START_SEED[level];
    } else if (m > st.lower) {
        s.kind = 1; s.unit = st.lower; s.fanout = uint16_t(m + st.lower - 1) / st.lower;
        s.start = st.v == SIMD ? START_SEED[17] : START_SEED[level];
    } else {
        s.kind = 2; s.unit = st.leaf; s.fanout = uint16_t(m + st.leaf - 1) / st.leaf;
        s.start = st.v == SIMD ? START_SEED[18] : START_SEED[level];
    }
    return s;
}

static inline uint32_t split_hash(Variant v, uint64_t key, uint64_t x, uint32_t m) {
    return v == SIMD ? remix_and_remap(key + x, m) : (uint32_t)remap16(remix(key + x), m);
}
static inline uint32_t split_part(const SplitShape &s, uint32_t h) {
    return s.kind == 0 ? (h >= s.unit ? 1u : 0u) : (uint32_t)(uint16_t(h) / s.unit);
}

// Predicate of one seed. sux flavour: per-part counts (ShockHash.h:455-462,484-493,518-527).
// SIMD flavour: higherLevel's count of h < split (SIMDShockHash.hpp:214-225) and aggrLevel's
// packed 8-bit counters, evaluated mod 2^32 exactly as the vector code adds them (:284-330).
static bool split_ok(Variant v, const SplitShape &s, const uint64_t *keys, uint32_t m, uint64_t x) {
    if (v != SIMD || s.kind == 0) {
        uint32_t cnt[32] = {0};
        for (uint32_t i = 0; i < m; i++) cnt[split_part(s, split_hash(v, keys[i], x, m))]++;
        if (s.kind == 0) return cnt[0] == s.unit;
        for (uint32_t p = 0; p + 1 < s.fanout; p++) if (cnt[p] != s.unit) return false;
        return true;
    }
    const uint32_t SPLIT = s.unit, fanout = s.fanout;
    uint32_t i;
    if (fanout <= 5) {
        uint32_t found = SPLIT;
        for (i = 1; i < fanout - 1; ++i) found |= SPLIT << (8 * i);
        if (fanout <= 4) found |= (m - (fanout - 1) * SPLIT) << (8 * i);
        uint32_t count = 0;
        for (uint32_t k = 0; k < m; k++) {
            uint32_t p = split_hash(v, keys[k], x, m) / SPLIT;
            count += (p < 4) ? (1u << (8 * p)) : 0u;   // aggr_level_count_lookup[4 + p]
        }
        return count == found;
    }
    const uint32_t found_low = (SPLIT << 24) | (SPLIT << 16) | (SPLIT << 8) | SPLIT;
    uint32_t found_high = SPLIT;
    for (i = 1; i < fanout - 5; ++i) found_high |= SPLIT << (8 * i);
    if (fanout <= 8) found_high |= (m - (fanout - 1) * SPLIT) << (8 * i);
    uint32_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < m; k++) {
        uint32_t p = split_hash(v, keys[k], x, m) / SPLIT;
        lo += (p < 4) ? (1u << (8 * p)) : 0u;
        hi += (p >= 4 && p < 8) ? (1u << (8 * (p - 4))) : 0u;
    }
    return lo == found_low && hi == found_high;
}

// Smallest x >= s.start with split_ok, then the stable counting-sort partition of the keys
// (ShockHash.h:464-469,495-499,528-532; SIMDShockHash.hpp:234-252,337-354). Returns x - start.
static uint64_t split_node(Variant v, const SplitShape &s, uint64_t *keys, uint32_t m, uint64_t *temp) {
    uint64_t x = s.start;
    while (!split_ok(v, s, keys, m, x)) x++;
    uint32_t pos[32];
    for (uint32_t p = 0, c = 0; p < s.fanout; p++, c += s.unit) pos[p] = c;
    for (uint32_t i = 0; i < m; i++) temp[pos[split_part(s, split_hash(v, keys[i], x, m))]++] = keys[i];
    std::memcpy(keys, temp, m * sizeof(uint64_t));
    return x - s.start;
}

// ---------------------------------------------------------------- Rice bit vector (RiceBitVector.h)

struct RiceBuilder {
    std::vector<uint64_t> w; uint64_t bits = 0;
    static size_t words_for(uint64_t b) { return (((b + 7) / 8) + 7 + 7) / 8; }  // resize rule :120,141,152
    void grow(uint64_t b) { if (w.size() < words_for(b)) w.resize(words_for(b), 0); }
    void append_fixed(uint64_t v, int k) {                                  // appendFixed :116-133
        grow(bits + k);
        for (int j = 0; j < k; j++, bits++) if ((v >> j) & 1) w[bits / 64] |= 1ull << (bits % 64);
    }
    void append_unary_all(const std::vector<uint32_t> &u) {                 // appendUnaryAll :135-149
        uint64_t inc = 0; for (uint32_t x : u) inc += x + 1;
        grow(bits + inc);
        for (uint32_t x : u) { bits += x; w[bits / 64] |= 1ull << (bits % 64); bits++; }
    }
    void append(const RiceBuilder &o) {                                     // appendRiceBitVector :151-169
        grow(bits + o.bits);
        for (uint64_t j = 0; j < o.bits; j++, bits++)
            if ((o.w[j / 64] >> (j % 64)) & 1) w[bits / 64] |= 1ull << (bits % 64);
    }
};

struct RiceReader {                                                          // Reader :198-263
    const std::vector<uint64_t> &w; uint64_t fixed = 0, unary = 0;
    explicit RiceReader(const std::vector<uint64_t> &w_) : w(w_) {}
    bool bit(uint64_t p) const { return p / 64 < w.size() && ((w[p / 64] >> (p % 64)) & 1); }
    void reset(uint64_t bit_pos, uint64_t unary_offset) { fixed = bit_pos; unary = bit_pos + unary_offset; }
    uint64_t next(int k) {
        uint64_t q = 0;
        while (!bit(unary)) { q++; unary++; }
        unary++;
        uint64_t r = 0;
        for (int j = 0; j < k; j++) if (bit(fixed + j)) r |= 1ull << j;
        fixed += k;
        return (q << k) | r;
    }
    void skip(uint64_t nodes, uint64_t fixed_len) {
        while (nodes) { if (bit(unary)) nodes--; unary++; }
        fixed += fixed_len;
    }
};

// ---------------------------------------------------------------- RecSplit DFS

struct NodeRec { uint32_t bucket, m; int32_t level; uint8_t kind; uint64_t code; };  // kind 3 = leaf

struct Build {
    Variant v; Strategy st; size_t n = 0, nbuckets = 0;
    bool rot = false;  // ROTATION_FITTING (ShockHash / SIMDShockHash): full leaves use sh1_leaf_rot
    std::vector<uint64_t> memo, size_acc, pos_acc, words;
    uint64_t bits = 0;
    std::vector<std::pair<uint64_t, uint8_t>> ribbon;
    std::unordered_map<uint64_t, uint8_t> retr;
    std::vector<NodeRec> nodes;
    Build(Variant v_, uint32_t L) : v(v_), st(v_, L) { memo = golomb_memo(st); }
    int golomb(size_t m) const { return (int)(memo[m] >> 27); }
    uint64_t skip_bits(size_t m) const { return memo[m] & 0xFFFF; }
    uint64_t skip_nodes(size_t m) const { return (memo[m] >> 16) & 0x7FF; }
};

struct Ctx {
    const Build &b; RiceBuilder rb; std::vector<uint32_t> unary;
    std::vector<std::pair<uint64_t, uint8_t>> ribbon; std::vector<NodeRec> nodes;
    std::vector<uint64_t> temp; uint32_t bucket = 0;
    explicit Ctx(const Build &b_) : b(b_), temp(MAX_BUCKET + 64) {}
    void emit(uint64_t code, size_t m) {
        int k = b.golomb(m);
        rb.append_fixed(code, k);
        unary.push_back((uint32_t)(code >> k));
    }
};

// Leaves: ShockHash.h:404-439, SIMDShockHash.hpp:507-542 (x from start_seed[19]), ShockHash2.h:272-278.
static void leaf(Ctx &c, uint64_t *keys, uint32_t m, int level) {
    uint64_t code;
    if (c.b.v == SH2) {
        code = sh2_find_seed(keys, m);
        uint8_t bits[128];
        sh2_retrieval(keys, m, code, bits);
        for (uint32_t k = 0; k < m; k++) c.ribbon.emplace_back(keys[k], bits[k]);
    } else if (c.b.rot && m == c.b.st.leaf) {
        const uint64_t start = c.b.v == SIMD ? START_SEED[19] : START_SEED[level];
        int32_t by_cell[64];
        uint64_t x = 0;
        uint32_t r = 0;
        sh1_leaf_rot(keys, m, start, &x, &r, by_cell);
        for (uint32_t i = 0; i < m; i++) {
            const uint64_t k = keys[by_cell[i]];
            const uint32_t cell = (k & 1) ? (i - r + m) % m : i;
            c.ribbon.emplace_back(k, cell < m / 2 ? 0 : 1);
        }
        code = (x - start) * m + r;  // ShockHash.h:401-402
    } else {
        uint64_t start = c.b.v == SIMD ? START_SEED[19] : START_SEED[level];
        int32_t by_cell[64];
        uint64_t x = sh1_leaf(keys, m, start, by_cell);
        for (uint32_t i = 0; i < m; i++) c.ribbon.emplace_back(keys[by_cell[i]], i < m / 2 ? 0 : 1);
        code = x - start;
    }
    c.emit(code, m);
    c.nodes.push_back({c.bucket, m, level, 3, code});
}

// recSplit: ShockHash.h:303-555, SIMDShockHash.hpp:563-617, ShockHash2.h:259-394.
static void rec_split(Ctx &c, uint64_t *keys, uint32_t m, int level) {
    const Strategy &st = c.b.st;
    if (m <= st.leaf) { leaf(c, keys, m, level); return; }
    SplitShape s = split_shape(st, m, level);
    uint64_t code = split_node(c.b.v, s, keys, m, c.temp.data());
    c.emit(code, m);
    c.nodes.push_back({c.bucket, m, level, (uint8_t)s.kind, code});
    if (s.kind == 0) {
        rec_split(c, keys, s.unit, level + 1);
        if (m - s.unit > 1) rec_split(c, keys + s.unit, m - s.unit, level + 1);
        return;
    }
    uint32_t i;
    for (i = 0; i < m - s.unit; i += s.unit) rec_split(c, keys + i, s.unit, level + 1);
    if (m - i > 1) rec_split(c, keys + i, m - i, level + 1);
}

// parallelPartition (Sorter.hpp:15-36): sort by .first, bucket b = remap128(first, nbuckets), the
// .second values in that order into `sorted`, bucket prefix sums into b->size_acc.
static void partition(Build *b, const uint64_t *h, size_t n, size_t bucket_size, std::vector<uint64_t> &sorted) {
    b->n = n;
    b->nbuckets = std::max<size_t>(1, (n + bucket_size - 1) / bucket_size);
    std::vector<std::pair<uint64_t, uint64_t>> kv(n);
    for (size_t i = 0; i < n; i++) kv[i] = {h[2 * i], h[2 * i + 1]};
    std::sort(kv.begin(), kv.end());
    sorted.assign(n, 0);
    b->size_acc.assign(b->nbuckets + 1, 0);
    b->pos_acc.assign(b->nbuckets + 1, 0);
    size_t i = 0;
    for (size_t bk = 0; bk < b->nbuckets; bk++) {
        b->size_acc[bk] = i;
        while (i < n && remap128(kv[i].first, b->nbuckets) == bk) { sorted[i] = kv[i].second; i++; }
    }
    b->size_acc[b->nbuckets] = n;
}

// compute_thread's per-bucket work (ShockHash.h:566-579) over a chosen set of buckets only, for a partition
// too large to build whole on the CPU: the nodes of those buckets' trees (codes, sizes, levels), in the order
// of `sel`. Descriptors, positions and retrieval input are not produced.
static Build *build_buckets(Variant v, uint32_t L, const uint64_t *h, size_t n, size_t bucket_size, const uint32_t *sel,
                            size_t nsel, int threads) {
    Build *b = new Build(v, L);
    std::vector<uint64_t> sorted;
    partition(b, h, n, bucket_size, sorted);
    if (threads < 1) threads = 1;
    std::vector<std::vector<NodeRec>> per(nsel);
    std::vector<std::thread> th;
    std::mutex mu;
    size_t next = 0;
    for (int t = 0; t < threads; t++) {
        th.emplace_back([&]() {
            Ctx c(*b);
            for (;;) {
                size_t k;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    if (next >= nsel) return;
                    k = next++;
                }
                const uint32_t bk = sel[k];
                if (bk >= b->nbuckets) continue;
                const size_t s = b->size_acc[bk + 1] - b->size_acc[bk];
                c.nodes.clear();
                if (s > 1) {
                    c.bucket = bk;
                    rec_split(c, &sorted[b->size_acc[bk]], (uint32_t)s, 0);
                }
                per[k] = c.nodes;
            }
        });
    }
    for (auto &x : th) x.join();
    for (auto &p : per) b->nodes.insert(b->nodes.end(), p.begin(), p.end());
    return b;
}

// hash_gen + compute_thread (ShockHash.h:557-688; ShockHash2.h:396-498; SIMDShockHash.hpp:619-760):
// partition by .first (Sorter.hpp:15-36), contiguous bucket ranges per thread, tid-ordered merge.
static Build *build(Variant v, uint32_t L, const uint64_t *h, size_t n, size_t bucket_size, int threads, bool rot = false) {
    Build *b = new Build(v, L);
    b->rot = rot && v != SH2;
    std::vector<uint64_t> sorted;
    partition(b, h, n, bucket_size, sorted);
    if (threads < 1) threads = 1;
    std::vector<Ctx *> ctx(threads);
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        ctx[t] = new Ctx(*b);
        auto work = [&, t]() {
            size_t beg = (size_t)t * b->nbuckets / threads;
            size_t end = (t == threads - 1) ? b->nbuckets : std::min(b->nbuckets, (size_t)(t + 1) * b->nbuckets / threads);
            Ctx &c = *ctx[t];
            for (size_t bk = beg; bk < end; bk++) {
                size_t s = b->size_acc[bk + 1] - b->size_acc[bk];
                if (s > 1) {
                    c.bucket = (uint32_t)bk;
                    rec_split(c, &sorted[b->size_acc[bk]], (uint32_t)s, 0);
                    c.rb.append_unary_all(c.unary);
                    c.unary.clear();
                }
                b->pos_acc[bk + 1] = c.rb.bits;
            }
        };
        if (threads == 1) work(); else th.emplace_back(work);
    }
    for (auto &x : th) x.join();
    RiceBuilder all;
    for (int t = 0; t < threads; t++) {
        size_t beg = (size_t)t * b->nbuckets / threads;
        size_t end = (t == threads - 1) ? b->nbuckets : std::min(b->nbuckets, (size_t)(t + 1) * b->nbuckets / threads);
        uint64_t off = all.bits;
        if (t > 0) for (size_t bk = beg + 1; bk < end + 1; bk++) b->pos_acc[bk] += off;
        all.append(ctx[t]->rb);
        b->ribbon.insert(b->ribbon.end(), ctx[t]->ribbon.begin(), ctx[t]->ribbon.end());
        b->nodes.insert(b->nodes.end(), ctx[t]->nodes.begin(), ctx[t]->nodes.end());
        delete ctx[t];
    }
    all.append_fixed(1, 1);  // sentinel, ShockHash.h:652
    b->words = all.w;
    b->bits = all.bits;
    for (auto &p : b->ribbon) b->retr[p.first] = p.second;
    return b;
}

// operator()(hash128_t): ShockHash.h:199-270, SIMDShockHash.hpp:776-846, ShockHash2.h:178-230.
static uint64_t query(const Build &b, uint64_t first, uint64_t second) {
    const Strategy &st = b.st;
    const size_t bucket = remap128(first, b.nbuckets);
    uint64_t cum = b.size_acc[bucket];
    size_t m = b.size_acc[bucket + 1] - cum;
    RiceReader r(b.words);
    r.reset(b.pos_acc[bucket], b.skip_bits(m));
    int level = 0;
    auto h = [&](uint64_t d, uint64_t seed_add, size_t mm) -> uint64_t {
        return b.v == SIMD ? remix_and_remap(second + d + seed_add, (uint32_t)mm) : remap16(remix(second + d + seed_add), mm);
    };
    while (m > st.upper) {
        uint64_t d = r.next(b.golomb(m));
        uint64_t hm = h(d, START_SEED[level], m);
        uint32_t split = (uint16_t(m / 2 + st.upper - 1) / st.upper) * st.upper;
        if (hm < split) m = split;
        else { r.skip(b.skip_nodes(split), b.skip_bits(split)); m -= split; cum += split; }
        level++;
    }
    if (m > st.lower) {
        uint64_t d = r.next(b.golomb(m));
        uint64_t hm = h(d, b.v == SIMD ? START_SEED[17] : START_SEED[level], m);
        uint64_t part = uint16_t(hm) / st.lower;
        m = std::min<size_t>(st.lower, m - part * st.lower);
        cum += st.lower * part;
        if (part) r.skip(b.skip_nodes(st.lower) * part, b.skip_bits(st.lower) * part);
        if (b.v != SIMD) level++;
    }
    if (m > st.leaf) {
        uint64_t d = r.next(b.golomb(m));
        uint64_t hm = h(d, b.v == SIMD ? START_SEED[18] : START_SEED[level], m);
        uint64_t part = uint16_t(hm) / st.leaf;
        m = std::min<size_t>(st.leaf, m - part * st.leaf);
        cum += st.leaf * part;
        if (part) r.skip(part, b.skip_bits(st.leaf) * part);
        if (b.v != SIMD) level++;
    }
    uint64_t code = r.next(b.golomb(m));
    auto it = b.retr.find(second);
    uint64_t bit = it == b.retr.end() ? 0 : it->second;
    if (b.v == SH2) return cum + sh2_query_cell((uint32_t)m, code, second, bit);
    if (b.rot && m == st.leaf) return cum + sh1_rot_query_cell((uint32_t)m, code, b.v == SIMD ? START_SEED[19] : START_SEED[level], second, bit);
    uint64_t seed = code + (b.v == SIMD ? START_SEED[19] : START_SEED[level]);
    Cells cc = sh1_cells(second, seed, (uint32_t)m);
    return cum + (bit == 0 ? cc.c1 : cc.c2);
}

// ---------------------------------------------------------------- SpookyHash V2
// sux `spooky(k, len, seed)` = SpookyHash::Hash128 with both state words = seed, returned as
// {h1, h0} (ShockHash.h:129). Public-domain SpookyV2 (Bob Jenkins): the short path below 192 bytes
// (ShortMix / ShortEnd over 32-byte rounds) and the long path (Mix per 96-byte block, End over the
// zero-padded last block that ends in its length, three EndPartial rounds). The short path is pinned by
// SpookyV2's own self-test table (Hash32 of bytes (i + 128) & 255, lengths 0..63: tests/golden); the long path
// has no vector source here (parity unpinned: it agrees with the library's two restatements only).
static inline uint64_t rot64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline void short_mix(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h2 = rot64(h2, 50); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 52); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 30); h0 += h1; h2 ^= h0;  h1 = rot64(h1, 41); h1 += h2; h3 ^= h1;
    h2 = rot64(h2, 54); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 48); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 38); h0 += h1; h2 ^= h0;  h1 = rot64(h1, 37); h1 += h2; h3 ^= h1;
    h2 = rot64(h2, 62); h2 += h3; h0 ^= h2;  h3 = rot64(h3, 34); h3 += h0; h1 ^= h3;
    h0 = rot64(h0, 5);  h0 += h1; h2 ^= h0;  h1 = rot64(h1, 36); h1 += h2; h3 ^= h1;
}
static inline void short_end(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h3 ^= h2; h2 = rot64(h2, 15); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 52); h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 26); h1 += h0;  h2 ^= h1; h1 = rot64(h1, 51); h2 += h1;
    h3 ^= h2; h2 = rot64(h2, 28); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 9);  h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 47); h1 += h0;  h2 ^= h1; h1 = rot64(h1, 54); h2 += h1;
    h3 ^= h2; h2 = rot64(h2, 32); h3 += h2;  h0 ^= h3; h3 = rot64(h3, 25); h0 += h3;
    h1 ^= h0; h0 = rot64(h0, 63); h1 += h0;
}
static void spooky_short(const uint8_t *p, size_t len, uint64_t &o1, uint64_t &o2) {
    const uint64_t C = 0xdeadbeefdeadbeefull;
    auto ld64 = [](const uint8_t *q) { uint64_t v; std::memcpy(&v, q, 8); return v; };
    auto ld32 = [](const uint8_t *q) { uint32_t v; std::memcpy(&v, q, 4); return (uint64_t)v; };
    size_t rem = len % 32;
    uint64_t a = o1, b = o2, c = C, d = C;
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
        case 0: c += C; d += C;
    }
    short_end(a, b, c, d);
    o1 = a; o2 = b;
}
static void spooky_mix(const uint64_t *data, uint64_t *s) {
    static const int R[12] = {11, 32, 43, 31, 17, 28, 39, 57, 55, 54, 22, 46};
    for (int i = 0; i < 12; i++) {
        s[i] += data[i];
        s[(i + 2) % 12] ^= s[(i + 10) % 12];
        s[(i + 11) % 12] ^= s[i];
        s[i] = rot64(s[i], R[i]);
        s[(i + 11) % 12] += s[(i + 1) % 12];
    }
}
static void spooky_end_partial(uint64_t *h) {
    static const int R[12] = {44, 15, 34, 21, 38, 33, 10, 13, 38, 53, 42, 54};
    for (int i = 0; i < 12; i++) {
        h[(i + 11) % 12] += h[(i + 1) % 12];
        h[(i + 2) % 12] ^= h[(i + 11) % 12];
        h[(i + 1) % 12] = rot64(h[(i + 1) % 12], R[i]);
    }
}
static void spooky128(const uint8_t *p, size_t len, uint64_t &o1, uint64_t &o2) {
    if (len < 192) { spooky_short(p, len, o1, o2); return; }
    uint64_t h[12], blk[12];
    for (int i = 0; i < 12; i += 3) { h[i] = o1; h[i + 1] = o2; h[i + 2] = 0xdeadbeefdeadbeefull; }
    const size_t whole = len / 96;
    for (size_t k = 0; k < whole; k++) { std::memcpy(blk, p + 96 * k, 96); spooky_mix(blk, h); }
    const size_t rem = len - 96 * whole;
    uint8_t last[96] = {0};
    std::memcpy(last, p + 96 * whole, rem);
    last[95] = (uint8_t)rem;
    std::memcpy(blk, last, 96);
    for (int i = 0; i < 12; i++) h[i] += blk[i];
    spooky_end_partial(h); spooky_end_partial(h); spooky_end_partial(h);
    o1 = h[0]; o2 = h[1];
}

}  // namespace orc

// ================================================================ C ABI (ctypes)
using namespace orc;

// ---------------------------------------------------------------- ShockHash2Flat (include/ShockHash2Flat.h)
// k-perfect bucketing of the keys' 64-bit hashes (mhc) into N / k buckets of exactly k keys, ShockHash2's leaf
// search inside each bucket, and a ShockHash2<k> fallback over the keys the thresholds bump. Restated from
// ShockHash2Flat.h:37-302 for keys given by their mhc (the reference computes mhc = MurmurHash64(key) of string
// keys at :70-75; that hash is an un-vendored dependency and stays outside this restatement).
struct Flat {
    uint32_t k = 0, tbits = 0, trange = 0, seed_bits = 0;
    uint64_t max_seed = 0;
    size_t N = 0, nbuckets = 0, layers = 2;
    std::vector<uint32_t> mapping;               // THRESHOLD_MAPPING
    std::vector<uint32_t> thr;                   // compacted threshold per bucket
    std::vector<uint64_t> seed;                  // stored seed per bucket (0: in seed_fb)
    std::map<size_t, uint64_t> seed_fb;          // seedsFallback
    std::vector<size_t> layer_bases;             // layerBases
    Build *fallback = nullptr;                   // fallbackPhf: ShockHash2<k> over the decimal strings of mhc
    std::vector<size_t> free_pos;                // freePositions (select1(i + 1) - i of freePositionsBv)
    std::unordered_map<uint64_t, uint8_t> retr;  // the ribbon's answers for the keys
    size_t bumped = 0;                           // keys handed to the fallback
    ~Flat() { delete fallback; }
};

static uint32_t floor_log2(uint64_t x) { uint32_t r = 0; while (x >>= 1) r++; return r; }

// _fill_mapping<THRESHOLD_RANGE> (ShockHash2Flat.h:17-35).
static std::vector<uint32_t> flat_mapping(uint32_t range) {
    const uint64_t one_third = 0xFFFFFFFFull / 3;
    std::vector<uint32_t> a(range, 0);
    if (range == 1) return a;
    if (range == 2) { a[0] = 0; a[1] = 0xFFFFFFFFu; return a; }
    a[0] = 0;
    a[1] = (uint32_t)one_third;
    const uint64_t ir = range - 3;
    for (uint64_t i = 0; i < ir; i++) a[2 + i] = (uint32_t)(2 * one_third + one_third * i / ir);
    a[range - 1] = 0xFFFFFFFFu;
    return a;
}

// compact_threshold (:191-199).
static uint32_t flat_compact(const Flat &f, uint32_t t) {
    for (uint32_t i = 0; i < f.trange; i++)
        if (t <= f.mapping[i]) return i;
    return f.mapping.back();
}

struct FlatKey { uint64_t mhc; uint32_t bucket, threshold; };  // KeyInfo (ShockHash2FlatBase.h)

static inline uint64_t flat_fallback_query(const Flat &f, uint64_t mhc) {
    const std::string s = std::to_string(mhc);
    uint64_t h0 = 0, h1 = 0;
    spooky128((const uint8_t *)s.data(), s.size(), h0, h1);
    return query(*f.fallback, h1, h0);  // sux spooky() = {h1, h0} = {first, second}
}

// evaluateKPerfect (:270-299): (bucket, stored seed).
static std::pair<size_t, uint64_t> flat_evaluate(const Flat &f, uint64_t mhc) {
    for (size_t layer = 0; layer < f.layers; layer++) {
        if (layer != 0) mhc = remix(mhc);
        const size_t base = f.layer_bases[layer], size = f.layer_bases[layer + 1] - base;
        const uint32_t bucket = fastrange32((uint32_t)mhc, (uint32_t)size);
        const uint32_t t = (uint32_t)(mhc >> 32);
        if (t <= f.mapping[f.thr[base + bucket]]) return {base + bucket, f.seed[base + bucket]};
    }
    const size_t phf = flat_fallback_query(f, mhc);
    const size_t bucket = f.free_pos[phf];
    if (bucket >= f.nbuckets) return {bucket - f.nbuckets + (size_t)f.k * f.nbuckets, 1};
    return {bucket, f.seed[bucket]};
}

// Constructor (:64-172) from the keys' mhc values.
static Flat *flat_build(uint32_t k, const uint64_t *mhcs, size_t n, int threads) {
    Flat *f = new Flat();
    f->k = k;
    f->tbits = floor_log2(k) - 1;
    f->trange = 1u << f->tbits;
    f->seed_bits = (uint32_t)std::ceil(0.442 * k - 0.2);
    f->max_seed = 1ull << f->seed_bits;
    f->mapping = flat_mapping(f->trange);
    f->N = n;
    f->nbuckets = n / k;
    const size_t keys_in_end = n - f->nbuckets * k;
    size_t buckets_this_layer = (size_t)std::ceil(0.9 * f->nbuckets);
    std::vector<FlatKey> hashes(n);
    for (size_t i = 0; i < n; i++)
        hashes[i] = {mhcs[i], fastrange32((uint32_t)mhcs[i], (uint32_t)buckets_this_layer), (uint32_t)(mhcs[i] >> 32)};
    std::vector<FlatKey> all = hashes;
    f->layer_bases.push_back(0);
    f->thr.assign(f->nbuckets, 0);
    f->seed.assign(f->nbuckets, 0);
    if (buckets_this_layer == 0) f->layers = 0;
    for (size_t layer = 0; layer < f->layers; layer++) {
        const size_t layer_base = f->layer_bases.back();
        if (layer != 0) {
            buckets_this_layer = (size_t)(0.9 * (double)(hashes.size() / k));
            buckets_this_layer = std::min(buckets_this_layer, f->nbuckets - layer_base);
            if (buckets_this_layer == 0) { f->layers = 1; break; }
            for (auto &h : hashes) {  // rehash
                h.mhc = remix(h.mhc);
                h.bucket = fastrange32((uint32_t)h.mhc, (uint32_t)buckets_this_layer);
                h.threshold = (uint32_t)(h.mhc >> 32);
            }
        }
        f->layer_bases.push_back(layer_base + buckets_this_layer);
        // sort_keyInfo (src/Sorter.cpp:10-12): by bucket, then threshold
        std::stable_sort(hashes.begin(), hashes.end(), [](const FlatKey &a, const FlatKey &b) {
            return ((uint64_t)a.bucket << 32 | a.threshold) < ((uint64_t)b.bucket << 32 | b.threshold);
        });
        std::vector<FlatKey> bumped;
        // flushBucket (:201-228)
        auto flush = [&](size_t start, size_t end, size_t bidx) {
            const size_t size = end - start, b = layer_base + bidx;
            if (size <= k) {
                f->thr[b] = f->trange - 1;
                for (size_t q = size; q < k; q++) f->free_pos.push_back(b);
                return;
            }
            const uint32_t last = flat_compact(*f, hashes[start + k - 1].threshold);
            const uint32_t first_bumped = flat_compact(*f, hashes[start + k].threshold);
            if (first_bumped == last && last == 0) throw std::runtime_error("ShockHash2Flat: threshold below 0");
            const uint32_t t = first_bumped == last ? last - 1 : last;
            f->thr[b] = t;
            const uint32_t uncompressed = f->mapping[t];
            for (size_t l = 0; l < size; l++)
                if (hashes[start + l].threshold > uncompressed) {
                    bumped.push_back(hashes[start + l]);
                    if (l < k) f->free_pos.push_back(b);
                }
        };
        size_t bucket_start = 0, prev = 0;
        for (size_t i = 0; i < hashes.size(); i++) {
            while (hashes[i].bucket != prev) { flush(bucket_start, i, prev); prev++; bucket_start = i; }
        }
        while (prev < buckets_this_layer) { flush(bucket_start, hashes.size(), prev); prev++; bucket_start = hashes.size(); }
        hashes = bumped;
    }
    // fallback PHF over the bumped keys (:124-129)
    f->bumped = hashes.size();
    std::vector<uint64_t> fh(2 * hashes.size());
    for (size_t i = 0; i < hashes.size(); i++) {
        const std::string s = std::to_string(hashes[i].mhc);
        uint64_t h0 = 0, h1 = 0;
        spooky128((const uint8_t *)s.data(), s.size(), h0, h1);
        fh[2 * i] = h1;
        fh[2 * i + 1] = h0;
    }
    f->fallback = build(SH2, k, fh.data(), hashes.size(), 2000, threads);
    // free positions of the buckets no layer handled and of the last, partial bucket (:130-146)
    const size_t additional = hashes.size() - f->free_pos.size(), handled = f->layer_bases.back();
    {
        size_t i = 0;
        for (; i < additional - keys_in_end; i++) f->free_pos.push_back(handled + i / k);
        for (; i < additional; i++) f->free_pos.push_back(f->nbuckets + i);
    }
    // base case in every bucket (:148-169): keys in input order, ShockHash2 leaf search and its retrieval bits
    std::vector<std::vector<uint64_t>> contents(f->nbuckets);
    for (const FlatKey &key : all) {
        const size_t b = flat_evaluate(*f, key.mhc).first;
        if (b < f->nbuckets) contents[b].push_back(key.mhc);
    }
    for (auto &c : contents)
        if (c.size() != k) throw std::runtime_error("ShockHash2Flat: a bucket does not hold k keys");
    std::vector<uint64_t> seeds(f->nbuckets);
    std::vector<std::vector<uint8_t>> bits(f->nbuckets, std::vector<uint8_t>(k));
    std::vector<std::thread> th;
    const int nt = std::max(1, threads);
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (size_t b = t; b < f->nbuckets; b += nt) {
                seeds[b] = sh2_find_seed(contents[b].data(), k);
                sh2_retrieval(contents[b].data(), k, seeds[b], bits[b].data());
            }
        });
    for (auto &x : th) x.join();
    for (size_t b = 0; b < f->nbuckets; b++) {
        for (uint32_t q = 0; q < k; q++) f->retr[contents[b][q]] = bits[b][q];
        uint64_t sd = seeds[b];
        if (sd >= f->max_seed || sd == 0) { f->seed_fb[b] = sd; sd = 0; }
        f->seed[b] = sd;
    }
    return f;
}

// operator()(uint64_t) (:254-267).
static uint64_t flat_query(const Flat &f, uint64_t mhc) {
    auto [bucket, sd] = flat_evaluate(f, mhc);
    if (bucket >= f.nbuckets) return bucket;
    if (sd == 0) sd = f.seed_fb.at(bucket);
    auto it = f.retr.find(mhc);
    const uint64_t bit = it == f.retr.end() ? 0 : it->second;
    return bucket * f.k + sh2_query_cell(f.k, sd, mhc, bit);
}

extern "C" {

uint64_t orc_remix(uint64_t z) { return remix(z); }
uint32_t orc_fastrange32(uint32_t x, uint32_t n) { return fastrange32(x, n); }
uint64_t orc_fastrange64(uint64_t x, uint64_t n) { return fastrange64(x, n); }
uint64_t orc_remap16(uint64_t x, uint64_t n) { return remap16(x, n); }
uint64_t orc_remap128(uint64_t x, uint64_t n) { return remap128(x, n); }
uint32_t orc_remix32(uint32_t z) { return remix32(z); }
uint32_t orc_remix_and_remap(uint64_t x, uint32_t n) { return remix_and_remap(x, n); }
uint64_t orc_start_seed(int i) { return START_SEED[i]; }
uint64_t orc_pair_triangular(uint64_t x, uint64_t y) { return pair_triangular(x, y); }
void orc_unpair_triangular(uint64_t z, uint64_t *x, uint64_t *y) { unpair_triangular(z, *x, *y); }
uint64_t orc_pair_elegant(uint64_t x, uint64_t y) { return pair_elegant(x, y); }
void orc_unpair_elegant(uint64_t z, uint64_t *x, uint64_t *y) { unpair_elegant(z, *x, *y); }
int orc_float_accurate_triangular(uint64_t x, uint64_t y) { return float_accurate_triangular(x, y); }
int orc_float_accurate_elegant(uint64_t x, uint64_t y) { return float_accurate_elegant(x, y); }

void orc_spooky128(const void *data, size_t len, uint64_t seed, uint64_t *first, uint64_t *second) {
    uint64_t h0 = seed, h1 = seed;
    spooky128((const uint8_t *)data, len, h0, h1);
    *first = h1; *second = h0;   // sux spooky() returns {h1, h0}
}

// Fingerprints of many byte strings (string i = data[offs[i], offs[i+1])): out[2i] = first, out[2i+1] =
// second, as the string constructors compute them (ShockHash.h:165,175; spooky(k, len, 0)).
void orc_spooky128_batch(const uint8_t *data, const uint64_t *offs, size_t n, uint64_t seed, uint64_t *out, int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([=]() {
            for (size_t i = t; i < n; i += threads) {
                uint64_t h0 = seed, h1 = seed;
                spooky128(data + offs[i], offs[i + 1] - offs[i], h0, h1);
                out[2 * i] = h1;
                out[2 * i + 1] = h0;
            }
        });
    for (auto &x : th) x.join();
}

// Whether seeds[l] satisfies the ShockHash1 leaf predicate for leaf l (every cell covered and the cuckoo
// insertion in leaf order succeeds, ShockHash.h:411-433): ok[l] = 1 or 0.
void orc_sh1_valid_batch(const uint64_t *keys, size_t n_leaves, uint32_t m, const uint64_t *seeds, uint8_t *ok, int threads) {
    if (threads < 1) threads = 1;
    const uint64_t all = (m == 64) ? ~0ull : ((1ull << m) - 1);
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([=]() {
            for (size_t l = t; l < n_leaves; l += threads) {
                const uint64_t *k = keys + l * m;
                uint64_t mask = 0;
                for (uint32_t i = 0; i < m; i++) {
                    Cells c = sh1_cells(k[i], seeds[l], m);
                    mask |= (1ull << c.c1) | (1ull << c.c2);
                }
                ok[l] = mask == all && sh1_construct(k, m, seeds[l], nullptr);
            }
        });
    for (auto &x : th) x.join();
}

// Sequential edge insertion; returns the number of edges accepted before the first failure
// (== n on success). pred: 0 CuckooUnionFind, 1 UnionFind, 2 TinyBinaryCuckooHashTable::insert.
size_t orc_pseudoforest_prefix(int pred, const uint32_t *c1, const uint32_t *c2, size_t n, size_t n_cells) {
    XorCuckoo x(n_cells); TreeUF u(n_cells); IndexCuckoo t(n_cells, n);
    for (size_t i = 0; i < n; i++) {
        bool ok = pred == 0 ? x.add(c1[i], c2[i]) : pred == 1 ? u.add(c1[i], c2[i]) : t.insert((int32_t)i, c1[i], c2[i]);
        if (!ok) return i;
    }
    return n;
}

void orc_sh1_cells(uint64_t key, uint64_t seed, uint32_t m, uint32_t *c1, uint32_t *c2) {
    Cells c = sh1_cells(key, seed, m); *c1 = c.c1; *c2 = c.c2;
}

// One ShockHash1 leaf; returns the absolute seed, key_by_cell (m entries) optional.
uint64_t orc_sh1_leaf(const uint64_t *keys, uint32_t m, uint64_t start, int32_t *key_by_cell) {
    return sh1_leaf(keys, m, start, key_by_cell);
}

// Many equal-size isolated leaves (bijections/ShockHash1.h semantics when start = 0).
void orc_sh1_leaf_batch(const uint64_t *keys, size_t n_leaves, uint32_t m, uint64_t start, uint64_t *seeds, int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([=]() {
            for (size_t l = t; l < n_leaves; l += threads) seeds[l] = sh1_leaf(keys + l * m, m, start, nullptr);
        });
    for (auto &x : th) x.join();
}

uint64_t orc_sh2_find_seed(const uint64_t *keys, uint32_t m) { return sh2_find_seed(keys, m); }

void orc_sh2_leaf_batch(const uint64_t *keys, const uint32_t *offsets, const uint32_t *sizes, size_t n_leaves,
                        uint64_t *codes, int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([=]() {
            for (size_t l = t; l < n_leaves; l += threads) codes[l] = sh2_find_seed(keys + offsets[l], sizes[l]);
        });
    for (auto &x : th) x.join();
}

void orc_sh2_retrieval(const uint64_t *keys, uint32_t m, uint64_t code, uint8_t *bits) { sh2_retrieval(keys, m, code, bits); }
uint64_t orc_sh2_query_cell(uint32_t m, uint64_t code, uint64_t key, uint64_t retrieved) {
    return sh2_query_cell(m, code, key, retrieved);
}
uint64_t orc_sh2_hash(uint32_t m, uint64_t key, uint64_t seed) { return sh2_hash(m, key, seed); }

// Split node search + partition (keys permuted in place). variant 0 sux flavour, 1 SIMD flavour.
// Returns the code (x - start); shape fields returned for inspection. A node of at most `leaf` keys
// is a leaf, not a split: returns UINT64_MAX with kind = -1 instead of searching for an impossible split.
uint64_t orc_split(int variant, int leaf, int level, uint64_t *keys, uint32_t m,
                   int *kind, uint32_t *fanout, uint32_t *unit, uint64_t *start) {
    if (m <= (uint32_t)leaf) {
        if (kind) *kind = -1;
        return UINT64_MAX;
    }
    Strategy st((Variant)variant, (uint32_t)leaf);
    SplitShape s = split_shape(st, m, level);
    std::vector<uint64_t> temp(m + 64);
    uint64_t code = split_node((Variant)variant, s, keys, m, temp.data());
    if (kind) *kind = s.kind;
    if (fanout) *fanout = s.fanout;
    if (unit) *unit = s.unit;
    if (start) *start = s.start;
    return code;
}
int orc_split_ok(int variant, int leaf, int level, const uint64_t *keys, uint32_t m, uint64_t x) {
    Strategy st((Variant)variant, (uint32_t)leaf);
    return split_ok((Variant)variant, split_shape(st, m, level), keys, m, x);
}

void orc_golomb_memo(int variant, int leaf, uint64_t *out /* 3000 */) {
    Strategy st((Variant)variant, (uint32_t)leaf);
    std::vector<uint64_t> m = golomb_memo(st);
    std::memcpy(out, m.data(), m.size() * sizeof(uint64_t));
}

void *orc_build(int variant, int leaf, const uint64_t *hashes, size_t n, size_t bucket_size, int threads) {
    try { return build((Variant)variant, (uint32_t)leaf, hashes, n, bucket_size, threads); }
    catch (...) { return nullptr; }
}
// ROTATION_FITTING = true builds of ShockHash / SIMDShockHash (ShockHash.h:316-403, SIMDShockHash.hpp:419-506).
void *orc_build_rot(int variant, int leaf, const uint64_t *hashes, size_t n, size_t bucket_size, int threads) {
    try { return build((Variant)variant, (uint32_t)leaf, hashes, n, bucket_size, threads, true); }
    catch (...) { return nullptr; }
}
// One rotation-fitted leaf: absolute x, rotation r and the leaf index of the key in each cell.
void orc_sh1_leaf_rot(const uint64_t *keys, uint32_t m, uint64_t start, uint64_t *x, uint32_t *r, int32_t *key_by_cell) {
    sh1_leaf_rot(keys, m, start, x, r, key_by_cell);
}
uint64_t orc_sh1_rot_query_cell(uint32_t m, uint64_t code, uint64_t start, uint64_t key, uint64_t bit) {
    return sh1_rot_query_cell(m, code, start, key, bit);
}
// Nodes of the chosen buckets only (orc_build_nnodes / orc_build_nodes read them; nothing else is built).
void *orc_build_buckets(int variant, int leaf, const uint64_t *hashes, size_t n, size_t bucket_size, const uint32_t *sel,
                        size_t nsel, int threads) {
    return build_buckets((Variant)variant, (uint32_t)leaf, hashes, n, bucket_size, sel, nsel, threads);
}
void orc_build_free(void *b) { delete (Build *)b; }
size_t orc_build_nbuckets(void *b) { return ((Build *)b)->nbuckets; }
size_t orc_build_nwords(void *b) { return ((Build *)b)->words.size(); }
uint64_t orc_build_bits(void *b) { return ((Build *)b)->bits; }
void orc_build_words(void *b, uint64_t *out) { auto &w = ((Build *)b)->words; std::memcpy(out, w.data(), w.size() * 8); }
void orc_build_acc(void *b, uint64_t *size_acc, uint64_t *pos_acc) {
    Build *x = (Build *)b;
    std::memcpy(size_acc, x->size_acc.data(), x->size_acc.size() * 8);
    std::memcpy(pos_acc, x->pos_acc.data(), x->pos_acc.size() * 8);
}
void orc_build_ribbon(void *b, uint64_t *keys, uint8_t *bits) {
    Build *x = (Build *)b;
    for (size_t i = 0; i < x->ribbon.size(); i++) { keys[i] = x->ribbon[i].first; bits[i] = x->ribbon[i].second; }
}
size_t orc_build_nnodes(void *b) { return ((Build *)b)->nodes.size(); }
void orc_build_nodes(void *b, uint32_t *bucket, uint32_t *m, int32_t *level, uint8_t *kind, uint64_t *code) {
    Build *x = (Build *)b;
    for (size_t i = 0; i < x->nodes.size(); i++) {
        const NodeRec &r = x->nodes[i];
        bucket[i] = r.bucket; m[i] = r.m; level[i] = r.level; kind[i] = r.kind; code[i] = r.code;
    }
}
void orc_build_query(void *b, const uint64_t *hashes, size_t n, uint64_t *out) {
    Build *x = (Build *)b;
    for (size_t i = 0; i < n; i++) out[i] = query(*x, hashes[2 * i], hashes[2 * i + 1]);
}


// ShockHash2Flat<k> built from the keys' mhc values (ShockHash2Flat.h:64-172), queried by mhc (:254-267).
void *orc_flat_build(uint32_t k, const uint64_t *mhcs, size_t n, int threads) {
    try { return flat_build(k, mhcs, n, threads); }
    catch (...) { return nullptr; }
}
void orc_flat_free(void *f) { delete (Flat *)f; }
void orc_flat_query(void *f, const uint64_t *mhcs, size_t n, uint64_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = flat_query(*(Flat *)f, mhcs[i]);
}
// nbuckets, layers, layer bases (3 entries), keys handed to the fallback, seeds kept in the fallback map
void orc_flat_info(void *fp, uint64_t *out7) {
    const Flat &f = *(Flat *)fp;
    out7[0] = f.nbuckets; out7[1] = f.layers;
    for (int i = 0; i < 3; i++) out7[2 + i] = i < (int)f.layer_bases.size() ? f.layer_bases[i] : 0;
    out7[5] = f.bumped; out7[6] = f.seed_fb.size();
}
// (bucket, stored seed) of each key (evaluateKPerfect) and THRESHOLD_MAPPING of k
void orc_flat_evaluate(void *f, const uint64_t *mhcs, size_t n, uint64_t *bucket, uint64_t *seed) {
    for (size_t i = 0; i < n; i++) { auto r = flat_evaluate(*(Flat *)f, mhcs[i]); bucket[i] = r.first; seed[i] = r.second; }
}
// The compacted threshold kept by every bucket of the layers (setThreshold's values), layer_bases.back() entries.
void orc_flat_thresholds(void *fp, uint32_t *out) {
    const Flat &f = *(Flat *)fp;
    const size_t handled = f.layer_bases.empty() ? 0 : f.layer_bases.back();
    for (size_t b = 0; b < handled; b++) out[b] = f.thr[b];
}
// freePositions as the construction lists them (ShockHash2Flat.h:124-147), one entry per key handed to the fallback.
void orc_flat_free_positions(void *fp, uint64_t *out) {
    const Flat &f = *(Flat *)fp;
    for (size_t i = 0; i < f.free_pos.size(); i++) out[i] = f.free_pos[i];
}
void orc_flat_mapping(uint32_t range, uint32_t *out) { auto m = flat_mapping(range); std::memcpy(out, m.data(), m.size() * 4); }
}  // extern "C"
