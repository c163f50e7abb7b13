// Bipartite ShockHash2 leaf seed search for one wavefront (gfx950 wave64), m <= 64.
//
// Replaces BijectionsShockHash2<m, true, QuadSplitCandidateFinderList>::findSeed (m >= 10) and
// <m, true, BasicSeedCandidateFinder>::findSeed (m < 10), include/shockhash/ShockHash2-internal.h:
// 386-546, dispatched by src/ShockHash2-precompiled.cpp:24-47.
//
// Semantics reproduced (H = ceil(m/2) cells per half, FULL = 2^H - 1):
//  * half-seed s: maskA(s) = OR over even keys of 1 << fastrange64(remix(key + s), H), maskB(s) the
//    same over odd keys (:435-447).
//  * candidates, in ascending code: for s = 0, 1, ...: pairs (a, s), a < s, with
//    maskA(a) | maskB(s) == FULL (code s*s + a), then (s, b), b <= s, with maskA(s) | maskB(b) == FULL
//    (code s*s + s + b) (:449-470). Even keys hash with the first half-seed, odd keys with the second.
//    For m < 10 the candidates are the seeds s with maskA(s) | maskB(s) == FULL, code s (:78-98).
//  * candidate j >= 1 is paired with i = 0 .. j-1 in order; a pair is skipped when some key is the
//    only key of its cell in both candidates (isolated-vertex filter, :34-45,514-518); otherwise it
//    succeeds when the graph with edges (hash_j(key) + m/2, hash_i(key)) is a pseudoforest. The first
//    success returns pairTriangular(code_j, code_i) (:519-541), the smallest successful code.
//
// Two implementations of that search, with identical results:
//  * the register path (sh2_leaf_search_regs): half-seed masks of the first SH2_HS_BLOCKS * 64
//    half-seeds and the isolated-key masks of the first SH2_CAND_BLOCKS * 64 candidates live
//    in VGPRs (lane l of block b holds entry 64 b + l), so candidate extraction and the isolated-vertex
//    filter are ballots over registers. Keys are staged in LDS split by parity and read back as
//    broadcasts. A leaf that needs more half-seeds or candidates (at m = 64 the median is ~100
//    half-seeds and ~40 candidates; ~0.15% of leaves need more than 128 candidates, none of 20,000
//    simulated leaves more than 153 candidates or 221 half-seeds) reports SH2_OVERFLOW and the first j whose
//    pairs are not all tested, and is continued by
//  * the general path (sh2_leaf_search_general), whose tables live in the wave's global scratch area
//    and are bounded only by SH2_MAX_HALF_SEEDS / SH2_MAX_CANDIDATES. It rebuilds the half-seed and
//    candidate tables and tests pairs from (resume j, 0) on (the pairs before it already failed).
// Leaves of 65..128 keys (ShockHash2<L> with L up to MAX_LEAF_SIZE2 = 128, include/ShockHash2.h:29) run the
// general path with two keys per lane (KPL = 2): 64-bit half-seed masks (H <= 64 cells per half), two-word
// isolated-key masks, 128-byte cell rows and a union-find over up to 128 vertices. Their tables grow with the
// leaf size (about (2^k(m) * 2)^(1/4) half-seeds, k(m) = bij_memo2[m]: ~110 at m = 64, ~14,000 at m = 128), so
// their capacities are run-time values the host derives from the largest leaf (Sh2Caps).
// Both write each candidate's cell indices (64 bytes per 64 keys) to scratch and queue the pairs that pass the
// isolated filter in LDS in (j, i) order; 64 queued pairs are tested at a time, one pair per lane,
// with a per-lane union-find (at most one cycle per component, UnionFind.h:58-67 semantics; the smaller
// root becomes the parent, a root's byte carries its cycle flag) whose bytes are interleaved in LDS. Any exact pseudoforest predicate gives
// the reference's result: CuckooUnionFind, UnionFind and the textbook component count agree.
#pragma once
#include <type_traits>

#include "device_common.h"

namespace shgpu {

constexpr uint32_t SH2_MAX_HALF_SEEDS = 1024;  // half-seed bound of leaves up to 64 keys (~2^7 needed at m = 64)
constexpr uint32_t SH2_MAX_CANDIDATES = 4096;
constexpr uint32_t SH2_HS_BLOCKS = 4;    // register path: half-seeds [0, 256)
constexpr uint32_t SH2_CAND_BLOCKS = 3;  // register path: candidates [0, 192)
constexpr uint64_t SH2_OVERFLOW = SHGPU_NOT_FOUND - 1;
constexpr uint32_t SH2_MAX_LEAF = 128;   // MAX_LEAF_SIZE2 (include/ShockHash2.h:29)

// Capacities of the general path's tables: half-seeds hashed and candidates kept per leaf.
struct Sh2Caps {
    uint32_t max_hs, max_cand;
};
constexpr Sh2Caps SH2_NARROW_CAPS{SH2_MAX_HALF_SEEDS, SH2_MAX_CANDIDATES};

__host__ __device__ constexpr size_t sh2_align256(size_t b) { return (b + 255) / 256 * 256; }
// Scratch bytes of one wave: half-seed masks A and B, candidate codes, isolated-key masks, cell rows.
__host__ __device__ constexpr size_t sh2_scratch_bytes(int kpl, Sh2Caps c) {
    return sh2_align256((size_t)c.max_hs * 2 * (kpl == 1 ? 4 : 8)) + sh2_align256((size_t)c.max_cand * 4) +
           sh2_align256((size_t)c.max_cand * 8 * kpl) + (size_t)c.max_cand * 64 * kpl;
}
constexpr size_t SH2_SCRATCH_BYTES = sh2_scratch_bytes(1, SH2_NARROW_CAPS);

template <int KPL> struct Sh2Types;
template <> struct Sh2Types<1> {
    using Mask = uint32_t;   // half-seed mask, H <= 32 cells
    using Queued = uint32_t; // pending pair j << 16 | i (candidates < 2^16)
};
template <> struct Sh2Types<2> {
    using Mask = uint64_t;   // H <= 64 cells
    using Queued = uint64_t; // j << 32 | i
};

template <int KPL>
struct Sh2LdsT {
    uint32_t par[16 * KPL * 64];  // union-find bytes: vertex x of lane l at byte x * 64 + l
    typename Sh2Types<KPL>::Queued queue[128];  // pending pairs (j, i)
    uint32_t hist[64];
    __align__(16) uint64_t keysA[64];  // register path (KPL 1): even keys, then
    __align__(16) uint64_t keysB[64];  // odd keys, in leaf order
};
using Sh2Lds = Sh2LdsT<1>;

template <int KPL>
struct Sh2ScratchT {
    using Mask = typename Sh2Types<KPL>::Mask;
    Mask *hsA, *hsB;
    uint32_t *code;
    uint64_t *iso;    // KPL words per candidate
    uint8_t *hashes;  // 64 * KPL bytes per candidate
    Sh2Caps caps;
    __device__ Sh2ScratchT(uint8_t *base, Sh2Caps c) : caps(c) {
        hsA = (Mask *)base;
        hsB = hsA + c.max_hs;
        code = (uint32_t *)(base + sh2_align256((size_t)c.max_hs * 2 * sizeof(Mask)));
        iso = (uint64_t *)((uint8_t *)code + sh2_align256((size_t)c.max_cand * 4));
        hashes = (uint8_t *)iso + sh2_align256((size_t)c.max_cand * 8 * KPL);
    }
};
using Sh2Scratch = Sh2ScratchT<1>;

template <int KPL>
__device__ __forceinline__ typename Sh2Types<KPL>::Queued sh2_pack(uint32_t j, uint32_t i) {
    return KPL == 1 ? (typename Sh2Types<KPL>::Queued)((j << 16) | i) : (typename Sh2Types<KPL>::Queued)(((uint64_t)j << 32) | i);
}
template <int KPL>
__device__ __forceinline__ uint32_t sh2_pj(typename Sh2Types<KPL>::Queued q) { return KPL == 1 ? (uint32_t)(q >> 16) : (uint32_t)((uint64_t)q >> 32); }
template <int KPL>
__device__ __forceinline__ uint32_t sh2_pi(typename Sh2Types<KPL>::Queued q) { return KPL == 1 ? (uint32_t)(q & 0xFFFF) : (uint32_t)q; }

__device__ __forceinline__ void mem_sync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

// Per-lane union-find over vertices [0, 64) in LDS, one byte per vertex: byte x of lane l at x * 64 + l
// (one shift-add per address). A byte below 0x80 is the parent of a non-root vertex; a root holds 0x80,
// plus 0x40 once its component contains a cycle, plus its rank (union by rank: trees stay at most 6 deep).
constexpr uint32_t UF_ROOT = 0x80, UF_CYCLIC = 0x40, UF_RANK = 0x3F;

// Tests the 64 queued pairs (one per lane). Returns the lane mask of pseudoforest pairs. Each lane
// streams its two candidates' cell bytes four keys at a time from the scratch area. An edge inside a
// component with a cycle, or between two components with cycles, kills the lane (UnionFind.h:58-67:
// a pseudoforest component holds at most one cycle); otherwise the root of higher rank becomes the
// parent. Both ends of an edge walk to their roots in one loop, so a key costs as many dependent LDS
// reads as the deeper of its two walks. Vertices are below m/2 + ceil(m/2) = m <= 64 * KPL.
template <int KPL>
__device__ __forceinline__ uint64_t sh2_uf_batch(const uint8_t *hashes, Sh2LdsT<KPL> &L, uint32_t m) {
    const uint32_t lane = lane_id();
    const auto pr = L.queue[lane];
    const uint32_t j = sh2_pj<KPL>(pr), i = sh2_pi<KPL>(pr);
    const uint32_t *pj = (const uint32_t *)(hashes + (size_t)j * 64 * KPL);
    const uint32_t *pi = (const uint32_t *)(hashes + (size_t)i * 64 * KPL);
#pragma unroll
    for (uint32_t q = 0; q < 16 * KPL; q++) L.par[q * 64 + lane] = UF_ROOT * 0x01010101u;
    wave_sync();
    uint8_t *lp = (uint8_t *)L.par + lane;
    const uint32_t shift = m / 2;
    bool alive = true;
    uint32_t wj = pj[0], wi = pi[0];
    const uint32_t nq = (m + 3) / 4;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t cj = wj, ci = wi;
        if (q + 1 < nq) { wj = pj[q + 1]; wi = pi[q + 1]; }  // prefetch the next four keys
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) {
            if (4 * q + r < m && alive) {
                uint32_t x = ((cj >> (8 * r)) & 0xFF) + shift;
                uint32_t y = (ci >> (8 * r)) & 0xFF;
                uint32_t bx = lp[x << 6], by = lp[y << 6];
                while (bx < UF_ROOT || by < UF_ROOT) {
                    if (bx < UF_ROOT) { x = bx; bx = lp[x << 6]; }
                    if (by < UF_ROOT) { y = by; by = lp[y << 6]; }
                }
                const bool same = x == y, cx = (bx & UF_CYCLIC) != 0, cy = (by & UF_CYCLIC) != 0;
                const uint32_t rx = bx & UF_RANK, ry = by & UF_RANK;
                const bool x_up = rx >= ry;
                const uint32_t root = x_up ? x : y, child = x_up ? y : x;
                const uint32_t rank = rx == ry ? rx + 1 : (x_up ? rx : ry);
                // same component: both stores write the root (x == y) with its cycle flag set
                lp[child << 6] = (uint8_t)(same ? UF_ROOT | UF_CYCLIC | rx : root);
                lp[root << 6] = (uint8_t)(same ? UF_ROOT | UF_CYCLIC | rx : UF_ROOT | (cx || cy ? UF_CYCLIC : 0u) | rank);
                alive = same ? !cx : !(cx && cy);
            }
        }
        if ((q & 1) && ballot(alive) == 0) break;
    }
    const uint64_t res = ballot(alive);
    wave_sync();
    return res;
}

// Cells of this lane's keys (lane + 64 q < m) in candidate (A, B) and the candidate's isolated-key mask
// (iso[q]: keys 64 q .. 64 q + 63); the cells go to row c of the scratch cell table.
template <int KPL>
__device__ __forceinline__ void sh2_candidate_cells(uint8_t *hashes, Sh2LdsT<KPL> &L, uint32_t c, const uint64_t (&kv)[KPL],
                                                    uint32_t m, uint32_t H, uint32_t A, uint32_t B, uint64_t (&iso)[KPL]) {
    const uint32_t lane = lane_id();
    uint32_t h[KPL];
#pragma unroll
    for (int q = 0; q < KPL; q++)
        h[q] = lane + 64 * q < m ? fastrange64(remix(kv[q] + (uint64_t)(((kv[q] & 1) == 0) ? A : B)), H) : 0;
    L.hist[lane] = 0;
    wave_sync();
#pragma unroll
    for (int q = 0; q < KPL; q++)
        if (lane + 64 * q < m) atomicAdd(&L.hist[h[q]], 1u);
    wave_sync();
#pragma unroll
    for (int q = 0; q < KPL; q++) iso[q] = ballot(lane + 64 * q < m && L.hist[h[q]] == 1);
    wave_sync();
#pragma unroll
    for (int q = 0; q < KPL; q++) hashes[(size_t)c * 64 * KPL + 64 * q + lane] = lane + 64 * q < m ? (uint8_t)h[q] : 0;
}

// ------------------------------------------------------------------ general path (scratch tables)

// Appends candidate (A, B) with the given code: cells of every key and its isolated-key mask.
template <int KPL>
__device__ __forceinline__ void sh2_add_candidate(const Sh2ScratchT<KPL> &g, Sh2LdsT<KPL> &L, uint32_t c,
                                                  const uint64_t (&kv)[KPL], uint32_t m, uint32_t H, uint32_t A, uint32_t B,
                                                  uint32_t code) {
    uint64_t iso[KPL];
    sh2_candidate_cells<KPL>(g.hashes, L, c, kv, m, H, A, B, iso);
    g.code[c] = code;  // every lane stores the same wave-uniform values
#pragma unroll
    for (int q = 0; q < KPL; q++) g.iso[(size_t)c * KPL + q] = iso[q];
}

// Returns the ShockHash2 leaf code, or SHGPU_NOT_FOUND when a table capacity is exceeded.
// Pairs (j, i) with j < j_start are not tested: the register path has tested them without success.
template <int KPL>
__device__ __forceinline__ uint64_t sh2_leaf_search_general(const uint64_t *keys, uint32_t m, uint8_t *scratch,
                                                            Sh2LdsT<KPL> &L, Sh2Caps caps, uint64_t t_start,
                                                            uint32_t j_start) {
    using Mask = typename Sh2Types<KPL>::Mask;
    const uint32_t lane = lane_id();
    if (m < 2) return 0;
    const Sh2ScratchT<KPL> g(scratch, caps);
    const uint32_t H = (m + 1) / 2;
    const Mask FULL = H >= 8 * sizeof(Mask) ? ~(Mask)0 : (((Mask)1 << H) - 1);
    const bool basic = m < 10;
    uint64_t kv[KPL];
#pragma unroll
    for (int q = 0; q < KPL; q++) kv[q] = lane + 64 * q < m ? keys[lane + 64 * q] : 0;
    uint32_t n_hs = 0, s_next = 0, n_cand = 0, jq = j_start > 1 ? j_start : 1, iq = 0, qlen = 0;
    for (;;) {
        // Make candidate jq available.
        while (n_cand <= jq) {
            if (s_next == n_hs) {
                if (n_hs + 64 > caps.max_hs) return SHGPU_NOT_FOUND;
                const uint64_t s = n_hs + lane;
                Mask ma = 0, mb = 0;
#pragma unroll
                for (int q = 0; q < KPL; q++) {
                    const uint32_t kn = m - 64 * q < 64 ? m - 64 * q : 64;
                    for (uint32_t k = 0; k < kn && 64 * q < m; k++) {
                        const uint64_t key = readlane64(kv[q], k);
                        const Mask bit = (Mask)1 << fastrange64(remix(key + s), H);
                        if ((key & 1) == 0) ma |= bit; else mb |= bit;
                    }
                }
                g.hsA[n_hs + lane] = ma;
                g.hsB[n_hs + lane] = mb;
                n_hs += 64;
                mem_sync();
            }
            const uint32_t s = s_next++;
            const Mask ms_a = g.hsA[s], ms_b = g.hsB[s];
            if (basic) {
                if ((ms_a | ms_b) == FULL) {
                    if (n_cand >= caps.max_cand) return SHGPU_NOT_FOUND;
                    sh2_add_candidate<KPL>(g, L, n_cand++, kv, m, H, s, s, s);
                    mem_sync();
                }
                continue;
            }
            // (a, s), a < s: code s*s + a
            for (uint32_t c0 = 0; c0 < s; c0 += 64) {
                const uint32_t a = c0 + lane;
                uint64_t hit = ballot(a < s && (g.hsA[a] | ms_b) == FULL);
                while (hit) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(hit);
                    hit &= hit - 1;
                    if (n_cand >= caps.max_cand) return SHGPU_NOT_FOUND;
                    sh2_add_candidate<KPL>(g, L, n_cand++, kv, m, H, c0 + l, s, s * s + c0 + l);
                }
            }
            // (s, b), b <= s: code s*s + s + b
            for (uint32_t c0 = 0; c0 <= s; c0 += 64) {
                const uint32_t b = c0 + lane;
                uint64_t hit = ballot(b <= s && (ms_a | g.hsB[b]) == FULL);
                while (hit) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(hit);
                    hit &= hit - 1;
                    if (n_cand >= caps.max_cand) return SHGPU_NOT_FOUND;
                    sh2_add_candidate<KPL>(g, L, n_cand++, kv, m, H, s, c0 + l, s * s + s + c0 + l);
                }
            }
            mem_sync();
        }
        // Queue pairs (jq, i), i in [iq, iq + 64) passing the isolated-vertex filter.
        age_priority(t_start);
        {
            const uint32_t i = iq + lane;
            bool pass = i < jq;
#pragma unroll
            for (int q = 0; q < KPL; q++)
                pass = pass && (g.iso[(size_t)i * KPL + q] & g.iso[(size_t)jq * KPL + q]) == 0;
            const uint64_t b = ballot(pass);
            if (pass) L.queue[qlen + lanes_below(b)] = sh2_pack<KPL>(jq, i);
            qlen += (uint32_t)__builtin_popcountll(b);
            iq += 64;
            if (iq >= jq) { jq++; iq = 0; }
            wave_sync();
        }
        if (qlen >= 64) {
            const uint64_t ok = sh2_uf_batch<KPL>(g.hashes, L, m);
            if (ok) {
                const uint32_t l = (uint32_t)__builtin_ctzll(ok);
                const auto pr = L.queue[l];
                const uint64_t cj = g.code[sh2_pj<KPL>(pr)], ci = g.code[sh2_pi<KPL>(pr)];
                return cj * (cj - 1) / 2 + ci;  // pairTriangular (PairingFunction.h:47-50)
            }
            const uint32_t rest = qlen - 64;
            const auto moved = lane < rest ? L.queue[64 + lane] : L.queue[0];
            wave_sync();
            if (lane < rest) L.queue[lane] = moved;
            qlen = rest;
            wave_sync();
        }
    }
}

// ------------------------------------------------------------------ register path

// Four (or two) per-lane registers addressed by a wave-uniform block number through compare-and-select.
// Separate members rather than an array: the compiler turns a dynamically indexed array into scratch
// memory, while these stay in VGPRs.
template <class T>
struct Regs4 {
    T r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    __device__ __forceinline__ T at(uint32_t b) const { return b == 0 ? r0 : b == 1 ? r1 : b == 2 ? r2 : r3; }
    __device__ __forceinline__ void set_if(bool c, uint32_t b, T v) {
        r0 = c && b == 0 ? v : r0;
        r1 = c && b == 1 ? v : r1;
        r2 = c && b == 2 ? v : r2;
        r3 = c && b == 3 ? v : r3;
    }
};
template <class T>
struct Regs2 {
    T r0 = 0, r1 = 0;
    __device__ __forceinline__ T at(uint32_t b) const { return b == 0 ? r0 : r1; }
    __device__ __forceinline__ void set_if(bool c, uint32_t b, T v) {
        r0 = c && b == 0 ? v : r0;
        r1 = c && b == 1 ? v : r1;
    }
};
template <class T>
struct Regs3 {
    T r0 = 0, r1 = 0, r2 = 0;
    __device__ __forceinline__ T at(uint32_t b) const { return b == 0 ? r0 : b == 1 ? r1 : r2; }
    __device__ __forceinline__ void set_if(bool c, uint32_t b, T v) {
        r0 = c && b == 0 ? v : r0;
        r1 = c && b == 1 ? v : r1;
        r2 = c && b == 2 ? v : r2;
    }
};
static_assert(SH2_HS_BLOCKS == 4 && SH2_CAND_BLOCKS == 3, "register tables are Regs4 / Regs3");

// OR of 1 << fastrange64(remix(key + s), H) over n staged keys (broadcast LDS reads, two keys per read).
__device__ __forceinline__ uint32_t sh2_half_mask(const uint64_t *stage, uint32_t n, uint64_t s, uint32_t H) {
    const ulonglong2 *st2 = (const ulonglong2 *)stage;
    uint32_t mask = 0;
    uint32_t k = 0;
    for (; k + 4 <= n; k += 4) {
        const ulonglong2 k0 = st2[k / 2], k1 = st2[k / 2 + 1];
        mask |= 1u << fastrange64(remix(k0.x + s), H);
        mask |= 1u << fastrange64(remix(k0.y + s), H);
        mask |= 1u << fastrange64(remix(k1.x + s), H);
        mask |= 1u << fastrange64(remix(k1.y + s), H);
    }
    for (; k < n; k++) mask |= 1u << fastrange64(remix(stage[k] + s), H);
    return mask;
}

// Returns the ShockHash2 leaf code, or SH2_OVERFLOW when the leaf needs more half-seeds or candidates
// than the registers hold; then every pair (j, i) with j < resume_j has been tested without success.
__device__ __forceinline__ uint64_t sh2_leaf_search_regs(const uint64_t *keys, uint32_t m, uint8_t *scratch,
                                                         Sh2Lds &L, uint64_t t_start, uint32_t &resume_j) {
    const uint32_t lane = lane_id();
    if (m < 2) return 0;
    const Sh2Scratch g(scratch, SH2_NARROW_CAPS);
    const uint32_t H = (m + 1) / 2;
    const uint32_t FULL = H >= 32 ? 0xFFFFFFFFu : ((1u << H) - 1u);
    const bool basic = m < 10;
    const bool act = lane < m;
    const uint64_t kv = act ? keys[lane] : 0;
    // stage even keys in keysA and odd keys in keysB (the masks are ORs, the order within a set is free)
    const bool even = act && (kv & 1) == 0;
    const uint64_t be = ballot(even), bo = ballot(act && !even);
    const uint32_t nA = (uint32_t)__builtin_popcountll(be), nB = (uint32_t)__builtin_popcountll(bo);
    wave_sync();
    if (even) L.keysA[lanes_below(be)] = kv;
    else if (act) L.keysB[lanes_below(bo)] = kv;
    wave_sync();

    Regs4<uint32_t> hsA, hsB;
    Regs3<uint64_t> isoR;
    uint32_t n_blocks = 0, s_next = 0, n_cand = 0, jq = 1, qlen = 0;
    // Overflow happens while candidate jq is made available: the pairs of every j < jq are queued, and the
    // queued pairs not yet tested start at the first queue entry's j.
    auto overflow = [&]() -> uint64_t {
        resume_j = qlen ? sh2_pj<1>(L.queue[0]) : jq;
        return SH2_OVERFLOW;
    };

    // Appends candidate (A, B) with the given code; false when the register table is full. The code goes to
    // the scratch table (read once, for the winner).
    auto add = [&](uint32_t A, uint32_t B, uint32_t code) -> bool {
        if (n_cand >= SH2_CAND_BLOCKS * 64) return false;
        uint64_t iso[1];
        const uint64_t kvs[1] = {kv};
        sh2_candidate_cells<1>(g.hashes, L, n_cand, kvs, m, H, A, B, iso);
        isoR.set_if(lane == (n_cand & 63), n_cand >> 6, iso[0]);
        g.code[n_cand] = code;  // same value in every lane
        n_cand++;
        return true;
    };

    for (;;) {
        // Make candidate jq available.
        while (n_cand <= jq) {
            if (s_next == 64 * n_blocks) {
                if (n_blocks == SH2_HS_BLOCKS) return overflow();
                const uint64_t s = 64 * n_blocks + lane;
                const uint32_t ma = sh2_half_mask(L.keysA, nA, s, H), mb = sh2_half_mask(L.keysB, nB, s, H);
                hsA.set_if(true, n_blocks, ma);
                hsB.set_if(true, n_blocks, mb);
                n_blocks++;
            }
            const uint32_t s = s_next++;
            const uint32_t sb = s >> 6, sl = s & 63;
            const uint32_t ms_a = (uint32_t)__builtin_amdgcn_readlane((int)hsA.at(sb), (int)sl);
            const uint32_t ms_b = (uint32_t)__builtin_amdgcn_readlane((int)hsB.at(sb), (int)sl);
            if (basic) {
                if ((ms_a | ms_b) == FULL && !add(s, s, s)) return overflow();
                continue;
            }
            // Hits of (a, s), a < s (code s*s + a) and of (s, b), b <= s (code s*s + s + b), one ballot per
            // register block; then the candidates in code order (block, then lane, side (a, s) first).
            uint64_t hA0 = 0, hA1 = 0, hA2 = 0, hA3 = 0, hB0 = 0, hB1 = 0, hB2 = 0, hB3 = 0;
            hA0 = ballot(lane < s && (hsA.r0 | ms_b) == FULL);
            hB0 = ballot(lane <= s && (ms_a | hsB.r0) == FULL);
            if (sb >= 1) {
                hA1 = ballot(64 + lane < s && (hsA.r1 | ms_b) == FULL);
                hB1 = ballot(64 + lane <= s && (ms_a | hsB.r1) == FULL);
            }
            if (sb >= 2) {
                hA2 = ballot(128 + lane < s && (hsA.r2 | ms_b) == FULL);
                hB2 = ballot(128 + lane <= s && (ms_a | hsB.r2) == FULL);
            }
            if (sb >= 3) {
                hA3 = ballot(192 + lane < s && (hsA.r3 | ms_b) == FULL);
                hB3 = ballot(192 + lane <= s && (ms_a | hsB.r3) == FULL);
            }
            for (;;) {
                uint32_t x, side;
#define SH2_TAKE(H, BASE, SIDE) { x = BASE + (uint32_t)__builtin_ctzll(H); H &= H - 1; side = SIDE; }
                if (hA0) SH2_TAKE(hA0, 0, 0)
                else if (hA1) SH2_TAKE(hA1, 64, 0)
                else if (hA2) SH2_TAKE(hA2, 128, 0)
                else if (hA3) SH2_TAKE(hA3, 192, 0)
                else if (hB0) SH2_TAKE(hB0, 0, 1)
                else if (hB1) SH2_TAKE(hB1, 64, 1)
                else if (hB2) SH2_TAKE(hB2, 128, 1)
                else if (hB3) SH2_TAKE(hB3, 192, 1)
                else break;
#undef SH2_TAKE
                const bool ok = side == 0 ? add(x, s, s * s + x) : add(s, x, s * s + s + x);
                if (!ok) return overflow();
            }
        }
        // Queue the pairs (jq, i), i < jq, that pass the isolated-vertex filter, 64 i at a time, and test
        // whenever 64 are queued.
        age_priority(t_start);
        const uint64_t iso_j = readlane64(isoR.at(jq >> 6), jq & 63);
        for (uint32_t b = 0; 64 * b < jq; b++) {
            const uint32_t i = 64 * b + lane;
            const bool pass = i < jq && (isoR.at(b) & iso_j) == 0;
            const uint64_t bits = ballot(pass);
            if (pass) L.queue[qlen + lanes_below(bits)] = sh2_pack<1>(jq, i);
            qlen += (uint32_t)__builtin_popcountll(bits);
            wave_sync();
            if (qlen >= 64) {
                const uint64_t ok = sh2_uf_batch<1>(g.hashes, L, m);
                if (ok) {
                    const uint32_t pr = L.queue[__builtin_ctzll(ok)];
                    const uint64_t cj = g.code[sh2_pj<1>(pr)], ci = g.code[sh2_pi<1>(pr)];
                    return cj * (cj - 1) / 2 + ci;  // pairTriangular (PairingFunction.h:47-50)
                }
                const uint32_t rest = qlen - 64;
                const uint32_t moved = lane < rest ? L.queue[64 + lane] : 0;
                wave_sync();
                if (lane < rest) L.queue[lane] = moved;
                qlen = rest;
                wave_sync();
            }
        }
        jq++;
    }
}

// Tuning flag: search every ShockHash2 leaf with the general path only (tests compare both paths).
constexpr uint32_t SH2_FLAG_GENERAL_ONLY = 1;

// The ShockHash2 leaf code of m <= 64 keys, or SHGPU_NOT_FOUND when the general path's bounds are exceeded. The
// wave's issue priority rises with the time since t_start (task_clock) and is reset on return.
__device__ __forceinline__ uint64_t sh2_leaf_search(const uint64_t *keys, uint32_t m, uint8_t *scratch, Sh2Lds &L,
                                                    uint32_t flags, uint64_t t_start) {
    uint64_t r = SH2_OVERFLOW;
    uint32_t resume_j = 1;
    if (!(flags & SH2_FLAG_GENERAL_ONLY)) r = sh2_leaf_search_regs(keys, m, scratch, L, t_start, resume_j);
    if (r == SH2_OVERFLOW) r = sh2_leaf_search_general<1>(keys, m, scratch, L, SH2_NARROW_CAPS, t_start, uniform(resume_j));
    reset_priority();
    return r;
}

// The same for 64 < m <= 128 keys (two keys per lane, tables of the given capacities).
__device__ __forceinline__ uint64_t sh2_leaf_search_wide(const uint64_t *keys, uint32_t m, uint8_t *scratch, Sh2LdsT<2> &L,
                                                         Sh2Caps caps, uint64_t t_start) {
    const uint64_t r = sh2_leaf_search_general<2>(keys, m, scratch, L, caps, t_start, 1);
    reset_priority();
    return r;
}

// LDS of one wave that searches leaves of up to 64 keys, or up to 128 (WIDE) with either path.
template <bool WIDE>
union Sh2LdsAny {
    Sh2Lds narrow;
    typename std::conditional<WIDE, Sh2LdsT<2>, char>::type wide;
};

// Either search by the leaf's size; `caps` and the scratch size are those of the widest leaf (narrow leaves use the
// start of the wave's area).
template <bool WIDE>
__device__ __forceinline__ uint64_t sh2_leaf_search_any(const uint64_t *keys, uint32_t m, uint8_t *scratch,
                                                        Sh2LdsAny<WIDE> &L, Sh2Caps caps, uint32_t flags, uint64_t t_start) {
    if constexpr (WIDE) {
        if (m > 64) return sh2_leaf_search_wide(keys, m, scratch, L.wide, caps, t_start);
    }
    return sh2_leaf_search(keys, m, scratch, L.narrow, flags, t_start);
}

}  // namespace shgpu
