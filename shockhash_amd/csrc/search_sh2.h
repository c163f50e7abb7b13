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
// Mapping: half-seed masks are computed 64 at a time (lanes over half-seeds). Candidates are
// extracted with lanes over the earlier half-seeds (a linear scan of the mask table, as the
// reference's CandidateList does). Each candidate's 64 cell indices and isolated-key mask go to a
// per-wave global scratch area. Pairs passing the isolated filter are queued in LDS in (j, i) order and
// tested 64 at a time, one pair per lane, with a per-lane union-find (tree flags, UnionFind.h:58-67
// semantics) whose parent bytes are interleaved in LDS. Any exact pseudoforest predicate gives the
// reference's result: CuckooUnionFind, UnionFind and the textbook component count agree.
#pragma once
#include "device_common.h"

namespace shgpu {

constexpr uint32_t SH2_MAX_HALF_SEEDS = 1024;  // half-seed bound (reference expects ~2^7 at m = 64)
constexpr uint32_t SH2_MAX_CANDIDATES = 4096;
constexpr size_t SH2_SCRATCH_BYTES =
    SH2_MAX_HALF_SEEDS * 4 * 2 + SH2_MAX_CANDIDATES * (4 + 8 + 64);

struct Sh2Lds {
    uint32_t par[16 * 64];  // union-find parents: byte x of lane l at dword (x >> 2) * 64 + l
    uint32_t queue[128];    // pending pairs j << 16 | i
    uint32_t hist[64];
};

struct Sh2Scratch {
    uint32_t *hsA, *hsB, *code;
    uint64_t *iso;
    uint8_t *hashes;
    __device__ explicit Sh2Scratch(uint8_t *base) {
        hsA = (uint32_t *)base;
        hsB = hsA + SH2_MAX_HALF_SEEDS;
        code = hsB + SH2_MAX_HALF_SEEDS;
        iso = (uint64_t *)(code + SH2_MAX_CANDIDATES);
        hashes = (uint8_t *)(iso + SH2_MAX_CANDIDATES);
    }
};

__device__ __forceinline__ void mem_sync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t uf_addr(uint32_t x, uint32_t lane) { return ((x >> 2) * 64 + lane) * 4 + (x & 3); }

__device__ __forceinline__ uint32_t uf_find(const uint8_t *par, uint32_t x, uint32_t lane) {
    for (;;) {
        const uint32_t p = par[uf_addr(x, lane)];
        if (p == x) return x;
        x = p;
    }
}

// Tests the 64 queued pairs (one per lane). Returns the lane mask of pseudoforest pairs. Each lane
// streams its two candidates' cell bytes four keys at a time from the scratch area.
__device__ __forceinline__ uint64_t sh2_uf_batch(const Sh2Scratch &g, Sh2Lds &L, uint32_t m) {
    const uint32_t lane = lane_id();
    const uint32_t pr = L.queue[lane];
    const uint32_t j = pr >> 16, i = pr & 0xFFFF;
    const uint32_t *pj = (const uint32_t *)(g.hashes + (size_t)j * 64);
    const uint32_t *pi = (const uint32_t *)(g.hashes + (size_t)i * 64);
    uint8_t *par = (uint8_t *)L.par;
#pragma unroll
    for (uint32_t q = 0; q < 16; q++) {
        const uint32_t x = 4 * q;
        L.par[q * 64 + lane] = x | ((x + 1) << 8) | ((x + 2) << 16) | ((x + 3) << 24);
    }
    wave_sync();
    const uint32_t shift = m / 2;
    uint64_t tree = ~0ull;
    bool alive = true;
    uint32_t wj = pj[0], wi = pi[0];
    const uint32_t nq = (m + 3) / 4;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t cj = wj, ci = wi;
        if (q + 1 < nq) { wj = pj[q + 1]; wi = pi[q + 1]; }  // prefetch the next four keys
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) {
            if (4 * q + r < m && alive) {
                const uint32_t u = ((cj >> (8 * r)) & 0xFF) + shift;
                const uint32_t v = (ci >> (8 * r)) & 0xFF;
                const uint32_t ru = uf_find(par, u, lane);
                const uint32_t rv = uf_find(par, v, lane);
                par[uf_addr(ru, lane)] = (uint8_t)rv;
                const bool tu = (tree >> ru) & 1, tv = (tree >> rv) & 1;
                if (!(tu && tv && ru != rv)) tree &= ~(1ull << rv);
                alive = tu || tv;
            }
        }
        if ((q & 1) && ballot(alive) == 0) break;
    }
    const uint64_t res = ballot(alive);
    wave_sync();
    return res;
}

// Appends candidate (A, B) with the given code: cells of every key and its isolated-key mask.
__device__ __forceinline__ void sh2_add_candidate(const Sh2Scratch &g, Sh2Lds &L, uint32_t c, uint64_t kv, uint32_t m,
                                         uint32_t H, uint32_t A, uint32_t B, uint32_t code) {
    const uint32_t lane = lane_id();
    const bool act = lane < m;
    uint32_t h = 0;
    if (act) h = fastrange64(remix(kv + (uint64_t)(((kv & 1) == 0) ? A : B)), H);
    L.hist[lane] = 0;
    wave_sync();
    if (act) atomicAdd(&L.hist[h], 1u);
    wave_sync();
    const uint64_t iso = ballot(act && L.hist[h] == 1);
    wave_sync();
    g.hashes[(size_t)c * 64 + lane] = act ? (uint8_t)h : 0;
    g.code[c] = code;  // every lane stores the same wave-uniform values
    g.iso[c] = iso;
}

// Returns the ShockHash2 leaf code, or SHGPU_NOT_FOUND when a scratch bound is exceeded.
__device__ __forceinline__ uint64_t sh2_leaf_search(const uint64_t *keys, uint32_t m, uint8_t *scratch, Sh2Lds &L) {
    const uint32_t lane = lane_id();
    if (m < 2) return 0;
    const Sh2Scratch g(scratch);
    const uint32_t H = (m + 1) / 2;
    const uint32_t FULL = H >= 32 ? 0xFFFFFFFFu : ((1u << H) - 1u);
    const bool basic = m < 10;
    const uint64_t kv = lane < m ? keys[lane] : 0;
    uint32_t n_hs = 0, s_next = 0, n_cand = 0, jq = 1, iq = 0, qlen = 0;
    for (;;) {
        // Make candidate jq available.
        while (n_cand <= jq) {
            if (s_next == n_hs) {
                if (n_hs + 64 > SH2_MAX_HALF_SEEDS) return SHGPU_NOT_FOUND;
                const uint64_t s = n_hs + lane;
                uint32_t ma = 0, mb = 0;
                for (uint32_t k = 0; k < m; k++) {
                    const uint64_t key = readlane64(kv, k);
                    const uint32_t bit = 1u << fastrange64(remix(key + s), H);
                    if ((key & 1) == 0) ma |= bit; else mb |= bit;
                }
                g.hsA[n_hs + lane] = ma;
                g.hsB[n_hs + lane] = mb;
                n_hs += 64;
                mem_sync();
            }
            const uint32_t s = s_next++;
            const uint32_t ms_a = g.hsA[s], ms_b = g.hsB[s];
            if (basic) {
                if ((ms_a | ms_b) == FULL) {
                    if (n_cand >= SH2_MAX_CANDIDATES) return SHGPU_NOT_FOUND;
                    sh2_add_candidate(g, L, n_cand++, kv, m, H, s, s, s);
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
                    if (n_cand >= SH2_MAX_CANDIDATES) return SHGPU_NOT_FOUND;
                    sh2_add_candidate(g, L, n_cand++, kv, m, H, c0 + l, s, s * s + c0 + l);
                }
            }
            // (s, b), b <= s: code s*s + s + b
            for (uint32_t c0 = 0; c0 <= s; c0 += 64) {
                const uint32_t b = c0 + lane;
                uint64_t hit = ballot(b <= s && (ms_a | g.hsB[b]) == FULL);
                while (hit) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(hit);
                    hit &= hit - 1;
                    if (n_cand >= SH2_MAX_CANDIDATES) return SHGPU_NOT_FOUND;
                    sh2_add_candidate(g, L, n_cand++, kv, m, H, s, c0 + l, s * s + s + c0 + l);
                }
            }
            mem_sync();
        }
        // Queue pairs (jq, i), i in [iq, iq + 64) passing the isolated-vertex filter.
        {
            const uint64_t iso_j = g.iso[jq];
            const uint32_t i = iq + lane;
            const bool pass = i < jq && (g.iso[i] & iso_j) == 0;
            const uint64_t b = ballot(pass);
            if (pass) L.queue[qlen + lanes_below(b)] = (jq << 16) | i;
            qlen += (uint32_t)__builtin_popcountll(b);
            iq += 64;
            if (iq >= jq) { jq++; iq = 0; }
            wave_sync();
        }
        if (qlen >= 64) {
            const uint64_t ok = sh2_uf_batch(g, L, m);
            if (ok) {
                const uint32_t l = (uint32_t)__builtin_ctzll(ok);
                const uint32_t pr = L.queue[l];
                const uint64_t cj = g.code[pr >> 16], ci = g.code[pr & 0xFFFF];
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
}

}  // namespace shgpu
