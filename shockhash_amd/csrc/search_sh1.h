// ShockHash1 leaf seed search for one wavefront (gfx950 wave64).
//
// Replaces the leaf loop of ShockHash::recSplit (include/ShockHash.h:404-434), the non-rotation
// leafLevel of SIMDShockHash (include/SIMDShockHash.hpp:507-542) and the isolated base case
// BijectionsShockHash1::calculateBijection (benchmark/bijections/ShockHash1.h:13-35). All three
// return the smallest seed x >= start whose candidate cells
//   r = remix(key + x); c1 = fastrange32(lo32 r, m/2); c2 = fastrange32(hi32 r, (m+1)/2) + m/2
// (TinyBinaryCuckooHashTable.h:138-149) cover all m cells and form a pseudoforest (the cuckoo
// insertion of TinyBinaryCuckooHashTable::construct succeeds exactly then).
//
// Mapping: lanes over seeds (S seeds per lane, window of 64*S consecutive seeds), keys staged in LDS
// and read back as wave-wide broadcasts. The coverage mask is split into two 32-bit halves (cells
// [0, m/2) and [m/2, m)), one v_lshl_or each. Candidates that cover every cell go, in seed order,
// through a wave-cooperative pseudoforest test (lanes over keys, min-label propagation in LDS).
#pragma once
#include "device_common.h"

namespace shgpu {

// LDS of one wave for the windowed searches: the key stage (split and leaf windows) and the labels and
// counters of pseudoforest_wave().
struct PfLds {
    __align__(16) uint64_t stage[STAGE_KEYS];
    uint32_t lab[64];
    uint32_t cnt[128];
};

// Exact pseudoforest test of the multigraph with vertices [0, nv) and one edge (a, b) per active
// lane (nv <= 64, at most 64 edges). Every component must have #edges <= #vertices. Components are
// found by min-label propagation with pointer jumping; labels are always vertices of the same
// component and only decrease, so the loop ends (at most nv*nv rounds, a handful in practice).
__device__ __forceinline__ bool pseudoforest_wave(uint32_t a, uint32_t b, bool act, uint32_t nv, PfLds &L) {
    const uint32_t lane = lane_id();
    if (lane < nv) { L.lab[lane] = lane; L.cnt[lane] = 0; L.cnt[64 + lane] = 0; }
    wave_sync();
    for (uint32_t it = 0; it < 64 * 64; it++) {
        uint32_t la = 0, lb = 0;
        if (act) { la = L.lab[a]; lb = L.lab[b]; }
        const bool diff = act && la != lb;
        if (ballot(diff) == 0) break;
        if (diff) {
            const uint32_t mn = la < lb ? la : lb;
            atomicMin(&L.lab[a], mn);
            atomicMin(&L.lab[b], mn);
            atomicMin(&L.lab[la], mn);
            atomicMin(&L.lab[lb], mn);
        }
        wave_sync();
        if (lane < nv) {
            const uint32_t l = L.lab[lane];
            const uint32_t l2 = L.lab[l];
            L.lab[lane] = l2;
        }
        wave_sync();
    }
    if (act) atomicAdd(&L.cnt[L.lab[a]], 1u);            // edges per component
    if (lane < nv) atomicAdd(&L.cnt[64 + L.lab[lane]], 1u);  // vertices per component
    wave_sync();
    const bool bad = lane < nv && L.cnt[lane] > L.cnt[64 + lane];
    const bool ok = ballot(bad) == 0;
    wave_sync();
    return ok;
}

__device__ __forceinline__ uint32_t mask_bits(uint32_t n) { return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u); }

// Evaluates one key for S seeds: OR the candidate cells into the two half masks.
template <int S>
__device__ __forceinline__ void sh1_eval(uint64_t key, const uint64_t (&seed)[S], uint32_t halfA, uint32_t halfB,
                                         uint32_t (&mA)[S], uint32_t (&mB)[S]) {
#pragma unroll
    for (int s = 0; s < S; s++) {
        const uint64_t r = remix(key + seed[s]);
        mA[s] |= 1u << __umulhi((uint32_t)r, halfA);
        mB[s] |= 1u << __umulhi((uint32_t)(r >> 32), halfB);
    }
}

// Stages the m <= 64 keys of a leaf in L.stage (read back as broadcasts) and returns this lane's key.
__device__ __forceinline__ uint64_t sh1_stage(const uint64_t *keys, uint32_t m, PfLds &L) {
    const uint32_t lane = lane_id();
    const uint64_t kv = lane < m ? keys[lane] : 0;
    wave_sync();
    if (lane < m) L.stage[lane] = kv;
    wave_sync();
    return kv;
}

// One window of a leaf search over keys staged by sh1_stage (kv = this lane's key): seeds
// start + woff + [0, 64*S), m >= 2 keys (m <= 64). Returns the offset (seed - start) of the smallest
// qualifying seed in the window or SHGPU_NOT_FOUND.
template <int S>
__device__ __forceinline__ uint64_t sh1_window_staged(uint64_t kv, uint32_t m, uint64_t start, uint64_t woff, PfLds &L) {
    const uint32_t lane = lane_id();
    const uint32_t halfA = m / 2, halfB = (m + 1) / 2;
    const uint32_t fullA = mask_bits(halfA), fullB = mask_bits(halfB);
    uint64_t seed[S];
    uint32_t mA[S], mB[S];
#pragma unroll
    for (int s = 0; s < S; s++) { seed[s] = start + woff + (uint64_t)(s * 64 + lane); mA[s] = 0; mB[s] = 0; }
    const ulonglong2 *st2 = (const ulonglong2 *)L.stage;
    if (m == 64) {
#pragma unroll
        for (uint32_t k = 0; k < 32; k++) {
            const ulonglong2 kk = st2[k];
            sh1_eval<S>(kk.x, seed, halfA, halfB, mA, mB);
            sh1_eval<S>(kk.y, seed, halfA, halfB, mA, mB);
        }
    } else {
        uint32_t k = 0;
        for (; k + 4 <= m; k += 4) {
            const ulonglong2 k0 = st2[k / 2], k1 = st2[k / 2 + 1];
            sh1_eval<S>(k0.x, seed, halfA, halfB, mA, mB);
            sh1_eval<S>(k0.y, seed, halfA, halfB, mA, mB);
            sh1_eval<S>(k1.x, seed, halfA, halfB, mA, mB);
            sh1_eval<S>(k1.y, seed, halfA, halfB, mA, mB);
        }
        for (; k < m; k++) sh1_eval<S>(L.stage[k], seed, halfA, halfB, mA, mB);
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        uint64_t cand = ballot(mA[s] == fullA && mB[s] == fullB);
        while (cand) {
            const uint32_t l = (uint32_t)__builtin_ctzll(cand);
            cand &= cand - 1;
            const uint64_t off = woff + (uint64_t)(s * 64 + l);
            const uint64_t x = start + off;
            uint32_t a = 0, b = 0;
            if (lane < m) {
                const uint64_t r = remix(kv + x);
                a = __umulhi((uint32_t)r, halfA);
                b = __umulhi((uint32_t)(r >> 32), halfB) + halfA;
            }
            if (pseudoforest_wave(a, b, lane < m, m, L)) return off;
        }
    }
    return SHGPU_NOT_FOUND;
}

// ROTATION_FITTING full leaf (m == LEAF_SIZE): include/ShockHash.h:316-403, include/SIMDShockHash.hpp:419-506.
// Keys with even value (group A) take cells of seed x & ~amask (amask 7 above 32 keys, GROUP_A_HF_MASK at
// ShockHash.h:141), odd keys (group B) cells of seed x rotated by r (cell + r mod m). The reference's result is
// the smallest x, and for it the smallest r, whose cells cover all m cells and admit the cuckoo insertions (a
// pseudoforest). Lanes over seeds as in sh1_window_staged: full 64-bit coverage masks of A and B per seed, the
// rotations r whose A | rotl(B, r) covers every cell collected in a mask, then candidates tested in (x, r)
// order with the wave's pseudoforest test. Returns (seed offset << 6) | r of the first success in the window
// (increasing in (x, r), so the window protocol's minimum is the reference's result), or SHGPU_NOT_FOUND.
__device__ __forceinline__ uint64_t rot_mask_of(uint32_t m) { return m > 32 ? 7ull : 0ull; }
__device__ __forceinline__ uint64_t rotl_cells(uint64_t v, uint32_t r, uint32_t m, uint64_t all) {
    return r == 0 ? v : ((v << r) | (v >> (m - r))) & all;
}

template <int S>
__device__ __forceinline__ uint64_t sh1_rot_window_staged(uint64_t kv, uint32_t m, uint64_t start, uint64_t woff, PfLds &L) {
    const uint32_t lane = lane_id();
    const uint32_t halfA = m / 2, halfB = (m + 1) / 2;
    const uint64_t all = m == 64 ? ~0ull : ((1ull << m) - 1), amask = rot_mask_of(m);
    uint64_t seed[S], ma[S], mb[S];
#pragma unroll
    for (int s = 0; s < S; s++) { seed[s] = start + woff + (uint64_t)(s * 64 + lane); ma[s] = 0; mb[s] = 0; }
    for (uint32_t k = 0; k < m; k++) {
        const uint64_t key = L.stage[k];
        const bool odd = (key & 1) != 0;  // wave-uniform
#pragma unroll
        for (int s = 0; s < S; s++) {
            const uint64_t r = remix(key + (odd ? seed[s] : (seed[s] & ~amask)));
            const uint64_t bits = (1ull << __umulhi((uint32_t)r, halfA)) | (1ull << (__umulhi((uint32_t)(r >> 32), halfB) + halfA));
            if (odd) mb[s] |= bits;
            else ma[s] |= bits;
        }
    }
    uint64_t rok[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        rok[s] = 0;
        for (uint32_t r = 0; r < m; r++)
            if ((ma[s] | rotl_cells(mb[s], r, m, all)) == all) rok[s] |= 1ull << r;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        uint64_t cand = ballot(rok[s] != 0);
        while (cand) {
            const uint32_t l = (uint32_t)__builtin_ctzll(cand);
            cand &= cand - 1;
            const uint64_t off = woff + (uint64_t)(s * 64 + l);
            const uint64_t x = start + off;
            uint64_t rs = readlane64(rok[s], l);
            // this lane's key (lane < m) under seed x: its cells before rotation
            uint32_t a = 0, b = 0;
            const bool odd = (kv & 1) != 0;
            if (lane < m) {
                const uint64_t r = remix(kv + (odd ? x : (x & ~amask)));
                a = __umulhi((uint32_t)r, halfA);
                b = __umulhi((uint32_t)(r >> 32), halfB) + halfA;
            }
            while (rs) {
                const uint32_t rot = (uint32_t)__builtin_ctzll(rs);
                rs &= rs - 1;
                uint32_t ra = a, rb = b;
                if (odd) {
                    ra = a + rot >= m ? a + rot - m : a + rot;
                    rb = b + rot >= m ? b + rot - m : b + rot;
                }
                if (pseudoforest_wave(ra, rb, lane < m, m, L)) return (off << 6) | rot;
            }
        }
    }
    return SHGPU_NOT_FOUND;
}

// Sequential leaf search by one wave. Returns the absolute seed, or SHGPU_NOT_FOUND when no seed in
// [start, start + max_trials) qualifies.
template <int S>
__device__ __forceinline__ uint64_t sh1_leaf_search(const uint64_t *keys, uint32_t m, uint64_t start,
                                                    uint64_t max_trials, PfLds &L) {
    if (m < 2) return start;
    const uint64_t kv = sh1_stage(keys, m, L);
    for (uint64_t woff = 0; woff < max_trials; woff += 64 * S) {
        const uint64_t off = sh1_window_staged<S>(kv, m, start, woff, L);
        if (off != SHGPU_NOT_FOUND) return off < max_trials ? start + off : SHGPU_NOT_FOUND;
    }
    return SHGPU_NOT_FOUND;
}

}  // namespace shgpu
