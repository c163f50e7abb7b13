// RecSplit split-seed search and stable partition for one wavefront (gfx950 wave64).
//
// Replaces the split loops of ShockHash::recSplit (include/ShockHash.h:451-548), ShockHash2::recSplit
// (include/ShockHash2.h:290-387) and SIMDShockHash::higherLevel/aggrLevel
// (include/SIMDShockHash.hpp:201-365). A node of m keys with part size `unit` searches the smallest
// seed x >= start such that, with h(key) = hash(key + x) mod-mapped to [0, m), every part but the
// last receives exactly `unit` keys (kind 0: exactly `unit` keys with h < unit).
//
// Hash flavours:
//   sux  : h = ((remix(key + x) & (2^48 - 1)) * m) >> 48             (sux::remap16 o remix)
//   SIMD : h = ((fmix32(lo32(key + x) ^ hi32(key + x)) & 0xFFFF) * m) >> 16
// Part counting is done with cumulative thresholds: part p < fanout-1 holds exactly `unit` keys for
// all p  <=>  #{h < (p+1)*unit} == (p+1)*unit for all p. Each test h < t is evaluated without the
// multiply as v < ceil(t * 2^w / m), where v is the w-bit value before scaling (w = 48 or 16).
// SIMDShockHash counts with packed 8-bit lanes mod 2^32; for every fanout/size the splitting
// strategies produce (fanout <= 4, part sizes < 256 in the 4-way case) the packed test accepts
// exactly the seeds whose plain counts match, so plain counts give the same minimal seed.
//
// Exact power-of-two fanouts. When m = F * unit with F = 2 or 4 (the full lower-aggregation nodes,
// fanout-2 nodes that halve exactly, upper aggregation of small leaves), the part of a key is
// floor(v * F / 2^w): the top one or two bits of the w-bit value. For the sux flavour those are bits
// 46..47 of remix(z), which the final xor-shift (z ^ z >> 31) leaves untouched, so they are bits 46..47
// of the last 64-bit product; only the low 16 bits of that product's high word are needed, which
// 24-bit multiplies give exactly. Counting then needs no compares: F = 2 counts the top bit; F = 4
// adds 1 << (8 * (3 - part)) into packed 8-bit fields (one v_alignbyte_b32 of a constant, which reads only
// the low two bits of its shift operand), and the packed sum equals unit * 0x01010101 (mod 2^32)
// exactly when every part holds unit keys, because each part's deviation from unit lies in
// [-unit, 3 * unit] and 3 * unit < 256 (the low field forces d0 = 0, then d1 = 0, then d2 = 0, and the
// counts sum to m).
//
// Mapping: lanes over seeds (S seeds per lane, window 64*S), keys staged in LDS and read back as
// wave-wide broadcasts; counters live in registers.
#pragma once
#include "device_common.h"

namespace shgpu {

// Counting modes of a split window.
constexpr int SPLIT_THRESHOLDS = 0;  // cumulative threshold compares (any shape)
constexpr int SPLIT_POW2_F2 = 1;     // m = 2 * unit: count the top bit
constexpr int SPLIT_POW2_F4 = 2;     // m = 4 * unit, unit <= 85: packed 8-bit counts of the top two bits

struct SplitShapeDev {
    uint32_t m, unit, fanout, nt;  // nt = number of thresholds (1 for kind 0, fanout-1 otherwise)
    uint32_t t[3];                 // target cumulative counts
    uint64_t thr[3];               // thresholds in the pre-scaling domain
    int simd;
    int mode;                      // SPLIT_THRESHOLDS / SPLIT_POW2_F2 / SPLIT_POW2_F4
};

__device__ __forceinline__ SplitShapeDev make_split_shape(uint32_t m, uint32_t unit, uint32_t kind, uint32_t hash) {
    SplitShapeDev s;
    s.m = m;
    s.unit = unit;
    s.simd = hash == SHGPU_SPLIT_SIMD;
    s.fanout = kind == SHGPU_NODE_SPLIT2 ? 2u : (m + unit - 1) / unit;
    s.nt = s.fanout - 1;
    const uint32_t w = s.simd ? 16 : 48;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const uint32_t tp = (uint32_t)(p + 1) * unit;
        s.t[p] = tp;
        // ceil(tp * 2^w / m); tp < m <= 3000 so tp * 2^48 < 2^60.
        s.thr[p] = (((uint64_t)tp << w) + m - 1) / m;
    }
    s.mode = SPLIT_THRESHOLDS;
    if (s.fanout == 2 && m == 2 * unit) s.mode = SPLIT_POW2_F2;
    if (s.fanout == 4 && m == 4 * unit && 3 * unit < 256) s.mode = SPLIT_POW2_F4;
    return s;
}

// Mix13 of z up to the low 16 bits of the high word of the last product (bits 46..47 of remix(z) are
// bits 14..15 of the result).
__device__ __forceinline__ uint32_t remix_top16(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    const uint32_t lo = (uint32_t)z, hi = (uint32_t)(z >> 32);
    // high word of lo64(z * C) = mulhi(lo, C_lo) + lo * C_hi + hi * C_lo (mod 2^32); mod 2^16 the
    // cross terms need only the low 16 bits of their operands, so two 24-bit multiply-adds give them
    return mad24(hi, 0x133111ebu, mad24(lo, 0x94d049bbu, __umulhi(lo, 0x133111ebu)));
}

// x >> 14 as one instruction. Written out because for the SIMD flavour the compiler otherwise distributes
// the shift over fmix32's last xor-shift ((z >> 14) ^ (z >> 30): three instructions instead of two).
__device__ __forceinline__ uint32_t shr14(uint32_t x) {
    uint32_t d;
    asm("v_lshrrev_b32 %0, 14, %1" : "=v"(d) : "v"(x));
    return d;
}

// Value before the final scaling: 48-bit (sux) or 16-bit (SIMD) fraction of the hash.
__device__ __forceinline__ uint64_t split_value(uint64_t z, int simd) {
    if (simd) return remix32((uint32_t)(z >> 32) ^ (uint32_t)z) & 0xFFFFu;
    return remix(z) & 0xFFFFFFFFFFFFull;
}

template <int S, int NT, int SIMD, int MODE>
__device__ __forceinline__ void split_eval(uint64_t key, const uint64_t (&seed)[S], const SplitShapeDev &sh,
                                           uint32_t (&c)[S][NT]) {
#pragma unroll
    for (int s = 0; s < S; s++) {
        const uint64_t z = key + seed[s];
        if (MODE != SPLIT_THRESHOLDS) {
            // the top bits of the w-bit value sit at bits 14..15 of `top`
            const uint32_t top = SIMD ? remix32((uint32_t)(z >> 32) ^ (uint32_t)z) : remix_top16(z);
            if (MODE == SPLIT_POW2_F2) c[s][0] += (top >> 15) & 1u;
            else c[s][0] += __builtin_amdgcn_alignbyte(0u, 0x01000000u, shr14(top));
        } else if (SIMD) {
            const uint32_t v = remix32((uint32_t)(z >> 32) ^ (uint32_t)z) & 0xFFFFu;
#pragma unroll
            for (int p = 0; p < NT; p++) c[s][p] += v < (uint32_t)sh.thr[p] ? 1u : 0u;
        } else {
            const uint64_t v = remix(z) & 0xFFFFFFFFFFFFull;
#pragma unroll
            for (int p = 0; p < NT; p++) c[s][p] += v < sh.thr[p] ? 1u : 0u;
        }
    }
}

// Stages the keys of a node of m <= STAGE_KEYS keys in `stage` for split windows called with staged = true.
__device__ __forceinline__ void split_stage(const uint64_t *keys, uint32_t m, uint64_t *stage) {
    wave_sync();
    for (uint32_t i = lane_id(); i < m; i += 64) stage[i] = keys[i];
    wave_sync();
}

// One window: seeds start + woff + [0, 64*S). Returns the offset (seed - start) of the first seed whose
// parts have the target sizes, or SHGPU_NOT_FOUND. Seeds are tried in (slot, lane) order. The keys are
// staged STAGE_KEYS at a time in the wave's LDS area `stage` and read back as broadcasts, two per LDS
// read; with staged = true (m <= STAGE_KEYS) the caller has staged all of them already.
template <int S, int NT, int SIMD, int MODE>
__device__ __forceinline__ uint64_t split_window_t(const uint64_t *keys, const SplitShapeDev &sh, uint64_t start,
                                                   uint64_t woff, uint64_t *stage, bool staged) {
    const uint32_t lane = lane_id();
    const uint32_t m = sh.m;
    uint64_t seed[S];
    uint32_t c[S][NT];
#pragma unroll
    for (int s = 0; s < S; s++) {
        seed[s] = start + woff + (uint64_t)(s * 64 + lane);
#pragma unroll
        for (int p = 0; p < NT; p++) c[s][p] = 0;
    }
    const ulonglong2 *st2 = (const ulonglong2 *)stage;
    for (uint32_t base = 0; base < m; base += STAGE_KEYS) {
        const uint32_t cnt = m - base < STAGE_KEYS ? m - base : STAGE_KEYS;
        if (!staged) {
            wave_sync();  // the previous block's reads are done
            for (uint32_t i = lane; i < cnt; i += 64) stage[i] = keys[base + i];
            wave_sync();
        }
        uint32_t k = 0;
        for (; k + 8 <= cnt; k += 8) {
#pragma unroll
            for (uint32_t q = 0; q < 4; q++) {
                const ulonglong2 kk = st2[k / 2 + q];
                split_eval<S, NT, SIMD, MODE>(kk.x, seed, sh, c);
                split_eval<S, NT, SIMD, MODE>(kk.y, seed, sh, c);
            }
        }
        for (; k < cnt; k++) split_eval<S, NT, SIMD, MODE>(stage[k], seed, sh, c);
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        bool ok = true;
        if (MODE == SPLIT_POW2_F2) ok = c[s][0] == sh.unit;  // keys in the upper half
        else if (MODE == SPLIT_POW2_F4) ok = c[s][0] == sh.unit * 0x01010101u;
        else
#pragma unroll
            for (int p = 0; p < NT; p++) ok = ok && c[s][p] == sh.t[p];
        const uint64_t b = ballot(ok);
        if (b) return woff + (uint64_t)(s * 64 + __builtin_ctzll(b));
    }
    return SHGPU_NOT_FOUND;
}

// HASH: SHGPU_SPLIT_SUX or SHGPU_SPLIT_SIMD when the flavour is known at compile time (only that
// flavour's loops are generated), -1 to follow sh.simd.
template <int S, int HASH = -1>
__device__ __forceinline__ uint64_t split_window(const uint64_t *keys, const SplitShapeDev &sh, uint64_t start,
                                                 uint64_t woff, uint64_t *stage, bool staged = false) {
    if (HASH == SHGPU_SPLIT_SIMD || (HASH < 0 && sh.simd)) {
        if (sh.mode == SPLIT_POW2_F2) return split_window_t<S, 1, 1, SPLIT_POW2_F2>(keys, sh, start, woff, stage, staged);
        if (sh.mode == SPLIT_POW2_F4) return split_window_t<S, 1, 1, SPLIT_POW2_F4>(keys, sh, start, woff, stage, staged);
        if (sh.nt == 1) return split_window_t<S, 1, 1, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
        if (sh.nt == 2) return split_window_t<S, 2, 1, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
        return split_window_t<S, 3, 1, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
    }
    if (sh.mode == SPLIT_POW2_F2) return split_window_t<S, 1, 0, SPLIT_POW2_F2>(keys, sh, start, woff, stage, staged);
    if (sh.mode == SPLIT_POW2_F4) return split_window_t<S, 1, 0, SPLIT_POW2_F4>(keys, sh, start, woff, stage, staged);
    if (sh.nt == 1) return split_window_t<S, 1, 0, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
    if (sh.nt == 2) return split_window_t<S, 2, 0, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
    return split_window_t<S, 3, 0, SPLIT_THRESHOLDS>(keys, sh, start, woff, stage, staged);
}

// Sequential search by one wave: the absolute seed, or SHGPU_NOT_FOUND below start + max_trials.
template <int S>
__device__ __forceinline__ uint64_t split_search(const uint64_t *keys, const SplitShapeDev &sh, uint64_t start,
                                                 uint64_t max_trials, uint64_t *stage) {
    const bool staged = sh.m <= STAGE_KEYS;
    if (staged) split_stage(keys, sh.m, stage);
    for (uint64_t woff = 0; woff < max_trials; woff += 64 * S) {
        const uint64_t off = split_window<S>(keys, sh, start, woff, stage, staged);
        if (off != SHGPU_NOT_FOUND) return off < max_trials ? start + off : SHGPU_NOT_FOUND;
    }
    return SHGPU_NOT_FOUND;
}

// Stable counting-sort partition of src[0..m) by part for seed x into dst (ShockHash.h:464-469,495-499,528-532;
// SIMDShockHash.hpp:234-252,337-354): part p starts at p*unit.
__device__ __forceinline__ void split_partition_to(const uint64_t *src, uint64_t *dst, const SplitShapeDev &sh, uint64_t x) {
    const uint32_t lane = lane_id();
    const uint32_t m = sh.m;
    uint32_t base[4] = {0, sh.unit, 2 * sh.unit, 3 * sh.unit};
    for (uint32_t chunk = 0; chunk < m; chunk += 64) {
        const bool act = chunk + lane < m;
        const uint64_t key = act ? src[chunk + lane] : 0;
        const uint64_t v = split_value(key + x, sh.simd);
        uint32_t part = 0;
#pragma unroll
        for (uint32_t p = 0; p < 3; p++) part += (p < sh.nt && v >= sh.thr[p]) ? 1u : 0u;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t p = 0; p < 4; p++) {
            const uint64_t b = ballot(act && part == p);
            if (act && part == p) pos = base[p] + lanes_below(b);
            base[p] += (uint32_t)__builtin_popcountll(b);
        }
        if (act) dst[pos] = key;
    }
}

// The same partition in place: into tmp, then copied back.
__device__ __forceinline__ void split_partition(uint64_t *keys, uint64_t *tmp, const SplitShapeDev &sh, uint64_t x) {
    split_partition_to(keys, tmp, sh, x);
    __threadfence_block();
    for (uint32_t i = lane_id(); i < sh.m; i += 64) keys[i] = tmp[i];
}

}  // namespace shgpu
