// One-bit retrieval on the device: a sharded ribbon (banded linear system over GF(2)) that returns, for every key
// of the set, the orientation bit of the key's leaf cell. It takes the place of the reference's
// SimpleRibbon<1, ...> construction (include/ShockHash.h:152-153,657-658; include/ShockHash2.h:134,491-497); query
// answers depend only on the bits it returns for the set's keys.
//
// Layout (shared with the host class shockhash::RibbonRetrieval, which answers the queries):
//  * S = max(1, ceil(n / 2048)) shards; key k belongs to shard fastrange64(h, S), h = remix(k + RB_SALT_SHARD).
//  * shard s holds n_s keys and m_s = rb_slots(n_s) one-bit slots (n_s + ceil(n_s / 20), at least 64, none for an
//    empty shard), at slot offset off[s] = m_0 + ... + m_{s-1} of one bit array; off has S + 1 entries.
//  * with the shard's seed byte d: g = remix(h + (d + 1) * GOLDEN), start = fastrange64(g, m_s - 63),
//    row = remix(g ^ RB_SALT_ROW) | 1; the key's bit is the parity of row AND slots [start, start + 64) of its
//    shard (bit j of row covers slot start + j).
// Construction, per shard (one wave each, independent): Gaussian elimination on the fly, the standard ribbon
// insertion with keys taken in order of their start slots: a row meets the stored row of its leading slot, is XORed
// with it and shifted to its next leading bit until it reaches a free slot; a row that vanishes with bit 1 makes the
// system inconsistent. In start order every row stays inside its own 64 slots, so the live band fits the lanes of
// one wave (k_rb_solve). Then back substitution from the last slot, free slots set to 0. A shard whose system is
// inconsistent (a few percent of shards at 1.05 slots per key) is solved again with its next seed byte by the same
// wave; repeated keys with different bits fail for every seed (SHGPU_ERR_SEARCH_LIMIT).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/shockhash_amd.h"
#include "device_common.h"
#include "ribbon.h"
#include "sort.h"

namespace {

constexpr uint64_t GOLDEN = SHGPU_RIBBON_SEED_STEP;
constexpr uint32_t RB_SEEDS = 256;  // seed bytes tried per shard

__device__ __forceinline__ uint32_t rb_slots_dev(uint32_t n) { return n == 0 ? 0u : max(RB_WIDTH, n + (n + 19) / 20); }

#define RB_GRID_STRIDE(i, n) for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// Sort key (shard << 1 | bit) and value (the key's hash h = remix(key + RB_SALT_SHARD)) of every key: sorted, the
// shard's keys lie together with their bits and hashes, read in order by the solve.
__global__ void k_rb_shard(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ bits, uint64_t n, uint64_t nshards,
                           uint64_t *skey, uint64_t *hval) {
    RB_GRID_STRIDE(i, n) {
        const uint64_t h = shgpu::remix(keys[i] + RB_SALT_SHARD);
        skey[i] = (__umul64hi(h, nshards) << 1) | (uint64_t)(bits[i] & 1u);
        hval[i] = h;
    }
}

// key_begin[s] = index of shard s's first key in the sorted order (key_begin[S] = n); sorted keys are shard << 1 | bit.
__global__ void k_rb_bounds(const uint64_t *__restrict__ sorted_key, uint64_t n, uint64_t nshards, uint32_t *key_begin) {
    RB_GRID_STRIDE(i, n + 1) {
        const uint64_t b = i < n ? sorted_key[i] >> 1 : nshards;
        const uint64_t lo = i > 0 ? (sorted_key[i - 1] >> 1) + 1 : 0;
        for (uint64_t s = lo; s <= b && s <= nshards; s++) key_begin[s] = (uint32_t)i;
    }
}

// Slot offsets: off[s] = sum of rb_slots(n_t) over t < s, by one block (PLAN-style chunked scan); also the largest
// shard's slot count (off[S + 1] holds it), and every seed byte starts at 0.
constexpr uint32_t RB_SCAN_THREADS = 1024;
__global__ void __launch_bounds__(RB_SCAN_THREADS) k_rb_offsets(const uint32_t *key_begin, uint32_t nshards, uint32_t *off,
                                                                uint8_t *seeds) {
    __shared__ uint32_t part[RB_SCAN_THREADS], mx[RB_SCAN_THREADS];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (nshards + RB_SCAN_THREADS - 1) / RB_SCAN_THREADS;
    const uint32_t c0 = min(nshards, t * chunk), c1 = min(nshards, c0 + chunk);
    uint32_t sum = 0, big = 0;
    for (uint32_t s = c0; s < c1; s++) {
        const uint32_t m = rb_slots_dev(key_begin[s + 1] - key_begin[s]);
        sum += m;
        big = max(big, m);
        seeds[s] = 0;
    }
    part[t] = sum;
    mx[t] = big;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0, m = 0;
        for (uint32_t k = 0; k < RB_SCAN_THREADS; k++) {
            const uint32_t v = part[k];
            part[k] = run;
            run += v;
            m = max(m, mx[k]);
        }
        off[nshards] = run;
        off[nshards + 1] = m;
    }
    __syncthreads();
    uint32_t run = part[t];
    for (uint32_t s = c0; s < c1; s++) {
        off[s] = run;
        run += rb_slots_dev(key_begin[s + 1] - key_begin[s]);
    }
}

// Clears the slot words and the two counters (failed shards, shard queue).
__global__ void k_rb_zero(uint64_t *words, uint64_t nwords, uint32_t *counters) {
    RB_GRID_STRIDE(i, nwords) words[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[threadIdx.x] = 0;
}

__device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t k) { return k ? (x << k) | (x >> (64 - k)) : x; }
__device__ __forceinline__ uint64_t rotr64(uint64_t x, uint32_t k) { return k ? (x >> k) | (x << (64 - k)) : x; }

// Start slot and coefficient row of a key with hash h = remix(key + RB_SALT_SHARD) under seed byte d (d1 = d + 1),
// nb = m - 63 start values.
__device__ __forceinline__ void rb_hash(uint64_t h, uint64_t d1, uint32_t nb, uint32_t &start, uint64_t &row) {
    const uint64_t g = shgpu::remix(h + d1 * GOLDEN);
    start = (uint32_t)__umul64hi(g, (uint64_t)nb);
    row = shgpu::remix(g ^ RB_SALT_ROW) | 1ull;
}

// Persistent waves (blocks of 64 threads) take shards from a queue counter, one at a time. Per seed byte d:
//  1. Counting sort of the shard's keys by start slot, stable in the shard's key order: a histogram of u16 counts in
//     LDS (32-bit atomics on the packed pairs), an exclusive scan, and a scatter whose rank among equal starts of one
//     batch of 64 keys is counted from the lower lanes. The sorted rows and (start | bit << 31) go to srow/smeta.
//  2. Insertion in start order. A row with start s reduced against rows of keys with starts <= s stays inside slots
//     [s, s + 63], so the live band is a 64-slot window held in the lanes, transposed and kept in reduced row
//     echelon form: a key's elimination is one ballot instead of a chain of dependent steps, and the band is never
//     read from memory. Slots that fall below the window are final and go to band[] packed as
//     (row >> 1) | bit << 63 (row relative to its pivot, whose bit 0 it drops; a free slot is 0).
//  3. Back substitution from the last slot over band[], 64 slots per load; the solution bits are ORed into words.
// A shard whose system is inconsistent is solved again with the next seed byte; the byte that works goes to seeds[],
// shards without one (repeated keys with different bits) are counted in *nfailed and write nothing to words.
__global__ void __launch_bounds__(64) k_rb_solve(const uint64_t *__restrict__ hs, const uint64_t *__restrict__ sk,
                                                 const uint32_t *__restrict__ key_begin,
                                                 const uint32_t *__restrict__ off, uint32_t nshards, uint8_t *seeds,
                                                 uint32_t *nfailed, uint32_t *queue, uint64_t *words, uint64_t *srow,
                                                 uint32_t *smeta, uint64_t *band) {
    extern __shared__ uint32_t hist32[];  // u16 count / cursor per start value, two per word
    uint16_t *hist = (uint16_t *)hist32;
    const uint32_t lane = shgpu::lane_id();
    for (;;) {
        const uint32_t sh = shgpu::wave_fetch_add(queue, 1u);
        if (sh >= nshards) break;
        const uint32_t k0 = key_begin[sh], k1 = key_begin[sh + 1];
        const uint32_t base_slot = off[sh], m = off[sh + 1] - base_slot;
        if (k1 == k0) continue;
        const uint32_t nb = m - RB_WIDTH + 1, nw = (nb + 1) / 2;
        uint64_t *sband = band + base_slot;
        bool ok = false, lost = false;
        uint32_t d = 0;
        for (; d < RB_SEEDS && !ok && !lost; d++) {
            const uint64_t d1 = (uint64_t)d + 1;
            // 1. counting sort by start
            for (uint32_t i = lane; i < nw; i += 64) hist32[i] = 0;
            __syncthreads();
            for (uint32_t kb = k0; kb < k1; kb += 64) {
                if (kb + lane < k1) {
                    uint32_t s;
                    uint64_t row;
                    rb_hash(hs[kb + lane], d1, nb, s, row);
                    atomicAdd(&hist32[s >> 1], 1u << ((s & 1) * 16));
                }
            }
            __syncthreads();
            {
                const uint32_t cw = (nw + 63) / 64, w0 = min(nw, lane * cw), w1 = min(nw, w0 + cw);
                uint32_t sum = 0;
                for (uint32_t w = w0; w < w1; w++) {
                    const uint32_t v = hist32[w];
                    sum += (v & 0xFFFFu) + (v >> 16);
                }
                uint32_t incl = sum;  // inclusive scan over the lanes
                for (uint32_t o = 1; o < 64; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
                    if (lane >= o) incl += y;
                }
                uint32_t run = incl - sum;
                for (uint32_t w = w0; w < w1; w++) {
                    const uint32_t v = hist32[w], lo = v & 0xFFFFu;
                    hist32[w] = run | ((run + lo) << 16);
                    run += lo + (v >> 16);
                }
            }
            __syncthreads();
            for (uint32_t kb = k0; kb < k1; kb += 64) {
                const uint32_t cnt = min(64u, k1 - kb);
                const bool act = lane < cnt;
                uint32_t s = 0xFFFFFFFFu, b = 0;
                uint64_t row = 0;
                if (act) {
                    rb_hash(hs[kb + lane], d1, nb, s, row);
                    b = (uint32_t)sk[kb + lane] & 1u;
                }
                uint32_t below = 0, above = 0;
                for (uint32_t j = 0; j < cnt; j++) {
                    const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)s, (int)j);
                    if (sj == s) {
                        below += j < lane;
                        above += j > lane;
                    }
                }
                const uint32_t cur = act ? hist[s] : 0u;
                shgpu::wave_sync();  // every lane has read its cursor before the cursors move
                if (act && above == 0) hist[s] = (uint16_t)(cur + below + 1);
                const uint32_t pos = cur + below;
                if (act) {
                    if (pos < k1 - k0) {
                        srow[k0 + pos] = row;
                        smeta[k0 + pos] = s | (b << 31);
                    } else {
                        lost = true;  // cannot happen with a consistent histogram; never write past the shard
                    }
                }
                shgpu::wave_sync();
            }
            lost = __ballot(lost) != 0;
            __syncthreads();  // sorted rows visible to every lane
            if (lost) break;
            // 2. insertion in start order over a 64-slot window kept in reduced row echelon form (every live row is 0
            //    at every other pivot). Slot p is bit p % 64 of the window's masks; rows are rotated so that bit q % 64
            //    stands for slot q, and held transposed: lane j holds column j (bit l: bit j of the row whose pivot
            //    is at window bit l). A new row then meets each pivot at most once, and all at once: its reduction
            //    is a ballot of the parities of (column AND covered pivots); it takes its lowest remaining slot as
            //    pivot and is cleared from the other rows by one XOR per lane. P: pivot slots, W: right-hand bits.
            uint64_t col = 0, P = 0, W = 0;
            uint32_t s0 = 0;
            const uint64_t lanebit = 1ull << lane;
            ok = true;
            // slots [s0, s1) leave the window: their rows are stored relative to their pivot (packed, one store per
            // lane), and cleared from the columns; slots past the window are free
            auto advance = [&](uint32_t s1) {
                const uint32_t adv = s1 - s0, nout = min(adv, 64u);
                uint64_t gone = 0, outv = 0;
                for (uint32_t k = 0; k < nout; k++) {
                    const uint32_t l = (s0 + k) & 63u;
                    const uint64_t row = __ballot((col >> l) & 1ull);
                    const uint64_t v = (rotr64(row, l) >> 1) | (((W >> l) & 1ull) << 63);
                    if (lane == l) outv = v;
                    gone |= 1ull << l;
                }
                if (gone & lanebit) sband[s0 + ((lane - s0) & 63u)] = outv;
                col &= ~gone;
                P &= ~gone;
                W &= ~gone;
                for (uint32_t qb = (s0 + 64) & ~63u; qb < s1; qb += 64) {
                    const uint32_t q = qb + lane;  // lane q % 64 writes slot q, as it does in the window
                    if (q >= s0 + 64 && q < s1) sband[q] = 0;
                }
                s0 = s1;
            };
            for (uint32_t kb = k0; kb < k1 && ok; kb += 64) {
                const uint32_t cnt = min(64u, k1 - kb);
                const uint32_t emeta = lane < cnt ? smeta[kb + lane] : 0u;
                const uint64_t erot = lane < cnt ? rotl64(srow[kb + lane], emeta & 63u) : 0;  // row in window bits
                for (uint32_t j = 0; j < cnt; j++) {
                    const uint32_t meta = (uint32_t)__builtin_amdgcn_readlane((int)emeta, (int)j);
                    const uint32_t s1 = meta & 0x7FFFFFFFu;
                    uint32_t b = meta >> 31;
                    const uint64_t r0 = shgpu::readlane64(erot, j);
                    if (s1 != s0) advance(s1);
                    const uint64_t x = r0 & P;  // pivots the row covers
                    const uint64_t r = r0 ^ __ballot(__builtin_popcountll(col & x) & 1);
                    b ^= (uint32_t)__builtin_popcountll(x & W) & 1u;
                    if (r == 0) {
                        if (b) {
                            ok = false;  // a redundant row is fine, a contradicting one is not
                            break;
                        }
                        continue;
                    }
                    const uint32_t pl = (s1 + (uint32_t)__builtin_ctzll(rotr64(r, s1 & 63u))) & 63u;  // new pivot
                    const uint64_t has = shgpu::readlane64(col, pl);  // rows with a 1 at the new pivot
                    if (r & lanebit) col ^= has | (1ull << pl);
                    if (b) W ^= has;
                    P |= 1ull << pl;
                    W |= (uint64_t)b << pl;
                }
            }
            if (!ok) continue;  // lost ends the seed loop
            advance(m);
            __syncthreads();  // the band visible to every lane
        }
        if (!ok) {
            if (lane == 0) atomicAdd(nfailed, 1u);
            continue;
        }
        if (lane == 0) seeds[sh] = (uint8_t)(d - 1);
        // 3. back substitution: acc holds z[s + 1 .. s + 64] in bits 0..63 before slot s; a slot's bit is
        //    parity(v & (z[s + 1 .. s + 63] | bit 63)), v = (row >> 1) | rhs << 63
        uint64_t acc = 0;
        for (int32_t top = (int32_t)m; top > 0; top -= 64) {
            const int32_t blk = top - 64 > 0 ? top - 64 : 0;
            const uint64_t vl = blk + (int32_t)lane < top ? sband[blk + lane] : 0;
            for (int32_t s = top - 1; s >= blk; s--) {
                const uint64_t v = shgpu::readlane64(vl, (uint32_t)(s - blk));
                const uint64_t z = (uint64_t)__builtin_popcountll(v & (acc | (1ull << 63))) & 1ull;
                acc = (acc << 1) | z;
            }
            // acc bits 0 .. top - blk - 1 are the slots blk .. top - 1
            const uint32_t nbits = (uint32_t)(top - blk);
            const uint64_t zb = nbits == 64 ? acc : acc & ((1ull << nbits) - 1);
            const uint64_t g = (uint64_t)base_slot + (uint64_t)blk, w = g >> 6, sft = g & 63;
            if (lane == 0 && zb) {
                atomicOr((unsigned long long *)&words[w], (unsigned long long)(zb << sft));
                if (sft && (zb >> (64 - sft))) atomicOr((unsigned long long *)&words[w + 1], (unsigned long long)(zb >> (64 - sft)));
            }
        }
        __syncthreads();
    }
}

struct RbBuf {
    void *p = nullptr;
    ~RbBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

uint64_t shgpu_ribbon_shards(uint64_t n) { return n == 0 ? 1 : (n + RB_SHARD_KEYS - 1) / RB_SHARD_KEYS; }

uint64_t shgpu_ribbon_max_words(uint64_t n) {
    const uint64_t s = shgpu_ribbon_shards(n);
    return (n + (n + 19 * s) / 20 + s * RB_WIDTH) / 64 + 2;
}

int shgpu_ribbon_build_impl(hipStream_t st, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint32_t *h_offsets,
                            uint8_t *h_seeds, uint64_t *h_words, uint64_t words_capacity, uint64_t *words_used,
                            std::string &err, void **cache, size_t *cache_bytes) {
    const uint64_t S = shgpu_ribbon_shards(n);
    if (n >= (1ull << 32)) { err = "ribbon retrieval: more than 2^32 - 1 keys"; return SHGPU_ERR_UNSUPPORTED; }
    const uint64_t maxw = shgpu_ribbon_max_words(n);
    if (words_capacity < maxw) { err = "ribbon retrieval: output capacity below " + std::to_string(maxw) + " words"; return SHGPU_ERR_ARG; }
    // one allocation: sort keys (shard << 1 | bit), hashes, both sorted (n u64 each), sort scratch, key_begin and
    // offsets (S + 2 u32 each), seeds (S), counters, words, band (u64 per slot)
    size_t sort_bytes = 0;
    if (shgpu_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, (int)std::max<uint64_t>(n, 1), 64, st) !=
        hipSuccess) {
        err = "ribbon retrieval: sort sizing failed";
        return SHGPU_ERR_HIP;
    }
    const size_t parts[] = {(n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, sort_bytes + 16, (S + 2) * 4, (S + 2) * 4,
                            S + 8, 64, maxw * 8, maxw * 64 * 8};
    size_t bytes = 0;
    for (size_t b : parts) bytes += (b + 255) / 256 * 256;
    RbBuf own;
    uint8_t *base;
    if (cache) {
        if (*cache_bytes < bytes) {
            if (*cache) (void)hipFree(*cache);
            *cache = nullptr;
            *cache_bytes = 0;
            if (!shgpu::device_malloc(cache, bytes)) {
                err = "ribbon retrieval: hipMalloc of " + std::to_string(bytes) + " bytes failed";
                return SHGPU_ERR_NOMEM;
            }
            *cache_bytes = bytes;
        }
        base = (uint8_t *)*cache;
    } else {
        if (!shgpu::device_malloc(&own.p, bytes)) { err = "ribbon retrieval: hipMalloc of " + std::to_string(bytes) + " bytes failed"; return SHGPU_ERR_NOMEM; }
        base = (uint8_t *)own.p;
    }
    void *ptr[11];
    size_t at = 0;
    for (int i = 0; i < 11; i++) {
        ptr[i] = base + at;
        at += (parts[i] + 255) / 256 * 256;
    }
    uint64_t *skey = (uint64_t *)ptr[0], *hval = (uint64_t *)ptr[1], *skey_sorted = (uint64_t *)ptr[2],
             *h_sorted = (uint64_t *)ptr[3];
    void *sort_tmp = ptr[4];
    uint32_t *key_begin = (uint32_t *)ptr[5], *off = (uint32_t *)ptr[6];
    uint8_t *seeds = (uint8_t *)ptr[7];
    uint32_t *nfailed = (uint32_t *)ptr[8];
    uint64_t *words = (uint64_t *)ptr[9];
    uint64_t *band = (uint64_t *)ptr[10];  // one packed row per slot (at most 64 * maxw slots)
    // after the shard sort, the unsorted arrays are dead: the solve's start-sorted rows and starts take their place
    uint64_t *srow = skey;
    uint32_t *smeta = (uint32_t *)hval;
#define RB_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) { err = std::string("ribbon retrieval: ") + hipGetErrorString(e_); return SHGPU_ERR_HIP; } \
    } while (0)
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256 + 1, 4096);
    int bits = 1;  // sort key bits: the shard id's and the retrieved bit
    while (bits < 63 && (1ull << bits) < S) bits++;
    bits++;
    if (n) {
        k_rb_shard<<<grid, 256, 0, st>>>(d_keys, d_bits, n, S, skey, hval);
        RB_HIP(hipGetLastError());
        size_t sb = sort_bytes;
        RB_HIP(shgpu_sort_pairs_u64(sort_tmp, sb, skey, skey_sorted, hval, h_sorted, (int)n, bits, st));
    }
    k_rb_bounds<<<grid, 256, 0, st>>>(skey_sorted, n, S, key_begin);
    RB_HIP(hipGetLastError());
    k_rb_offsets<<<1, RB_SCAN_THREADS, 0, st>>>(key_begin, (uint32_t)S, off, seeds);
    RB_HIP(hipGetLastError());
    k_rb_zero<<<(uint32_t)std::min<uint64_t>((maxw + 255) / 256, 4096), 256, 0, st>>>(words, maxw, nfailed);
    RB_HIP(hipGetLastError());
    RB_HIP(hipMemcpyAsync(h_offsets, off, (S + 1) * 4, hipMemcpyDeviceToHost, st));
    uint32_t max_slots = 0;
    RB_HIP(hipMemcpyAsync(&max_slots, off + S + 1, 4, hipMemcpyDeviceToHost, st));
    RB_HIP(hipStreamSynchronize(st));
    if (max_slots > RB_MAX_SLOTS) {
        err = "ribbon retrieval: a shard holds " + std::to_string(max_slots) + " slots, more than its u16 start counts allow (repeated keys?)";
        return SHGPU_ERR_SEARCH_LIMIT;
    }
    const size_t lds = max_slots >= RB_WIDTH ? (size_t)(max_slots - RB_WIDTH + 2) / 2 * 4 + 16 : 16;
    const uint32_t total_slots = h_offsets[S];
    const uint64_t nwords = (uint64_t)(total_slots + 63) / 64 + 1;  // one word of padding for the 64-slot windows
    // Persistent waves, RB_WAVES_PER_CU per CU, take shards from the queue: every CU gets the same load even when the
    // dispatcher would pack one-shard blocks unevenly, and a shard solved again runs while other shards go on.
    int dev = 0, cus = 0;
    RB_HIP(hipGetDevice(&dev));
    RB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(S, (uint64_t)std::max(cus, 1) * RB_WAVES_PER_CU));
    k_rb_solve<<<blocks, 64, lds, st>>>(h_sorted, skey_sorted, key_begin, off, (uint32_t)S, seeds, nfailed, nfailed + 1, words,
                                        srow, smeta, band);
    RB_HIP(hipGetLastError());
    uint32_t nf = 0;
    RB_HIP(hipMemcpyAsync(&nf, nfailed, 4, hipMemcpyDeviceToHost, st));
    RB_HIP(hipMemcpyAsync(h_seeds, seeds, S, hipMemcpyDeviceToHost, st));
    RB_HIP(hipMemcpyAsync(h_words, words, nwords * 8, hipMemcpyDeviceToHost, st));
    RB_HIP(hipStreamSynchronize(st));
#undef RB_HIP
    if (nf) {
        err = "ribbon retrieval: " + std::to_string(nf) + " shards inconsistent for every seed byte (repeated keys with different bits?)";
        return SHGPU_ERR_SEARCH_LIMIT;
    }
    *words_used = nwords;
    return SHGPU_OK;
}
