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
// insertion (rows kept in the shard's band, in LDS): a row meets the stored row of its leading slot, is XORed
// with it and shifted to its next leading bit until it reaches a free slot; a row that vanishes with bit 1 makes
// the system inconsistent. Then back substitution from the last slot, free slots set to 0. A shard whose system
// is inconsistent (a few percent of shards at 1.05 slots per key) is solved again with its next seed byte by the
// same wave; repeated keys with different bits fail for every seed (SHGPU_ERR_SEARCH_LIMIT).
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

// Shard of every key (the sort key) and its index (the value).
__global__ void k_rb_shard(const uint64_t *__restrict__ keys, uint64_t n, uint64_t nshards, uint64_t *shard, uint64_t *idx) {
    RB_GRID_STRIDE(i, n) {
        shard[i] = __umul64hi(shgpu::remix(keys[i] + RB_SALT_SHARD), nshards);
        idx[i] = i;
    }
}

// key_begin[s] = index of shard s's first key in the sorted order (key_begin[S] = n).
__global__ void k_rb_bounds(const uint64_t *__restrict__ sorted_shard, uint64_t n, uint64_t nshards, uint32_t *key_begin) {
    RB_GRID_STRIDE(i, n + 1) {
        const uint64_t b = i < n ? sorted_shard[i] : nshards;
        const uint64_t lo = i > 0 ? sorted_shard[i - 1] + 1 : 0;
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

__global__ void k_rb_zero(uint64_t *words, uint64_t nwords, uint32_t *counter) {
    RB_GRID_STRIDE(i, nwords) words[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
}

// One wave per shard (blocks of 64 threads, grid-stride over shards), rows and bits of the shard's band in dynamic
// LDS (9 bytes per slot). A shard whose system is inconsistent with its seed byte is solved again with the next one,
// in the same wave; the seed byte that works goes to seeds[]. Shards without one (repeated keys with different bits)
// are counted in *nfailed and write nothing. Solved shards OR their slot bits into words (zeroed beforehand).
// The elimination is sequential per shard and runs on wave-uniform values: the lanes first hash 64 keys at once
// (start slot, row, bit), then the wave inserts them one by one (readlane), walking a 64-slot window of the band
// held in the lanes.
// Back substitution reads 64 rows and bits per LDS access into the lanes and walks them by readlane.
__global__ void __launch_bounds__(64) k_rb_solve(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ bits,
                                                 const uint64_t *__restrict__ order, const uint32_t *__restrict__ key_begin,
                                                 const uint32_t *__restrict__ off, uint32_t nshards, uint8_t *seeds,
                                                 uint32_t *nfailed, uint64_t *words) {
    extern __shared__ uint64_t rows[];
    const uint32_t lane = shgpu::lane_id();
    for (uint32_t sh = blockIdx.x; sh < nshards; sh += gridDim.x) {
        const uint32_t k0 = key_begin[sh], k1 = key_begin[sh + 1];
        const uint32_t base_slot = off[sh], m = off[sh + 1] - base_slot;
        if (k1 == k0) continue;
        uint8_t *bb = (uint8_t *)(rows + m);
        bool ok = false;
        uint32_t d = 0;
        for (; d < RB_SEEDS && !ok; d++) {
            const uint64_t d1 = (uint64_t)d + 1;
            for (uint32_t i = lane; i < m; i += 64) {
                rows[i] = 0;
                bb[i] = 0;
            }
            shgpu::wave_sync();
            ok = true;
            for (uint32_t kb = k0; kb < k1 && ok; kb += 64) {
                // lane l hashes key kb + l: start slot, row, bit
                uint32_t start = 0, bit = 0;
                uint64_t row = 0;
                if (kb + lane < k1) {
                    const uint64_t i = order[kb + lane];
                    const uint64_t g = shgpu::remix(shgpu::remix(keys[i] + RB_SALT_SHARD) + d1 * GOLDEN);
                    start = (uint32_t)__umul64hi(g, (uint64_t)(m - RB_WIDTH + 1));
                    row = shgpu::remix(g ^ RB_SALT_ROW) | 1ull;
                    bit = bits[i] & 1u;
                }
                const uint32_t cnt = min(64u, k1 - kb);
                // The band window [base, base + 64) sits in the lanes (lane l: row and bit of slot base + l), so an
                // elimination step is a readlane, not a dependent LDS read. The next key's window is loaded while
                // this key is inserted and patched with this key's stored row if it falls inside.
                uint32_t nstart = (uint32_t)__builtin_amdgcn_readlane((int)start, 0);
                uint64_t nwr = rows[nstart + lane];  // start + 63 < m
                uint32_t nwb = bb[nstart + lane];
                for (uint32_t j = 0; j < cnt; j++) {
                    uint32_t sj = nstart, base = nstart;
                    uint64_t wr = nwr;
                    uint32_t wb = nwb;
                    if (j + 1 < cnt) {
                        nstart = (uint32_t)__builtin_amdgcn_readlane((int)start, (int)j + 1);
                        nwr = rows[nstart + lane];
                        nwb = bb[nstart + lane];
                    }
                    uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)bit, (int)j);
                    uint64_t r = shgpu::readlane64(row, j);
                    // eliminate against the stored rows until the row reaches a free slot or vanishes
                    bool stored = false;
                    for (;;) {
                        uint32_t o = sj - base;
                        if (o >= 64) {  // the row left the window: load the one starting at its leading slot
                            base = sj;
                            o = 0;
                            wr = sj + lane < m ? rows[sj + lane] : 0;
                            wb = sj + lane < m ? bb[sj + lane] : 0u;
                        }
                        const uint64_t c = shgpu::readlane64(wr, o);
                        if (c == 0) {
                            stored = true;
                            break;
                        }
                        r ^= c;
                        b ^= (uint32_t)__builtin_amdgcn_readlane((int)wb, (int)o);
                        if (r == 0) break;
                        const uint32_t t = (uint32_t)__builtin_ctzll(r);
                        r >>= t;
                        sj += t;
                    }
                    if (stored) {
                        if (lane == 0) {
                            rows[sj] = r;
                            bb[sj] = (uint8_t)b;
                        }
                        if (j + 1 < cnt && lane == sj - nstart) {  // the next key's window holds slot sj
                            nwr = r;
                            nwb = b;
                        }
                    } else if (b) {
                        ok = false;  // a redundant row is fine, a contradicting one is not
                        break;
                    }
                }
                shgpu::wave_sync();
            }
        }
        if (!ok) {
            if (lane == 0) atomicAdd(nfailed, 1u);
            continue;
        }
        if (lane == 0) seeds[sh] = (uint8_t)(d - 1);
        // Back substitution from the last slot: zw holds z[s + 1 .. s + 64] in bits 0..63. Rows and bits of slots
        // [blk, blk + 64) are read into the lanes (lane l: slot blk + l) and walked downwards by readlane.
        uint64_t zw = 0, acc = 0;
        uint64_t word = (base_slot + m - 1) >> 6;
        for (int32_t top = (int32_t)m; top > 0; top -= 64) {
            const int32_t blk = top - 64 > 0 ? top - 64 : 0;
            const int32_t mine = blk + (int32_t)lane;
            const uint64_t rl = mine < top ? rows[mine] : 0;
            const uint32_t bl = mine < top ? bb[mine] : 0u;
            for (int32_t s = top - 1; s >= blk; s--) {
                const uint64_t c = shgpu::readlane64(rl, (uint32_t)(s - blk));
                uint64_t z = 0;
                if (c) {
                    const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane((int)bl, s - blk);
                    z = (cb ^ (uint64_t)__builtin_popcountll((c >> 1) & zw)) & 1ull;
                }
                zw = (zw << 1) | z;
                const uint64_t p = base_slot + (uint64_t)s;
                if ((p >> 6) != word) {
                    if (acc && lane == 0) atomicOr((unsigned long long *)&words[word], (unsigned long long)acc);
                    acc = 0;
                    word = p >> 6;
                }
                acc |= z << (p & 63);
            }
        }
        if (acc && lane == 0) atomicOr((unsigned long long *)&words[word], (unsigned long long)acc);
        shgpu::wave_sync();
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
    // one allocation: shard ids, indices, sorted shard ids, sorted indices (n u64 each), sort scratch, key_begin and
    // offsets (S + 2 u32 each), seeds (S), counters, words
    size_t sort_bytes = 0;
    if (shgpu_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, (int)std::max<uint64_t>(n, 1), 64, st) !=
        hipSuccess) {
        err = "ribbon retrieval: sort sizing failed";
        return SHGPU_ERR_HIP;
    }
    const size_t parts[] = {(n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, sort_bytes + 16, (S + 2) * 4, (S + 2) * 4,
                            S + 8, 64, maxw * 8};
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
    void *ptr[10];
    size_t at = 0;
    for (int i = 0; i < 10; i++) {
        ptr[i] = base + at;
        at += (parts[i] + 255) / 256 * 256;
    }
    uint64_t *shard = (uint64_t *)ptr[0], *idx = (uint64_t *)ptr[1], *shard_sorted = (uint64_t *)ptr[2],
             *order = (uint64_t *)ptr[3];
    void *sort_tmp = ptr[4];
    uint32_t *key_begin = (uint32_t *)ptr[5], *off = (uint32_t *)ptr[6];
    uint8_t *seeds = (uint8_t *)ptr[7];
    uint32_t *nfailed = (uint32_t *)ptr[8];
    uint64_t *words = (uint64_t *)ptr[9];
#define RB_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) { err = std::string("ribbon retrieval: ") + hipGetErrorString(e_); return SHGPU_ERR_HIP; } \
    } while (0)
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256 + 1, 4096);
    int bits = 1;
    while (bits < 64 && (1ull << bits) < S) bits++;
    if (n) {
        k_rb_shard<<<grid, 256, 0, st>>>(d_keys, n, S, shard, idx);
        RB_HIP(hipGetLastError());
        size_t sb = sort_bytes;
        RB_HIP(shgpu_sort_pairs_u64(sort_tmp, sb, shard, shard_sorted, idx, order, (int)n, bits, st));
    }
    k_rb_bounds<<<grid, 256, 0, st>>>(shard_sorted, n, S, key_begin);
    RB_HIP(hipGetLastError());
    k_rb_offsets<<<1, RB_SCAN_THREADS, 0, st>>>(key_begin, (uint32_t)S, off, seeds);
    RB_HIP(hipGetLastError());
    k_rb_zero<<<(uint32_t)std::min<uint64_t>((maxw + 255) / 256, 4096), 256, 0, st>>>(words, maxw, nfailed);
    RB_HIP(hipGetLastError());
    RB_HIP(hipMemcpyAsync(h_offsets, off, (S + 1) * 4, hipMemcpyDeviceToHost, st));
    uint32_t max_slots = 0;
    RB_HIP(hipMemcpyAsync(&max_slots, off + S + 1, 4, hipMemcpyDeviceToHost, st));
    RB_HIP(hipStreamSynchronize(st));
    const size_t lds = (size_t)max_slots * 9 + 16;
    if (lds > RB_MAX_LDS) {
        err = "ribbon retrieval: a shard holds " + std::to_string(max_slots) + " slots, more than fit in LDS (repeated keys?)";
        return SHGPU_ERR_SEARCH_LIMIT;
    }
    const uint32_t total_slots = h_offsets[S];
    const uint64_t nwords = (uint64_t)(total_slots + 63) / 64 + 1;  // one word of padding for the 64-slot windows
    // every shard in one launch; LDS bounds the waves per CU, so one block per shard is resident or queued
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(S, 1u << 20));
    k_rb_solve<<<blocks, 64, lds, st>>>(d_keys, d_bits, order, key_begin, off, (uint32_t)S, seeds, nfailed, words);
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
