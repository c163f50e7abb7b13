// ShockHash2Flat's bucketing on the device: the threshold layers with bumping and the placement of every key into its
// bucket, which the reference runs key by key on the host (include/ShockHash2Flat.h:70-157,204-233,280-302).
//
// A layer of B buckets: every key gets bucket fastrange32(lo32(h), B) and threshold hi32(h) (h rehashed by remix
// after the first layer); one radix sort on (bucket << 32 | threshold) gives the reference's order (sort_keyInfo,
// src/Sorter.cpp:10-12, sorts on the same key; keys equal in it have equal thresholds, so they are kept or bumped
// together whatever their order); one thread per bucket chooses its threshold index (flushBucket,
// ShockHash2Flat.h:204-233) and how many keys it keeps, always a prefix of its sorted
// keys; the bumped suffixes, concatenated in bucket order, are the next layer's keys in the reference's order.
// Placement evaluates every key against the stored thresholds (evaluateKPerfect), takes the caller's buckets for
// the keys every layer bumped, and one more radix sort by bucket lays each bucket's keys out in input order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "../../include/shockhash_amd.h"
#include "device_common.h"
#include "flat.h"
#include "sort.h"

namespace {

using shgpu::remix;

constexpr uint64_t UNSET = ~0ull;
constexpr double OVERLOAD_FACTOR = 0.9;  // ShockHash2FlatBase.h

struct FlatMap {
    uint32_t t[FLAT_MAX_RANGE];
    uint32_t range;
};
struct FlatLayers {
    uint64_t base[FLAT_MAX_LAYERS + 1];
    uint32_t layers;
};

__device__ __forceinline__ uint32_t fastrange32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

// compact_threshold (ShockHash2Flat.h:194-202): the first index whose threshold value is at least x (the last value is
// 2^32 - 1, so one always is).
__device__ __forceinline__ uint32_t compact_threshold(const FlatMap &map, uint32_t x) {
    for (uint32_t i = 0; i < map.range; i++)
        if (x <= map.t[i]) return i;
    return map.range - 1;
}

__device__ __forceinline__ uint64_t gtid() { return blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// Layer keys: the (rehashed) hash, the sort key bucket << 32 | threshold, and the position as the sorted value.
__global__ void k_flat_keys(const uint64_t *__restrict__ src, uint64_t c, uint32_t B, int rehash, uint64_t *mhc,
                            uint64_t *skey, uint64_t *val) {
    for (uint64_t j = gtid(); j < c; j += gstride()) {
        uint64_t m = src[j];
        if (rehash) m = remix(m);
        mhc[j] = m;
        skey[j] = ((uint64_t)fastrange32((uint32_t)m, B) << 32) | (m >> 32);
        val[j] = j;
    }
}

// start[b] = first sorted position of bucket b, start[B] = c (every position writes the starts of the buckets from
// its predecessor's bucket, exclusive, to its own, inclusive).
__global__ void k_flat_bounds(const uint64_t *__restrict__ skey, uint64_t c, uint32_t B, uint64_t *start) {
    for (uint64_t i = gtid(); i <= c; i += gstride()) {
        const int64_t b = i < c ? (int64_t)(skey[i] >> 32) : (int64_t)B;
        const int64_t prev = i > 0 ? (int64_t)(skey[i - 1] >> 32) : -1;
        for (int64_t q = prev + 1; q <= b; q++) start[q] = i;
    }
}

// flushBucket (ShockHash2Flat.h:204-233) for every bucket of the layer: threshold index, keys kept (a prefix of the
// bucket's sorted keys), free positions, keys bumped.
__global__ void k_flat_flush(const uint64_t *__restrict__ skey, const uint64_t *__restrict__ start, uint32_t B, uint32_t k,
                             FlatMap map, uint64_t base, uint8_t *thr_out, uint32_t *free_out, uint64_t *kept_out,
                             uint64_t *bump_cnt, uint32_t *err) {
    for (uint64_t b = gtid(); b < B; b += gstride()) {
        const uint64_t s = start[b], size = start[b + 1] - s;
        uint32_t t = map.range - 1;
        uint64_t kept = size;
        if (size > k) {
            const uint32_t last = compact_threshold(map, (uint32_t)skey[s + k - 1]);
            const uint32_t first_bumped = compact_threshold(map, (uint32_t)skey[s + k]);
            if (first_bumped == last && last == 0) {
                atomicOr(err, 1u);  // the reference throws: threshold index below 0
                t = 0;
            } else {
                t = first_bumped == last ? last - 1 : last;
            }
            const uint32_t unc = map.t[t];
            kept = 0;
            while (kept < k && (uint32_t)skey[s + kept] <= unc) kept++;  // sorted: the kept keys are a prefix
        }
        thr_out[base + b] = (uint8_t)t;
        free_out[base + b] = (uint32_t)(k - kept);
        kept_out[b] = kept;
        bump_cnt[b] = size - kept;
    }
}

// The bumped suffix of every bucket, in sorted order, to the next layer's arrays (idx_cur NULL: the positions are
// the input indices).
__global__ void k_flat_bump(const uint64_t *__restrict__ skey, const uint64_t *__restrict__ val, uint64_t c,
                            const uint64_t *__restrict__ start, const uint64_t *__restrict__ kept,
                            const uint64_t *__restrict__ bump_off, const uint64_t *__restrict__ mhc_cur,
                            const uint64_t *__restrict__ idx_cur, uint64_t *mhc_next, uint64_t *idx_next) {
    for (uint64_t i = gtid(); i < c; i += gstride()) {
        const uint64_t b = skey[i] >> 32, l = i - start[b];
        if (l < kept[b]) continue;
        const uint64_t dst = bump_off[b] + l - kept[b], j = val[i];
        mhc_next[dst] = mhc_cur[j];
        idx_next[dst] = idx_cur ? idx_cur[j] : j;
    }
}

// evaluateKPerfect's layers (ShockHash2Flat.h:280-291) for every key: its bucket, or UNSET when every layer bumped it.
__global__ void k_flat_place(const uint64_t *__restrict__ mhcs, uint64_t n, FlatLayers L, const uint8_t *__restrict__ thr,
                             FlatMap map, uint64_t *bkey, uint64_t *val) {
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        uint64_t m = mhcs[i], bucket = UNSET;
        for (uint32_t layer = 0; layer < L.layers; layer++) {
            if (layer != 0) m = remix(m);
            const uint64_t base = L.base[layer];
            const uint32_t b = fastrange32((uint32_t)m, (uint32_t)(L.base[layer + 1] - base));
            if ((uint32_t)(m >> 32) <= map.t[thr[base + b]]) {
                bucket = base + b;
                break;
            }
        }
        bkey[i] = bucket;
        val[i] = i;
    }
}

// The caller's buckets of the keys every layer bumped; err bit 2 when such a key had a bucket or is out of range.
__global__ void k_flat_fallback(const uint64_t *__restrict__ fb_index, const uint64_t *__restrict__ fb_bucket, uint64_t nfb,
                                uint64_t n, uint64_t *bkey, uint32_t *err) {
    for (uint64_t j = gtid(); j < nfb; j += gstride()) {
        const uint64_t i = fb_index[j];
        if (i >= n || bkey[i] != UNSET) {
            atomicOr(err, 2u);
            continue;
        }
        bkey[i] = fb_bucket[j];
    }
}

// err bit 4 when a key has no bucket or a bucket value that names neither a full bucket (below nbuckets) nor a place
// in the partial bucket ([nbuckets * k, limit)).
__global__ void k_flat_check_set(const uint64_t *__restrict__ bkey, uint64_t n, uint64_t nbuckets, uint32_t k, uint64_t limit,
                                 uint32_t *err) {
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint64_t b = bkey[i];
        if (b >= limit || (b >= nbuckets && b < nbuckets * k)) atomicOr(err, 4u);
    }
}

// Sorted by bucket: position j < nbuckets * k must hold bucket j / k (every bucket exactly k keys, err bit 8
// otherwise), and its key goes to out[j].
__global__ void k_flat_gather(const uint64_t *__restrict__ bkey, const uint64_t *__restrict__ val, uint64_t n, uint32_t k,
                              uint64_t nbuckets, const uint64_t *__restrict__ mhcs, uint64_t *out, uint32_t *err) {
    const uint64_t full = nbuckets * k;
    for (uint64_t j = gtid(); j < n; j += gstride()) {
        if (j < full) {
            if (bkey[j] != j / k) atomicOr(err, 8u);
            out[j] = mhcs[val[j]];
        } else if (bkey[j] < nbuckets) {
            atomicOr(err, 8u);
        }
    }
}

__global__ void k_flat_zero(uint32_t *p) { p[threadIdx.x] = 0; }

struct Arena {
    uint8_t *base = nullptr;
    size_t at = 0;
    void *take(size_t bytes) {
        void *p = base + at;
        at += (bytes + 255) / 256 * 256;
        return p;
    }
};

// The working memory: the caller's cache, grown to `bytes` when smaller.
int working_memory(size_t bytes, void **cache, size_t *cache_bytes, std::string &err, const char *what) {
    if (*cache_bytes >= bytes) return SHGPU_OK;
    if (*cache) (void)hipFree(*cache);
    *cache = nullptr;
    *cache_bytes = 0;
    if (!shgpu::device_malloc(cache, bytes)) {
        err = std::string(what) + ": hipMalloc of " + std::to_string(bytes) + " bytes failed";
        return SHGPU_ERR_NOMEM;
    }
    *cache_bytes = bytes;
    return SHGPU_OK;
}

int bits_for(uint64_t v) {  // bits that hold every value below v
    int b = 1;
    while (b < 64 && (1ull << b) < v) b++;
    return b;
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)); }

#define FL_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) { err = std::string("ShockHash2Flat bucketing: ") + hipGetErrorString(e_); return SHGPU_ERR_HIP; } \
    } while (0)

int check_map(const uint32_t *mapping, uint32_t range, FlatMap &map, std::string &err) {
    bool ok = mapping && range >= 2 && range <= FLAT_MAX_RANGE && mapping[range - 1] == 0xFFFFFFFFu;
    for (uint32_t i = 1; ok && i < range; i++) ok = mapping[i - 1] <= mapping[i];  // compact_threshold searches it
    if (!ok) {
        err = "ShockHash2Flat bucketing: threshold mapping must have 2.." + std::to_string(FLAT_MAX_RANGE) +
              " non-decreasing values ending at 2^32 - 1";
        return SHGPU_ERR_ARG;
    }
    std::memset(&map, 0, sizeof map);
    std::memcpy(map.t, mapping, range * 4);
    map.range = range;
    return SHGPU_OK;
}

}  // namespace

int shgpu_flat_layers_impl(hipStream_t st, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                           uint32_t range, uint8_t *h_thresholds, uint32_t *h_free, uint64_t *h_layer_bases,
                           uint32_t *h_layers, uint64_t *h_bumped_mhcs, uint64_t *h_bumped_index, uint64_t *h_nbumped,
                           std::string &err, void **cache, size_t *cache_bytes) {
    FlatMap map;
    int rc = check_map(mapping, range, map, err);
    if (rc) return rc;
    if (k < 2 || n >= (1ull << 31)) { err = "ShockHash2Flat bucketing: k >= 2 and fewer than 2^31 keys"; return SHGPU_ERR_ARG; }
    const uint64_t nbuckets = n / k;
    // working memory: layer hashes, sort keys and values before and after the sort, next layer's hashes and indices,
    // this layer's indices, sort scratch, per bucket starts, kept, bump counts and offsets, thresholds, free counts,
    // the error word and the scan scratch
    size_t sort_bytes = 0, scan_bytes = 0;
    FL_HIP(shgpu_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, (int)std::max<uint64_t>(n, 1), 64, st));
    FL_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t)0,
                                   (size_t)std::max<uint64_t>(nbuckets + 1, 1), rocprim::plus<uint64_t>(), st));
    const uint64_t nb1 = nbuckets + 2;
    const size_t parts[] = {(n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8,
                            (n + 1) * 8, sort_bytes + 16, nb1 * 8, nb1 * 8, nb1 * 8, nb1 * 8, nb1, nb1 * 4, 64, scan_bytes + 16};
    size_t bytes = 0;
    for (size_t b : parts) bytes += (b + 255) / 256 * 256;
    if ((rc = working_memory(bytes, cache, cache_bytes, err, "ShockHash2Flat bucketing"))) return rc;
    Arena a{(uint8_t *)*cache, 0};
    uint64_t *mhc = (uint64_t *)a.take(parts[0]), *skey = (uint64_t *)a.take(parts[1]), *val = (uint64_t *)a.take(parts[2]);
    uint64_t *skey_s = (uint64_t *)a.take(parts[3]), *val_s = (uint64_t *)a.take(parts[4]);
    uint64_t *mhc_next = (uint64_t *)a.take(parts[5]), *idx_next = (uint64_t *)a.take(parts[6]), *idx_cur = (uint64_t *)a.take(parts[7]);
    void *sort_tmp = a.take(parts[8]);
    uint64_t *start = (uint64_t *)a.take(parts[9]), *kept = (uint64_t *)a.take(parts[10]), *bump_cnt = (uint64_t *)a.take(parts[11]),
             *bump_off = (uint64_t *)a.take(parts[12]);
    uint8_t *thr = (uint8_t *)a.take(parts[13]);
    uint32_t *freec = (uint32_t *)a.take(parts[14]);
    uint32_t *errw = (uint32_t *)a.take(parts[15]);
    void *scan_tmp = a.take(parts[16]);
    k_flat_zero<<<1, 16, 0, st>>>(errw);  // a kernel, not a memset node, should a caller capture this
    FL_HIP(hipGetLastError());

    std::vector<uint64_t> bases{0};
    uint32_t layers = 2;  // ShockHash2FlatBase.h
    uint64_t B = (uint64_t)std::ceil(OVERLOAD_FACTOR * (double)nbuckets);
    if (B == 0) layers = 0;
    const uint64_t *cur_mhc = d_mhcs;
    const uint64_t *cur_idx = nullptr;  // layer 0: the positions are the input indices
    uint64_t c = n;
    for (uint32_t layer = 0; layer < layers; layer++) {
        const uint64_t base = bases.back();
        if (layer != 0) {
            B = (uint64_t)(OVERLOAD_FACTOR * (double)(c / k));
            B = std::min<uint64_t>(B, nbuckets - base);
            if (B == 0) {
                layers = 1;
                break;
            }
        }
        bases.push_back(base + B);
        const uint32_t g = grid_for(c + 1);
        k_flat_keys<<<g, 256, 0, st>>>(cur_mhc, c, (uint32_t)B, layer != 0 ? 1 : 0, mhc, skey, val);
        FL_HIP(hipGetLastError());
        if (c) {
            size_t sb = sort_bytes;
            FL_HIP(shgpu_sort_pairs_u64(sort_tmp, sb, skey, skey_s, val, val_s, (int)c, 32 + bits_for(B), st));
        }
        k_flat_bounds<<<g, 256, 0, st>>>(skey_s, c, (uint32_t)B, start);
        FL_HIP(hipGetLastError());
        k_flat_flush<<<grid_for(B), 256, 0, st>>>(skey_s, start, (uint32_t)B, k, map, base, thr, freec, kept, bump_cnt, errw);
        FL_HIP(hipGetLastError());
        size_t scb = scan_bytes;
        FL_HIP(rocprim::exclusive_scan(scan_tmp, scb, bump_cnt, bump_off, (uint64_t)0, (size_t)B, rocprim::plus<uint64_t>(), st));
        uint64_t last_off = 0, last_cnt = 0;
        FL_HIP(hipMemcpyAsync(&last_off, bump_off + B - 1, 8, hipMemcpyDeviceToHost, st));
        FL_HIP(hipMemcpyAsync(&last_cnt, bump_cnt + B - 1, 8, hipMemcpyDeviceToHost, st));
        k_flat_bump<<<g, 256, 0, st>>>(skey_s, val_s, c, start, kept, bump_off, mhc, cur_idx, mhc_next, idx_next);
        FL_HIP(hipGetLastError());
        FL_HIP(hipStreamSynchronize(st));
        c = last_off + last_cnt;
        // the next layer reads the bumped keys; their indices move to idx_cur (idx_next is rewritten next)
        FL_HIP(hipMemcpyAsync(idx_cur, idx_next, c * 8, hipMemcpyDeviceToDevice, st));
        FL_HIP(hipMemcpyAsync(mhc, mhc_next, c * 8, hipMemcpyDeviceToDevice, st));
        cur_mhc = mhc;
        cur_idx = idx_cur;
    }
    uint32_t e = 0;
    FL_HIP(hipMemcpyAsync(&e, errw, 4, hipMemcpyDeviceToHost, st));
    const uint64_t handled = bases.back();
    if (handled) {
        FL_HIP(hipMemcpyAsync(h_thresholds, thr, handled, hipMemcpyDeviceToHost, st));
        FL_HIP(hipMemcpyAsync(h_free, freec, handled * 4, hipMemcpyDeviceToHost, st));
    }
    if (c && layers) {
        FL_HIP(hipMemcpyAsync(h_bumped_mhcs, mhc, c * 8, hipMemcpyDeviceToHost, st));
        FL_HIP(hipMemcpyAsync(h_bumped_index, idx_cur, c * 8, hipMemcpyDeviceToHost, st));
    } else if (c) {  // no layer at all: every key goes to the fallback, in input order
        FL_HIP(hipMemcpyAsync(h_bumped_mhcs, d_mhcs, c * 8, hipMemcpyDeviceToHost, st));
    }
    FL_HIP(hipStreamSynchronize(st));
    if (c && !layers)
        for (uint64_t i = 0; i < c; i++) h_bumped_index[i] = i;
    if (e & 1) { err = "ShockHash2Flat: threshold index below 0"; return SHGPU_ERR_ARG; }
    for (uint32_t l = 0; l <= layers; l++) h_layer_bases[l] = bases[l];
    *h_layers = layers;
    *h_nbumped = c;
    return SHGPU_OK;
}

int shgpu_flat_bucket_keys_impl(hipStream_t st, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                                uint32_t range, const uint64_t *h_layer_bases, uint32_t layers, const uint8_t *h_thresholds,
                                const uint64_t *h_fb_index, const uint64_t *h_fb_bucket, uint64_t nfb,
                                uint64_t *d_bucket_keys, std::string &err, void **cache, size_t *cache_bytes) {
    FlatMap map;
    int rc = check_map(mapping, range, map, err);
    if (rc) return rc;
    if (k < 2 || n >= (1ull << 31) || layers > FLAT_MAX_LAYERS) {
        err = "ShockHash2Flat placement: k >= 2, fewer than 2^31 keys, at most 2 layers";
        return SHGPU_ERR_ARG;
    }
    const uint64_t nbuckets = n / k;
    FlatLayers L;
    std::memset(&L, 0, sizeof L);
    L.layers = layers;
    for (uint32_t l = 0; l <= layers; l++) L.base[l] = h_layer_bases[l];
    if (L.base[0] != 0 || L.base[layers] > nbuckets) { err = "ShockHash2Flat placement: layer bases beyond the buckets"; return SHGPU_ERR_ARG; }
    for (uint32_t l = 0; l < layers; l++)
        if (L.base[l + 1] <= L.base[l] || L.base[l + 1] - L.base[l] >= (1ull << 32)) {
            err = "ShockHash2Flat placement: empty or oversized layer";
            return SHGPU_ERR_ARG;
        }
    const uint64_t handled = L.base[layers];
    size_t sort_bytes = 0;
    FL_HIP(shgpu_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, (int)std::max<uint64_t>(n, 1), 64, st));
    const size_t parts[] = {(n + 1) * 8, (n + 1) * 8, (n + 1) * 8, (n + 1) * 8, sort_bytes + 16, handled + 8,
                            (nfb + 1) * 8, (nfb + 1) * 8, 64};
    size_t bytes = 0;
    for (size_t b : parts) bytes += (b + 255) / 256 * 256;
    if ((rc = working_memory(bytes, cache, cache_bytes, err, "ShockHash2Flat placement"))) return rc;
    Arena a{(uint8_t *)*cache, 0};
    uint64_t *bkey = (uint64_t *)a.take(parts[0]), *val = (uint64_t *)a.take(parts[1]), *bkey_s = (uint64_t *)a.take(parts[2]),
             *val_s = (uint64_t *)a.take(parts[3]);
    void *sort_tmp = a.take(parts[4]);
    uint8_t *thr = (uint8_t *)a.take(parts[5]);
    uint64_t *fbi = (uint64_t *)a.take(parts[6]), *fbb = (uint64_t *)a.take(parts[7]);
    uint32_t *errw = (uint32_t *)a.take(parts[8]);
    k_flat_zero<<<1, 16, 0, st>>>(errw);  // a kernel, not a memset node, should a caller capture this
    FL_HIP(hipGetLastError());
    if (handled) FL_HIP(hipMemcpyAsync(thr, h_thresholds, handled, hipMemcpyHostToDevice, st));
    if (nfb) {
        FL_HIP(hipMemcpyAsync(fbi, h_fb_index, nfb * 8, hipMemcpyHostToDevice, st));
        FL_HIP(hipMemcpyAsync(fbb, h_fb_bucket, nfb * 8, hipMemcpyHostToDevice, st));
    }
    const uint32_t g = grid_for(n);
    k_flat_place<<<g, 256, 0, st>>>(d_mhcs, n, L, thr, map, bkey, val);
    FL_HIP(hipGetLastError());
    k_flat_fallback<<<grid_for(nfb), 256, 0, st>>>(fbi, fbb, nfb, n, bkey, errw);
    FL_HIP(hipGetLastError());
    // the partial bucket's keys take the values k * nbuckets + i, i below the number of free positions (< n)
    const uint64_t limit = nbuckets * k + n;
    k_flat_check_set<<<g, 256, 0, st>>>(bkey, n, nbuckets, k, limit, errw);
    FL_HIP(hipGetLastError());
    uint32_t e = 0;
    FL_HIP(hipMemcpyAsync(&e, errw, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (e) { err = "ShockHash2Flat placement: a key has no bucket, or a fallback key was placed twice"; return SHGPU_ERR_ARG; }
    if (n) {
        size_t sb = sort_bytes;
        FL_HIP(shgpu_sort_pairs_u64(sort_tmp, sb, bkey, bkey_s, val, val_s, (int)n, bits_for(limit), st));
    }
    k_flat_gather<<<g, 256, 0, st>>>(bkey_s, val_s, n, k, nbuckets, d_mhcs, d_bucket_keys, errw);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(&e, errw, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (e) { err = "ShockHash2Flat: a bucket does not hold k keys"; return SHGPU_ERR_ARG; }
    return SHGPU_OK;
}
