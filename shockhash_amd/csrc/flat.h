// ShockHash2Flat's bucketing on the device (flat.hip), called by the C ABI: the threshold layers with bumping
// (include/ShockHash2Flat.h:70-124, flushBucket :204-233) and the placement of every key into its bucket of k keys
// (include/ShockHash2Flat.h:148-157, evaluateKPerfect :280-302).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

constexpr uint32_t FLAT_MAX_RANGE = 64;  // threshold indices of k <= 128 (2^(floor(log2 k) - 1))
constexpr uint32_t FLAT_MAX_LAYERS = 2;

// The threshold layers of n keys' 64-bit hashes (device array d_mhcs) for buckets of k keys: per bucket of the
// layers its threshold index (h_thresholds, the first h_layer_bases[*h_layers] of n / k entries) and its free
// positions (h_free: k minus the keys it keeps), the layer bases (h_layer_bases, *h_layers + 1 entries), and the
// keys bumped from every layer in the order the reference leaves them, as their last layer's hash (h_bumped_mhcs)
// and their input index (h_bumped_index), *h_nbumped of them. mapping: the range threshold values.
int shgpu_flat_layers_impl(hipStream_t st, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                           uint32_t range, uint8_t *h_thresholds, uint32_t *h_free, uint64_t *h_layer_bases,
                           uint32_t *h_layers, uint64_t *h_bumped_mhcs, uint64_t *h_bumped_index, uint64_t *h_nbumped,
                           std::string &err, void **cache, size_t *cache_bytes);

// Every key's bucket (the first layer whose stored threshold admits it, else the placement the caller computed
// for the bumped keys: h_fb_bucket[j] for the key of input index h_fb_index[j]; values from (n / k) * k on name the
// last, partial bucket), checked so that every one of the n / k buckets receives exactly k keys, and the keys of
// bucket b in input order at d_bucket_keys[b * k .. (b + 1) * k).
int shgpu_flat_bucket_keys_impl(hipStream_t st, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                                uint32_t range, const uint64_t *h_layer_bases, uint32_t layers, const uint8_t *h_thresholds,
                                const uint64_t *h_fb_index, const uint64_t *h_fb_bucket, uint64_t nfb,
                                uint64_t *d_bucket_keys, std::string &err, void **cache, size_t *cache_bytes);
