// Device construction of the one-bit retrieval structure (ribbon.hip), called by the C ABI. The layout's constants
// live in shgpu_types.h, which the host query (shockhash::RibbonRetrieval) includes too.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "shgpu_types.h"

constexpr uint32_t RB_WIDTH = SHGPU_RIBBON_WIDTH;
constexpr uint64_t RB_SHARD_KEYS = SHGPU_RIBBON_SHARD_KEYS;
constexpr uint64_t RB_SALT_SHARD = SHGPU_RIBBON_SALT_SHARD;
constexpr uint64_t RB_SALT_ROW = SHGPU_RIBBON_SALT_ROW;
constexpr uint32_t RB_MAX_SLOTS = 65535;  // slots of one shard: its start counts and cursors are u16
constexpr uint32_t RB_WAVES_PER_CU = 16;  // persistent solve waves per CU

// Builds the structure for n keys and their bits (device arrays) on stream s; writes the S + 1 slot offsets, the S
// seed bytes and the packed slot bits (*words_used words, at most shgpu_ribbon_max_words(n)) to host memory.
// Returns an SHGPU status, message in err. Working memory: *cache (*cache_bytes bytes, grown as needed and kept by
// the caller) when cache is not NULL, else an allocation of this call.
int shgpu_ribbon_build_impl(hipStream_t s, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint32_t *h_offsets,
                            uint8_t *h_seeds, uint64_t *h_words, uint64_t words_capacity, uint64_t *words_used,
                            std::string &err, void **cache = nullptr, size_t *cache_bytes = nullptr);
