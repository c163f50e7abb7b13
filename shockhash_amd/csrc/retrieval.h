// Device construction of the one-bit retrieval structure (retrieval.hip), called by the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

// (shgpu_retrieval_cells_per_third, the layout's size rule, is part of the C ABI: include/shockhash_amd.h.)
// Builds the structure for n keys and their bits (device arrays) on stream s; writes the seed, the cells per
// third and ceil(3 * seg / 64) packed words to host memory. Returns an SHGPU status, message in err.
// Working memory: *cache (*cache_bytes bytes, grown as needed and kept by the caller) when cache is not
// NULL, else an allocation of this call.
int shgpu_retrieval_build_impl(hipStream_t s, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint64_t *seed,
                               uint64_t *seg, uint64_t *h_words, uint64_t words_capacity, std::string &err,
                               void **cache = nullptr, size_t *cache_bytes = nullptr);

#include "shgpu_types.h"
// Retrieval input of n leaves (one thread per leaf replays its cuckoo insertions); bipartite: ShockHash2.
int shgpu_leaf_orientation_impl(hipStream_t s, int bipartite, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                                uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits, std::string &err);
