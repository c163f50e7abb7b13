// Device replay of the leaves' cuckoo insertions (orientation.hip), called by the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "shgpu_types.h"
// Retrieval input of n leaves (one thread per leaf replays its cuckoo insertions); bipartite: ShockHash2.
int shgpu_leaf_orientation_impl(hipStream_t s, int bipartite, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                                uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits, std::string &err);
