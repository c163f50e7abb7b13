// The compiled ShockHash2 leaf seam with the reference's signatures (include/shockhash/ShockHash2-precompiled.h:8-10),
// implemented by libShockHash2Precompiled.so over the MI355X engine. A code base that links the reference's
// ShockHash2Precompiled library can link this one instead: ShockHash2::recSplit's leaf calls
// (include/ShockHash2.h:273) and queries (:228) keep their source.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace shockhash {
// Cell of `key` in a ShockHash2 leaf of `param` keys with code `seed`, given its retrieved bit; 0 for
// param <= 1 or param > 128 (the reference's dispatch covers leaf sizes 2..128).
size_t shockhash2query(size_t param, size_t seed, uint64_t key, size_t retrieved);
// Searches the leaf code of leafKeys (leafKeys.size() must equal param, 2 <= param <= 128) on the GPU and
// appends the leaf's retrieval input (key, bit) to ribbonInput in key order; returns the code. param <= 1 or
// param > 128: returns 0 and appends nothing. Throws shockhash::gpu::Error (with the engine's status) when the
// engine fails or the search exhausts its bounds (repeated keys), std::invalid_argument on a size mismatch.
size_t shockhash2construct(size_t param, std::vector<uint64_t> &leafKeys,
                           std::vector<std::pair<uint64_t, uint8_t>> &ribbonInput);
}  // namespace shockhash
