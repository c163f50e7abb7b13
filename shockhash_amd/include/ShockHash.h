// shockhash::ShockHash<LEAF_SIZE, ROTATION_FITTING> backed by the MI355X engine.
// Same constructors, operator(), size/getBits/printBits and descriptor layout as the reference class
// (include/ShockHash.h:134-691). Construction runs every seed search on the GPU; encoding, retrieval
// and queries run on the host.
#pragma once
#include <span>
#include <string>
#include <vector>

#include "shockhash/GpuRecSplit.h"

namespace shockhash {

template <size_t LEAF_SIZE, bool ROTATION_FITTING = false>
class ShockHash : public detail::GpuRecSplit<SHGPU_SHOCKHASH, LEAF_SIZE, ROTATION_FITTING> {
    static_assert(LEAF_SIZE <= MAX_LEAF_SIZE);

  public:
    ShockHash() {}
    ShockHash(const std::vector<std::string> &keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct_strings(keys, bucket_size, num_threads);
    }
    // The reference sorts the caller's array in place; this engine leaves it unchanged.
    ShockHash(std::vector<hash128_t> &keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct(keys.data(), keys.size(), bucket_size, num_threads);
    }
    size_t operator()(const hash128_t &hash) { return this->query(hash); }
    size_t operator()(const std::string &key) { return this->query(spooky(key.c_str(), key.size(), 0)); }
};

}  // namespace shockhash
