// shockhash::SIMDShockHash<LEAF_SIZE, ROTATION_FITTING> backed by the MI355X engine.
// Same constructors, operator() and descriptor layout as the reference class
// (include/SIMDShockHash.hpp:104-867): 32-bit fmix split hash, start seeds 17/18/19 for the
// aggregation levels and the leaves.
#pragma once
#include <fstream>
#include <span>
#include <string>
#include <vector>

#include "shockhash/GpuRecSplit.h"

namespace shockhash {

template <size_t LEAF_SIZE, bool ROTATION_FITTING = false>
class SIMDShockHash : public detail::GpuRecSplit<SHGPU_SIMD_SHOCKHASH, LEAF_SIZE, ROTATION_FITTING> {
    static_assert(LEAF_SIZE <= MAX_LEAF_SIZE);

  public:
    SIMDShockHash() {}
    SIMDShockHash(std::span<const std::string> keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct_strings(keys, bucket_size, num_threads);
    }
    SIMDShockHash(std::span<hash128_t> keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct(keys.data(), keys.size(), bucket_size, num_threads);
    }
    SIMDShockHash(std::ifstream &input, const size_t bucket_size, size_t num_threads = 1) {
        std::vector<hash128_t> h;
        for (std::string key; std::getline(input, key);) h.push_back(spooky(key.c_str(), key.size(), 0));
        this->construct(h.data(), h.size(), bucket_size, num_threads);
    }
    size_t operator()(const hash128_t &hash) { return this->query(hash); }
    size_t operator()(const std::string &key) { return this->query(spooky(key.c_str(), key.size(), 0)); }
};

}  // namespace shockhash
