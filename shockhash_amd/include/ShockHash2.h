// shockhash::ShockHash2<LEAF_SIZE> (bipartite ShockHash) backed by the MI355X engine.
// Same constructors, const operator() and descriptor layout as the reference class
// (include/ShockHash2.h:117-532). The GPU engine's ShockHash2 leaf search covers leaf sizes up to 64.
#pragma once
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "shockhash/GpuRecSplit.h"

namespace shockhash {

template <size_t LEAF_SIZE>
class ShockHash2 : public detail::GpuRecSplit<SHGPU_SHOCKHASH2, LEAF_SIZE> {
    static_assert(LEAF_SIZE <= MAX_LEAF_SIZE2);

  public:
    ShockHash2() {}
    ShockHash2(std::span<const std::string> keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct_strings(keys, bucket_size, num_threads);
    }
    // String keys given as views (not in the reference interface): the same function as from std::string keys,
    // for callers that hold the keys' bytes elsewhere and need not allocate a std::string per key.
    ShockHash2(std::span<const std::string_view> keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct_strings(keys, bucket_size, num_threads);
    }
    ShockHash2(std::span<hash128_t> keys, const size_t bucket_size, size_t num_threads = 1) {
        this->construct(keys.data(), keys.size(), bucket_size, num_threads);
    }
    size_t operator()(const hash128_t &hash) const { return this->query(hash); }
    size_t operator()(const std::string &key) const { return this->query(spooky(key.c_str(), key.size(), 0)); }
};

}  // namespace shockhash
