// Host-side handle of the MI355X seed-search engine (libshockhash_amd.so, C ABI in
// include/shockhash_amd.h). One engine per process and device, created on first use.
#pragma once
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/shockhash_amd.h"
#include "Hashing.h"

namespace shockhash {
namespace gpu {

class Error : public std::runtime_error {
  public:
    int status;
    Error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

inline void check(int rc, const char *what) {
    if (rc != SHGPU_OK)
        throw Error(rc, std::string(what) + ": " + shgpu_status_string(rc) + ": " + shgpu_last_error());
}

// Seed-search results of one construction, everything the host encoder needs.
struct Forest {
    uint64_t nbuckets = 0;
    std::vector<uint64_t> size_acc;  // nbuckets + 1
    std::vector<shgpu_node> nodes;   // DFS pre-order per bucket, buckets in order
    std::vector<uint64_t> codes;     // per node: found seed - start seed (ShockHash2 leaves: pair code)
    std::vector<uint64_t> keys;      // fingerprint .second in final leaf order
    double ms[4] = {0, 0, 0, 0};     // sort + bucketing, tree setup, seed search, search kernel
};

class Engine {
    shgpu_engine *e = nullptr;
    std::mutex mtx;

  public:
    explicit Engine(int device = 0) { check(shgpu_engine_create(device, &e), "shgpu_engine_create"); }
    ~Engine() { shgpu_engine_destroy(e); }
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    static Engine &instance() {
        static Engine eng(0);
        return eng;
    }

    shgpu_engine *handle() { return e; }

    Forest build(int variant, uint32_t leaf_size, const hash128_t *hashes, uint64_t n, uint64_t bucket_size) {
        static_assert(sizeof(hash128_t) == 16, "hash128_t must be two packed 64-bit words");
        std::lock_guard<std::mutex> lock(mtx);
        const shgpu_build_desc d = desc(variant, leaf_size, bucket_size);
        shgpu_build *b = nullptr;
        check(shgpu_build_run_host(e, &d, reinterpret_cast<const uint64_t *>(hashes), n, &b), "shgpu_build_run");
        return collect(b, n);
    }

    // String keys: packed and fingerprinted on the device (shgpu_hash_strings, SpookyHash V2 as
    // first_hash), then built from the device fingerprints.
    template <class Strings>
    Forest build_strings(int variant, uint32_t leaf_size, const Strings &keys, uint64_t bucket_size) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t n = keys.size();
        std::vector<uint64_t> offsets(n + 1, 0);
        for (uint64_t i = 0; i < n; i++) offsets[i + 1] = offsets[i] + keys[i].size();
        std::vector<uint8_t> data(offsets[n] + 8);
        for (uint64_t i = 0; i < n; i++) std::memcpy(data.data() + offsets[i], keys[i].data(), keys[i].size());
        void *d_data = nullptr, *d_off = nullptr, *d_hashes = nullptr;
        int rc = shgpu_device_alloc(e, data.size(), &d_data);
        if (rc == SHGPU_OK) rc = shgpu_device_alloc(e, offsets.size() * 8, &d_off);
        if (rc == SHGPU_OK) rc = shgpu_device_alloc(e, n * 16 + 16, &d_hashes);
        if (rc == SHGPU_OK) rc = shgpu_copy_to_device(e, d_data, data.data(), data.size());
        if (rc == SHGPU_OK) rc = shgpu_copy_to_device(e, d_off, offsets.data(), offsets.size() * 8);
        if (rc == SHGPU_OK) rc = shgpu_hash_strings(e, (const uint8_t *)d_data, (const uint64_t *)d_off, n, (uint64_t *)d_hashes);
        shgpu_build *b = nullptr;
        const shgpu_build_desc d = desc(variant, leaf_size, bucket_size);
        if (rc == SHGPU_OK) rc = shgpu_build_run(e, &d, (const uint64_t *)d_hashes, n, &b);
        const std::string msg = rc == SHGPU_OK ? "" : shgpu_last_error();
        if (d_data) shgpu_device_free(e, d_data);
        if (d_off) shgpu_device_free(e, d_off);
        if (d_hashes) shgpu_device_free(e, d_hashes);
        if (rc != SHGPU_OK) throw Error(rc, std::string("shgpu_build_run (strings): ") + shgpu_status_string(rc) + ": " + msg);
        return collect(b, n);
    }

  private:
    static shgpu_build_desc desc(int variant, uint32_t leaf_size, uint64_t bucket_size) {
        shgpu_build_desc d{};
        d.variant = variant;
        d.leaf_size = leaf_size;
        d.bucket_size = bucket_size;
        d.mode = SHGPU_MODE_PERSISTENT;
        d.copy_keys = 1;
        return d;
    }

    Forest collect(shgpu_build *b, uint64_t n) {
        Forest f;
        f.nbuckets = shgpu_build_nbuckets(b);
        f.size_acc.resize(f.nbuckets + 1);
        const uint64_t nn = shgpu_build_nnodes(b);
        f.nodes.resize(nn);
        f.codes.resize(nn);
        f.keys.resize(n);
        int rc = shgpu_build_bucket_sizes(b, f.size_acc.data());
        if (rc == SHGPU_OK) rc = shgpu_build_nodes(b, f.nodes.data(), f.codes.data());
        if (rc == SHGPU_OK && n) rc = shgpu_build_keys(b, f.keys.data());
        if (rc == SHGPU_OK) rc = shgpu_build_timings(b, f.ms);
        shgpu_build_free(b);
        check(rc, "shgpu_build results");
        return f;
    }

  public:
    // The one-bit retrieval structure of (key, bit) pairs, built on the device (shgpu_retrieval_build):
    // seed, cells per third and packed cell bits in the layout shockhash::BitRetrieval reads.
    void retrieval(const std::vector<std::pair<uint64_t, uint8_t>> &in, uint64_t &seed, uint64_t &seg,
                   std::vector<uint64_t> &words) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t n = in.size();
        std::vector<uint64_t> keys(n);
        std::vector<uint8_t> bits(n);
        for (uint64_t i = 0; i < n; i++) {
            keys[i] = in[i].first;
            bits[i] = in[i].second;
        }
        void *d_keys = nullptr, *d_bits = nullptr;
        int rc = shgpu_device_alloc(e, n * 8 + 8, &d_keys);
        if (rc == SHGPU_OK) rc = shgpu_device_alloc(e, n + 8, &d_bits);
        if (rc == SHGPU_OK && n) rc = shgpu_copy_to_device(e, d_keys, keys.data(), n * 8);
        if (rc == SHGPU_OK && n) rc = shgpu_copy_to_device(e, d_bits, bits.data(), n);
        words.assign(shgpu_retrieval_words(n), 0);
        if (rc == SHGPU_OK)
            rc = shgpu_retrieval_build(e, (const uint64_t *)d_keys, (const uint8_t *)d_bits, n, &seed, &seg, words.data(),
                                       words.size());
        if (d_keys) shgpu_device_free(e, d_keys);
        if (d_bits) shgpu_device_free(e, d_bits);
        check(rc, "shgpu_retrieval_build");
    }
};

}  // namespace gpu
}  // namespace shockhash
