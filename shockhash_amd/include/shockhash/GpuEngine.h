// Host-side handle of the MI355X seed-search engine (libshockhash_amd.so, C ABI in
// include/shockhash_amd.h). One engine per process and device, created on first use.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/shockhash_amd.h"
#include "Hashing.h"
#include "Parallel.h"

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

// Whether the class templates keep each build's retrieval input (the leaf orientations in the reference's
// order) on the host for verification; off by default (the reference keeps none), the tests turn it on.
inline bool keep_retrieval_input = false;

// Seed-search results of one construction, everything the host encoder needs.
struct Forest {
    uint64_t nbuckets = 0;
    std::vector<uint64_t> size_acc;  // nbuckets + 1
    std::vector<shgpu_node> nodes;   // DFS pre-order per bucket, buckets in order
    std::vector<uint64_t> codes;     // per node: found seed - start seed (ShockHash2 leaves: pair code)
    double ms[4] = {0, 0, 0, 0};     // sort + bucketing, tree setup, seed search, search kernel
    // The engine's build result while it holds the fingerprint .second values in final leaf order in device
    // memory (the orientation pass reads them there); released once the retrieval structure is built.
    std::shared_ptr<shgpu_build> device;
};

class Engine {
    shgpu_engine *e = nullptr;
    std::mutex mtx;
    // Device buffers of the orientation pass, grown as needed and kept between constructions.
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
    };
    Buf tasks_buf, out_keys_buf, out_bits_buf;
    Buf leaf_keys_buf, leaf_task_buf, leaf_code_buf;  // one ShockHash2 leaf at a time (sh2LeafCode)
    Buf flat_mhc_buf;                                 // ShockHash2Flat: the keys' hashes while a construction runs
    Buf str_data_buf, str_off_buf, str_hash_buf;      // string keys: bytes, offsets and fingerprints on the device
    Buf str_data_host, str_off_host;                  // ... and the packed bytes and offsets in page-locked memory
    Buf orient_host;                                  // the orientation pass's task list, in page-locked memory
    Buf ribbon_host;                                  // the retrieval structure as it comes back, page-locked
    int ensure(Buf &b, size_t bytes) {
        if (b.bytes >= bytes) return SHGPU_OK;
        if (b.p) shgpu_device_free(e, b.p);
        b.p = nullptr;
        b.bytes = 0;
        const int rc = shgpu_device_alloc(e, bytes, &b.p);
        if (rc == SHGPU_OK) b.bytes = bytes;
        return rc;
    }
    int ensure_host(Buf &b, size_t bytes) {
        if (b.bytes >= bytes) return SHGPU_OK;
        if (b.p) shgpu_host_free(e, b.p);
        b.p = nullptr;
        b.bytes = 0;
        const int rc = shgpu_host_alloc(e, bytes, &b.p);
        if (rc == SHGPU_OK) b.bytes = bytes;
        return rc;
    }

  public:
    explicit Engine(int device = 0) { check(shgpu_engine_create(device, &e), "shgpu_engine_create"); }
    ~Engine() {
        for (Buf *b : {&tasks_buf, &out_keys_buf, &out_bits_buf, &leaf_keys_buf, &leaf_task_buf, &leaf_code_buf, &flat_mhc_buf,
                       &str_data_buf, &str_off_buf, &str_hash_buf})
            if (b->p) shgpu_device_free(e, b->p);
        for (Buf *b : {&str_data_host, &str_off_host, &orient_host, &ribbon_host})
            if (b->p) shgpu_host_free(e, b->p);
        shgpu_engine_destroy(e);
    }
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    static Engine &instance() {
        static Engine eng(0);
        return eng;
    }

    shgpu_engine *handle() { return e; }

    Forest build(int variant, uint32_t leaf_size, const hash128_t *hashes, uint64_t n, uint64_t bucket_size,
                 bool rotation = false) {
        static_assert(sizeof(hash128_t) == 16, "hash128_t must be two packed 64-bit words");
        std::lock_guard<std::mutex> lock(mtx);
        const shgpu_build_desc d = desc(variant, leaf_size, bucket_size, rotation);
        shgpu_build *b = nullptr;
        check(shgpu_build_run_host(e, &d, reinterpret_cast<const uint64_t *>(hashes), n, &b), "shgpu_build_run");
        return collect(b, n);
    }

    // String keys: packed on host threads into page-locked memory, copied by DMA and fingerprinted on the device
    // (shgpu_hash_strings, SpookyHash V2 as first_hash), then built from the device fingerprints. The packed and
    // device buffers are kept between constructions.
    template <class Strings>
    Forest build_strings(int variant, uint32_t leaf_size, const Strings &keys, uint64_t bucket_size, bool rotation = false) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t n = keys.size();
        const size_t nt = hostThreads(n);
        std::vector<uint64_t> chunk(nt + 1, 0);  // bytes before each thread's range
        parallelRanges(n, nt, [&](size_t t, size_t lo, size_t hi) {
            uint64_t b = 0;
            for (size_t i = lo; i < hi; i++) b += keys[i].size();
            chunk[t + 1] = b;
        });
        for (size_t t = 0; t < nt; t++) chunk[t + 1] += chunk[t];
        const uint64_t total = chunk[nt];
        // the hashing kernel may read up to 8 bytes past the last key
        int rc = ensure_host(str_data_host, total + 8);
        if (rc == SHGPU_OK) rc = ensure_host(str_off_host, (n + 1) * 8);
        check(rc, "shgpu_host_alloc (string keys)");
        uint8_t *data = (uint8_t *)str_data_host.p;
        uint64_t *off = (uint64_t *)str_off_host.p;
        parallelRanges(n, nt, [&](size_t t, size_t lo, size_t hi) {
            uint64_t o = chunk[t];
            for (size_t i = lo; i < hi; i++) {
                off[i] = o;
                std::memcpy(data + o, keys[i].data(), keys[i].size());
                o += keys[i].size();
            }
        });
        off[n] = total;
        std::memset(data + total, 0, 8);
        rc = ensure(str_data_buf, total + 8);
        if (rc == SHGPU_OK) rc = ensure(str_off_buf, (n + 1) * 8);
        if (rc == SHGPU_OK) rc = ensure(str_hash_buf, n * 16 + 16);
        if (rc == SHGPU_OK) rc = shgpu_copy_to_device(e, str_data_buf.p, data, total + 8);
        if (rc == SHGPU_OK) rc = shgpu_copy_to_device(e, str_off_buf.p, off, (n + 1) * 8);
        if (rc == SHGPU_OK)
            rc = shgpu_hash_strings(e, (const uint8_t *)str_data_buf.p, (const uint64_t *)str_off_buf.p, n, (uint64_t *)str_hash_buf.p);
        shgpu_build *b = nullptr;
        const shgpu_build_desc d = desc(variant, leaf_size, bucket_size, rotation);
        if (rc == SHGPU_OK) rc = shgpu_build_run(e, &d, (const uint64_t *)str_hash_buf.p, n, &b);
        if (rc != SHGPU_OK)
            throw Error(rc, std::string("shgpu_build_run (strings): ") + shgpu_status_string(rc) + ": " + shgpu_last_error());
        return collect(b, n);
    }

    // The ShockHash2 leaf code of m keys (BijectionsShockHash2::findSeed, ShockHash2-internal.h:501-546),
    // searched on the device by shgpu_sh2_leaf_search: the compiled seam shockhash2construct runs one leaf per call.
    uint64_t sh2LeafCode(const uint64_t *keys, uint32_t m) {
        std::lock_guard<std::mutex> lock(mtx);
        const shgpu_leaf_task task{0, m, 0, 0};
        uint64_t code = SHGPU_NOT_FOUND;
        int rc = ensure(leaf_keys_buf, (size_t)m * 8 + 8);
        if (rc == SHGPU_OK) rc = ensure(leaf_task_buf, sizeof(task));
        if (rc == SHGPU_OK) rc = ensure(leaf_code_buf, 8);
        if (rc == SHGPU_OK && m) rc = shgpu_copy_to_device(e, leaf_keys_buf.p, keys, (size_t)m * 8);
        if (rc == SHGPU_OK) rc = shgpu_copy_to_device(e, leaf_task_buf.p, &task, sizeof(task));
        if (rc == SHGPU_OK)
            rc = shgpu_sh2_leaf_search(e, (const uint64_t *)leaf_keys_buf.p, (const shgpu_leaf_task *)leaf_task_buf.p, 1,
                                       (uint64_t *)leaf_code_buf.p, nullptr);
        if (rc == SHGPU_OK) rc = shgpu_copy_to_host(e, &code, leaf_code_buf.p, 8);
        check(rc, "shgpu_sh2_leaf_search (one leaf)");
        if (code == SHGPU_NOT_FOUND) throw Error(SHGPU_ERR_SEARCH_LIMIT, "ShockHash2 leaf search exhausted its tables (duplicate keys?)");
        return code;
    }

  private:
    static shgpu_build_desc desc(int variant, uint32_t leaf_size, uint64_t bucket_size, bool rotation) {
        shgpu_build_desc d{};
        d.rotation_fitting = rotation ? 1 : 0;
        d.variant = variant;
        d.leaf_size = leaf_size;
        d.bucket_size = bucket_size;
        d.mode = SHGPU_MODE_PERSISTENT;
        d.copy_keys = SHGPU_KEYS_DEVICE | SHGPU_NODES_HOST;  // keys stay on the device for the orientation pass
        return d;
    }

    Forest collect(shgpu_build *b, uint64_t /*n*/) {
        Forest f;
        f.device = std::shared_ptr<shgpu_build>(b, shgpu_build_free);
        f.nbuckets = shgpu_build_nbuckets(b);
        f.size_acc.resize(f.nbuckets + 1);
        const uint64_t nn = shgpu_build_nnodes(b);
        f.nodes.resize(nn);
        f.codes.resize(nn);
        int rc = shgpu_build_bucket_sizes(b, f.size_acc.data());
        if (rc == SHGPU_OK) rc = shgpu_build_nodes(b, f.nodes.data(), f.codes.data());
        if (rc == SHGPU_OK) rc = shgpu_build_timings(b, f.ms);
        check(rc, "shgpu_build results");
        return f;
    }

  public:
    // The retrieval input of a forest (every leaf's orientation, in the reference's order) and its one-bit
    // retrieval structure, both computed on the device (shgpu_leaf_orientation, shgpu_ribbon_build) from the build's
    // device-resident leaf-ordered keys; the structure comes back, and the input too when `keep` is set. rot_leaf:
    // the leaf size of a ROTATION_FITTING build (its full leaves' codes are x * leaf + r), else 0.
    void orientation_retrieval(int variant, const Forest &f, bool keep, std::vector<uint64_t> &input_keys,
                               std::vector<uint8_t> &input_bits, std::vector<uint32_t> &offsets,
                               std::vector<uint8_t> &seeds, std::vector<uint64_t> &words, double *ms2 = nullptr,
                               uint32_t rot_leaf = 0) {
        std::lock_guard<std::mutex> lock(mtx);
        check(ensure_host(orient_host, (f.nodes.size() + 1) * sizeof(shgpu_orient_task)), "shgpu_host_alloc (orientation tasks)");
        shgpu_orient_task *tasks = (shgpu_orient_task *)orient_host.p;
        size_t ntasks = 0;
        uint64_t total = 0;
        for (size_t i = 0; i < f.nodes.size(); i++) {
            const shgpu_node &nd = f.nodes[i];
            if (nd.kind != SHGPU_NODE_LEAF) continue;
            if (variant != SHGPU_SHOCKHASH2 && rot_leaf != 0 && nd.m == rot_leaf && nd.m >= 2)  // ShockHash.h:401-402
                tasks[ntasks++] = {total, f.codes[i] / nd.m + nd.start_seed, nd.offset, nd.m, (uint32_t)(f.codes[i] % nd.m) + 1, 0};
            else
                tasks[ntasks++] = {total, variant == SHGPU_SHOCKHASH2 ? f.codes[i] : f.codes[i] + nd.start_seed, nd.offset, nd.m, 0, 0};
            total += nd.m;
        }
        if (!f.device) throw Error(SHGPU_ERR_ARG, "leaf orientation: the forest holds no device build");
        const uint64_t *leaf_keys = nullptr;
        check(shgpu_build_device_keys(f.device.get(), &leaf_keys), "shgpu_build_device_keys");
        retrieval_from_device(variant, leaf_keys, tasks, ntasks, total, keep, input_keys, input_bits, offsets, seeds, words, ms2);
    }

    // ShockHash2Flat's threshold layers (shgpu_flat_layers) over the hashes mhcs, which stay on the device for
    // flat_leaves_retrieval. thresholds and free_counts get nbuckets = n / k entries, bumped_* up to n.
    void flat_layers(const uint64_t *mhcs, uint64_t n, uint32_t k, const uint32_t *mapping, uint32_t range,
                     std::vector<uint8_t> &thresholds, std::vector<uint32_t> &free_counts, std::vector<uint64_t> &layer_bases,
                     std::vector<uint64_t> &bumped_mhcs, std::vector<uint64_t> &bumped_index) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t nb = k ? n / k : 0;
        thresholds.assign(nb + 1, 0);
        free_counts.assign(nb + 1, 0);
        // room for every key to be bumped, left uninitialised (a zero-filled vector of n words costs more than the
        // layers themselves); the bumped keys, a few percent, are copied out
        std::unique_ptr<uint64_t[]> bm(new uint64_t[n + 1]), bi(new uint64_t[n + 1]);
        uint64_t bases[3] = {0, 0, 0}, nbumped = 0;
        uint32_t layers = 0;
        int rc = ensure(flat_mhc_buf, n * 8 + 8);
        if (rc == SHGPU_OK && n) rc = shgpu_copy_to_device(e, flat_mhc_buf.p, mhcs, n * 8);
        if (rc == SHGPU_OK)
            rc = shgpu_flat_layers(e, (const uint64_t *)flat_mhc_buf.p, n, k, mapping, range, thresholds.data(), free_counts.data(),
                                   bases, &layers, bm.get(), bi.get(), &nbumped);
        check(rc, "shgpu_flat_layers");
        layer_bases.assign(bases, bases + layers + 1);
        bumped_mhcs.assign(bm.get(), bm.get() + nbumped);
        bumped_index.assign(bi.get(), bi.get() + nbumped);
    }

    // After flat_layers: every key's bucket (shgpu_flat_bucket_keys; fb_index / fb_bucket place the bumped keys), then
    // every bucket's ShockHash2 leaf code and the retrieval structure, as sh2_leaves_retrieval, without the keys
    // leaving the device.
    void flat_leaves_retrieval(uint64_t n, uint32_t k, const uint32_t *mapping, uint32_t range,
                               const std::vector<uint64_t> &layer_bases, const std::vector<uint8_t> &thresholds,
                               const std::vector<uint64_t> &fb_index, const std::vector<uint64_t> &fb_bucket,
                               std::vector<uint64_t> &codes, bool keep, std::vector<uint64_t> &input_keys,
                               std::vector<uint8_t> &input_bits, std::vector<uint32_t> &offsets, std::vector<uint8_t> &seeds,
                               std::vector<uint64_t> &words, double *search_ms = nullptr) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t nb = k ? n / k : 0;
        int rc = ensure(leaf_keys_buf, nb * k * 8 + 8);
        if (rc == SHGPU_OK)
            rc = shgpu_flat_bucket_keys(e, (const uint64_t *)flat_mhc_buf.p, n, k, mapping, range, layer_bases.data(),
                                        (uint32_t)layer_bases.size() - 1, thresholds.data(), fb_index.data(), fb_bucket.data(),
                                        fb_index.size(), (uint64_t *)leaf_keys_buf.p);
        check(rc, "shgpu_flat_bucket_keys");
        leaves_retrieval_device(nb, k, codes, keep, input_keys, input_bits, offsets, seeds, words, search_ms);
    }

    // ShockHash2 leaves of k keys each, keys[i * k .. (i + 1) * k) (ShockHash2Flat's buckets): their codes
    // (BijectionsShockHash2::findSeed, one batched search on the device) and the retrieval structure of their
    // orientations (constructRetrieval, ShockHash2-internal.h:548-572), built as for a forest.
    void sh2_leaves_retrieval(const std::vector<uint64_t> &keys, uint32_t k, std::vector<uint64_t> &codes, bool keep,
                              std::vector<uint64_t> &input_keys, std::vector<uint8_t> &input_bits,
                              std::vector<uint32_t> &offsets, std::vector<uint8_t> &seeds, std::vector<uint64_t> &words) {
        std::lock_guard<std::mutex> lock(mtx);
        const uint64_t nb = k ? keys.size() / k : 0;
        int rc = ensure(leaf_keys_buf, keys.size() * 8 + 8);
        if (rc == SHGPU_OK && !keys.empty()) rc = shgpu_copy_to_device(e, leaf_keys_buf.p, keys.data(), keys.size() * 8);
        check(rc, "ShockHash2 leaf keys upload");
        leaves_retrieval_device(nb, k, codes, keep, input_keys, input_bits, offsets, seeds, words);
    }

  private:
    // The nb leaves of k keys in leaf_keys_buf: codes, orientation and retrieval (caller holds mtx); search_ms: the
    // batched leaf search alone (host clock around the call and its copy of the codes).
    void leaves_retrieval_device(uint64_t nb, uint32_t k, std::vector<uint64_t> &codes, bool keep,
                                 std::vector<uint64_t> &input_keys, std::vector<uint8_t> &input_bits,
                                 std::vector<uint32_t> &offsets, std::vector<uint8_t> &seeds, std::vector<uint64_t> &words,
                                 double *search_ms = nullptr) {
        std::vector<shgpu_leaf_task> lt(nb);
        for (uint64_t i = 0; i < nb; i++) lt[i] = {(uint32_t)(i * k), k, 0, 0};
        codes.assign(nb, SHGPU_NOT_FOUND);
        int rc = ensure(leaf_task_buf, nb * sizeof(shgpu_leaf_task) + 8);
        if (rc == SHGPU_OK) rc = ensure(leaf_code_buf, nb * 8 + 8);
        if (rc == SHGPU_OK && nb) rc = shgpu_copy_to_device(e, leaf_task_buf.p, lt.data(), nb * sizeof(shgpu_leaf_task));
        if (rc == SHGPU_OK) rc = shgpu_sync(e);
        const auto t0 = std::chrono::steady_clock::now();
        if (rc == SHGPU_OK && nb)
            rc = shgpu_sh2_leaf_search(e, (const uint64_t *)leaf_keys_buf.p, (const shgpu_leaf_task *)leaf_task_buf.p,
                                       (uint32_t)nb, (uint64_t *)leaf_code_buf.p, nullptr);
        if (rc == SHGPU_OK && nb) rc = shgpu_copy_to_host(e, codes.data(), leaf_code_buf.p, nb * 8);
        if (search_ms) *search_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        check(rc, "shgpu_sh2_leaf_search (buckets)");
        for (uint64_t c : codes)
            if (c == SHGPU_NOT_FOUND) throw Error(SHGPU_ERR_SEARCH_LIMIT, "ShockHash2 leaf search exhausted its tables (duplicate keys?)");
        check(ensure_host(orient_host, (nb + 1) * sizeof(shgpu_orient_task)), "shgpu_host_alloc (orientation tasks)");
        shgpu_orient_task *tasks = (shgpu_orient_task *)orient_host.p;
        for (uint64_t i = 0; i < nb; i++) tasks[i] = {i * k, codes[i], (uint32_t)(i * k), k, 0, 0};
        retrieval_from_device(SHGPU_SHOCKHASH2, (const uint64_t *)leaf_keys_buf.p, tasks, nb, nb * k, keep, input_keys,
                              input_bits, offsets, seeds, words, nullptr);
    }

    // Orientation replay of `tasks` over device keys, then the ribbon over the replay's output (caller holds mtx).
    void retrieval_from_device(int variant, const uint64_t *d_keys, const shgpu_orient_task *tasks, size_t ntasks,
                               uint64_t total, bool keep, std::vector<uint64_t> &input_keys,
                               std::vector<uint8_t> &input_bits, std::vector<uint32_t> &offsets,
                               std::vector<uint8_t> &seeds, std::vector<uint64_t> &words, double *ms2) {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = ensure(tasks_buf, ntasks * sizeof(shgpu_orient_task) + 8);
        if (rc == SHGPU_OK) rc = ensure(out_keys_buf, total * 8 + 8);
        if (rc == SHGPU_OK) rc = ensure(out_bits_buf, total + 8);
        void *d_tasks = tasks_buf.p, *d_out_keys = out_keys_buf.p, *d_out_bits = out_bits_buf.p;
        if (rc == SHGPU_OK && ntasks) rc = shgpu_copy_to_device(e, d_tasks, tasks, ntasks * sizeof(shgpu_orient_task));
        if (rc == SHGPU_OK)
            rc = shgpu_leaf_orientation(e, variant, d_keys, (const shgpu_orient_task *)d_tasks, ntasks,
                                        (uint64_t *)d_out_keys, (uint8_t *)d_out_bits);
        const auto t1 = std::chrono::steady_clock::now();
        // the structure comes back into page-locked memory (offsets, seed bytes, slot words) and is copied out
        const uint64_t shards = shgpu_ribbon_shards(total), maxw = shgpu_ribbon_max_words(total);
        const size_t off_b = ((shards + 1) * 4 + 7) / 8 * 8, seed_b = (shards + 7) / 8 * 8;  // 8-byte aligned parts
        if (rc == SHGPU_OK) rc = ensure_host(ribbon_host, off_b + seed_b + maxw * 8 + 8);
        uint64_t used = 0;
        if (rc == SHGPU_OK) {
            uint8_t *base = (uint8_t *)ribbon_host.p;
            uint32_t *h_off = (uint32_t *)base;
            uint8_t *h_seeds = base + off_b;
            uint64_t *h_words = (uint64_t *)(base + off_b + seed_b);
            rc = shgpu_ribbon_build(e, (const uint64_t *)d_out_keys, (const uint8_t *)d_out_bits, total, h_off, h_seeds,
                                    h_words, maxw, &used);
            if (rc == SHGPU_OK) {
                offsets.assign(h_off, h_off + shards + 1);
                seeds.assign(h_seeds, h_seeds + shards);
                words.assign(h_words, h_words + used);
            }
        }
        input_keys.assign(keep ? total : 0, 0);
        input_bits.assign(keep ? total : 0, 0);
        if (rc == SHGPU_OK && keep && total) rc = shgpu_copy_to_host(e, input_keys.data(), d_out_keys, total * 8);
        if (rc == SHGPU_OK && keep && total) rc = shgpu_copy_to_host(e, input_bits.data(), d_out_bits, total);
        const std::string msg = rc == SHGPU_OK ? "" : shgpu_last_error();
        if (ms2) {
            ms2[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            ms2[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        }
        if (rc != SHGPU_OK) throw Error(rc, std::string("leaf orientation / retrieval: ") + shgpu_status_string(rc) + ": " + msg);
    }
};

}  // namespace gpu
}  // namespace shockhash
