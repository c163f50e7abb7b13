/*
 * libshockhash_amd.so — MI355X (gfx950) construction engine for ShockHash's brute-force seed search.
 *
 * C ABI: plain pointers and sizes, int status codes, no exceptions across the boundary.
 * Device pointers ("d_" prefix) are HIP device allocations; `stream` is a hipStream_t or NULL for the
 * engine's own stream. Calls on one engine must not overlap (one submitting thread per engine).
 *
 * Reference seams these entry points replace (paths relative to the reference repository):
 *  - shgpu_sh1_leaf_search: the inline ShockHash1 leaf loop of ShockHash::recSplit
 *    (include/ShockHash.h:404-434), SIMDShockHash::leafLevel (include/SIMDShockHash.hpp:507-542) and
 *    BijectionsShockHash1::calculateBijection (benchmark/bijections/ShockHash1.h:13-35).
 *  - shgpu_sh2_leaf_search: the search half of
 *    size_t shockhash2construct(size_t m, std::vector<uint64_t>&, std::vector<std::pair<uint64_t,uint8_t>>&)
 *    (include/shockhash/ShockHash2-precompiled.h:9, src/ShockHash2-precompiled.cpp:49-61), i.e.
 *    BijectionsShockHash2::findSeed (include/shockhash/ShockHash2-internal.h:501-546).
 *  - shgpu_split_search: the inline split-seed loops + partitions of ShockHash::recSplit
 *    (include/ShockHash.h:451-548), ShockHash2::recSplit (include/ShockHash2.h:290-387) and
 *    SIMDShockHash::higherLevel/aggrLevel (include/SIMDShockHash.hpp:201-365).
 *  - shgpu_build_run: parallelPartition (include/shockhash/Sorter.hpp:15-36) plus every recSplit
 *    seed search of hash_gen/compute_thread (include/ShockHash.h:557-688, include/ShockHash2.h:396-498,
 *    include/SIMDShockHash.hpp:619-760); encoding stays with the caller.
 */
#ifndef SHOCKHASH_AMD_H
#define SHOCKHASH_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "../shockhash_amd/csrc/shgpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum {
    SHGPU_OK = 0,
    SHGPU_ERR_ARG = 1,               /* invalid argument */
    SHGPU_ERR_HIP = 2,               /* HIP runtime error (message in shgpu_last_error) */
    SHGPU_ERR_NOMEM = 3,             /* device allocation failed */
    SHGPU_ERR_SEARCH_LIMIT = 4,      /* a node exhausted its seed bound (duplicate keys?) */
    SHGPU_ERR_BUCKET_TOO_LARGE = 5,  /* bucket above the encodable maximum (2999 keys) */
    SHGPU_ERR_UNSUPPORTED = 6,       /* configuration the GPU engine does not implement */
    SHGPU_ERR_NO_DEVICE = 7,         /* no HIP device / extension unusable */
    SHGPU_ERR_ADMISSION = 8          /* request larger than the serving HBM budget / server stopped */
};

/* Construction variants (the three reference class templates). */
enum { SHGPU_SHOCKHASH = 0, SHGPU_SIMD_SHOCKHASH = 1, SHGPU_SHOCKHASH2 = 2 };

/* Scheduling forms of shgpu_build_run. */
enum { SHGPU_MODE_PERSISTENT = 0, SHGPU_MODE_PLAIN = 1 };

typedef struct shgpu_engine shgpu_engine;
typedef struct shgpu_build shgpu_build;

/* Schedule knobs (see shgpu_autotune). */
typedef struct shgpu_tuning {
    uint32_t sh1_seeds_per_lane;    /* 1, 2 or 4: ShockHash1 leaf window = 64 * this */
    uint32_t split_seeds_per_lane;  /* 1, 2 or 4: split window for aggregation nodes */
    uint32_t waves_per_cu;          /* persistent grid = CUs * this; 0 (the default): waves per CU and both
                                       seeds-per-lane values come from the engine's schedule class for the
                                       batch's key count (shgpu_schedule_class) */
    uint32_t flags;                 /* SHGPU_TUNE_* bits, 0 by default */
    uint64_t max_trials;            /* seed bound per split node / ShockHash1 leaf */
} shgpu_tuning;

/* shgpu_tuning.flags: search ShockHash2 leaves with the scratch-memory path only (by default leaves run on
 * the register path and fall back to it when they outgrow the registers; both give the same codes). */
#define SHGPU_TUNE_SH2_GENERAL_ONLY 1u
/* shgpu_tuning.flags: the persistent build takes every split node before any leaf (two-part queue)
 * instead of one FIFO over all nodes. */
#define SHGPU_TUNE_SPLITS_FIRST 2u
/* shgpu_tuning.flags: after a persistent build, compare the forest planned on the device with the host's
 * plan node by node (SHGPU_ERR_HIP on a difference); a check for tests, it copies every node back. */
#define SHGPU_TUNE_CHECK_PLAN 4u
/* shgpu_tuning.flags: ShockHash2 leaves above 64 keys first search with tables sized for a quarter of the
 * half-seeds they need on average, so most of them outgrow the tables and are finished by the second pass with
 * larger tables, which otherwise runs for about one leaf in 10^7; a check for tests, the codes are unchanged. */
#define SHGPU_TUNE_SH2_TIGHT_TABLES 8u

/* Last error message of the calling thread. */
const char *shgpu_last_error(void);
const char *shgpu_status_string(int status);

int shgpu_engine_create(int device, shgpu_engine **out);
void shgpu_engine_destroy(shgpu_engine *e);
/* CU count, peak engine clock (kHz) and gfx arch name of the engine's device. */
int shgpu_engine_info(shgpu_engine *e, int *cu_count, int *clock_khz, char *arch, size_t arch_len);
/* Schedule of the batched per-node entry points; shgpu_set_tuning also sets every variant's build
 * schedule. */
int shgpu_get_tuning(shgpu_engine *e, shgpu_tuning *out);
int shgpu_set_tuning(shgpu_engine *e, const shgpu_tuning *t);
/* Build schedule of one variant (SHGPU_SHOCKHASH / SHGPU_SIMD_SHOCKHASH / SHGPU_SHOCKHASH2). */
int shgpu_get_variant_tuning(shgpu_engine *e, int variant, shgpu_tuning *out);
int shgpu_set_variant_tuning(shgpu_engine *e, int variant, const shgpu_tuning *t);
/* Synchronises the engine's stream. */
int shgpu_sync(shgpu_engine *e);

/* Build schedule by size class: a build of a variant whose tuning has waves_per_cu 0 takes its waves per CU and
 * seeds per lane from the variant's first class with max_keys >= the batch's key count (the variant's largest
 * class takes larger batches). shgpu_engine_create loads the table from the JSON file named by the environment
 * variable SHGPU_SCHEDULE ("none": the built-in table), else from schedule_<arch>.json beside the library or in
 * its parent directory, else uses the built-in hand-measured table. No reference counterpart: the reference's
 * CPU build has no launch schedule. */
typedef struct shgpu_schedule_class {
    uint32_t variant;               /* SHGPU_SHOCKHASH / SHGPU_SIMD_SHOCKHASH / SHGPU_SHOCKHASH2 */
    uint32_t waves_per_cu;          /* 1..32 */
    uint32_t sh1_seeds_per_lane;    /* 1, 2 or 4 */
    uint32_t split_seeds_per_lane;  /* 1, 2 or 4 */
    uint64_t max_keys;              /* batches of at most this many keys */
    double tuned_ms;                /* search kernel ms of the tuning sample under this class (0: not tuned) */
    uint64_t sample_keys;           /* keys of that sample (0: not tuned) */
} shgpu_schedule_class;
/* Copies the engine's classes (sorted by variant, then max_keys); *count = their number. */
int shgpu_schedule_get(shgpu_engine *e, shgpu_schedule_class *out, uint32_t capacity, uint32_t *count);
/* Replaces the classes of every variant named in `classes`. */
int shgpu_schedule_set(shgpu_engine *e, const shgpu_schedule_class *classes, uint32_t count);
int shgpu_schedule_save(shgpu_engine *e, const char *path);
int shgpu_schedule_load(shgpu_engine *e, const char *path);
/* Where the engine's table came from: a file path or "built-in". */
const char *shgpu_schedule_source(shgpu_engine *e);
/* Parses the JSON form without a device (what shgpu_schedule_load reads). */
int shgpu_schedule_parse(const char *json, shgpu_schedule_class *out, uint32_t capacity, uint32_t *count);

/* Free and total bytes of the engine's device (hipMemGetInfo), e.g. to size a serving HBM budget. */
int shgpu_device_memory(shgpu_engine *e, size_t *free_bytes, size_t *total_bytes);
/* Device memory helpers (ordered on the engine's stream; the copies are synchronous). */
int shgpu_device_alloc(shgpu_engine *e, size_t bytes, void **d_out);
int shgpu_device_free(shgpu_engine *e, void *d);
/* Page-locked host memory (hipHostMalloc), from which shgpu_copy_to_device reads by DMA without staging; freed
 * with shgpu_host_free. */
int shgpu_host_alloc(shgpu_engine *e, size_t bytes, void **h_out);
int shgpu_host_free(shgpu_engine *e, void *h);
int shgpu_copy_to_device(shgpu_engine *e, void *d_dst, const void *h_src, size_t bytes);
int shgpu_copy_to_host(shgpu_engine *e, void *h_dst, const void *d_src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (k_copy_u64); bytes must be a multiple of 8. Used as
 * the known-size reference when HBM traffic counters are calibrated. Synchronous. */
int shgpu_device_copy(shgpu_engine *e, void *d_dst, const void *d_src, size_t bytes);

/* Batched ShockHash1 leaf search: d_seeds[i] = smallest seed >= start_seed of task i
 * (SHGPU_NOT_FOUND if none below start_seed + max_trials). m <= 64. */
int shgpu_sh1_leaf_search(shgpu_engine *e, const uint64_t *d_keys, const shgpu_leaf_task *d_tasks, uint32_t n_tasks,
                          uint64_t *d_seeds, void *stream);

/* Batched ShockHash2 leaf search: d_codes[i] = the findSeed code of task i (start_seed and
 * max_trials are ignored). m <= 128; reads the tasks back to size the search tables by the largest leaf and
 * synchronises when a leaf exceeds 64 keys. SHGPU_NOT_FOUND when a leaf's search exhausts its tables (repeated
 * keys). */
int shgpu_sh2_leaf_search(shgpu_engine *e, const uint64_t *d_keys, const shgpu_leaf_task *d_tasks, uint32_t n_tasks,
                          uint64_t *d_codes, void *stream);

/* Batched split-seed search; each task's keys are partitioned in place by the found seed.
 * d_seeds[i] = found seed (absolute). Tasks must not overlap. fanout <= 4, m <= 65535. */
int shgpu_split_search(shgpu_engine *e, uint64_t *d_keys, const shgpu_split_task *d_tasks, uint32_t n_tasks,
                       uint64_t *d_seeds, void *stream);

/* shgpu_split_search without any allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its partition staging and work counter live in a fixed per-device scratch array of the library
 * (SHGPU_NOALLOC_MAX_KEYS keys, 8 MiB per GPU, held until the process ends); it calls no hipMalloc/hipFree and
 * does not synchronise. n_keys: length of the d_keys array (at most SHGPU_NOALLOC_MAX_KEYS, else SHGPU_ERR_ARG).
 * The tasks are validated on the device instead of the host: a task that shgpu_split_search would reject (or one
 * reaching past n_keys) leaves its keys untouched and gets d_seeds[i] = SHGPU_NOT_FOUND, like a search that
 * exhausts its bound. Results equal shgpu_split_search's. Calls on one GPU share the scratch: calls that overlap
 * on one GPU (other streams, other engines, other threads) race and must be serialised by the caller. */
#define SHGPU_NOALLOC_MAX_KEYS (1u << 20)
int shgpu_split_search_noalloc(shgpu_engine *e, uint64_t *d_keys, uint64_t n_keys, const shgpu_split_task *d_tasks,
                               uint32_t n_tasks, uint64_t *d_seeds, void *stream);

/* Whole-construction seed search. d_hashes: n 128-bit fingerprints as (first, second) pairs. Sorts by
 * first, buckets by remap128(first, nbuckets), builds the RecSplit tree of every bucket and searches
 * every node's seed. Synchronous; the result owns copies of everything the encoder needs. */
typedef struct shgpu_build_desc {
    int variant;           /* SHGPU_SHOCKHASH / SHGPU_SIMD_SHOCKHASH / SHGPU_SHOCKHASH2 */
    uint32_t leaf_size;    /* template LEAF_SIZE */
    uint64_t bucket_size;  /* constructor bucket_size */
    int mode;              /* SHGPU_MODE_PERSISTENT or SHGPU_MODE_PLAIN */
    uint32_t plain_bucket_limit; /* plain mode: build only the first N buckets (0 = all) */
    int copy_keys;         /* SHGPU_KEYS_HOST: keep a host copy of the keys in final leaf order;
                              SHGPU_KEYS_DEVICE: keep a device copy (shgpu_build_device_keys);
                              SHGPU_NODES_HOST: copy the device-planned nodes back with the codes (otherwise
                              shgpu_build_nodes plans them again on the host on first request); any OR */
    int rotation_fitting;  /* ShockHash / SIMDShockHash ROTATION_FITTING (include/ShockHash.h:316-403,
                              include/SIMDShockHash.hpp:419-506): leaves of exactly leaf_size keys get the code
                              x * leaf_size + r of the rotation search; 0: off (ShockHash2 ignores it) */
} shgpu_build_desc;

#define SHGPU_KEYS_HOST 1
#define SHGPU_KEYS_DEVICE 2
#define SHGPU_NODES_HOST 4

int shgpu_build_run(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                    shgpu_build **out);
/* Autotuner for one explicit schedule: the sweep of shgpu_autotune_class (below); the winner becomes the
 * variant's tuning (waves_per_cu set, so the size classes no longer apply to the variant) and is returned.
 * Results are identical for every schedule; only the time differs. d_hashes: device fingerprints. */
int shgpu_autotune(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                   shgpu_tuning *best, double *best_ms);
/* Autotunes one class: builds the sample (device fingerprints) under candidate schedules in two stages, waves
 * per CU in {4, 8, 12, 16, 24} at the default seeds per lane, then every pair of leaf seeds per lane {1, 2, 4}
 * (ShockHash1 leaves only) and split seeds per lane {1, 2, 4} at the fastest wave count, twice each (once when a
 * build takes over a second); checks that every candidate gives the same codes (SHGPU_ERR_HIP otherwise), and
 * stores the one with the shortest search kernel as the variant's class for batches up to max_keys keys. */
int shgpu_autotune_class(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *d_hashes, uint64_t n,
                         uint64_t max_keys, shgpu_schedule_class *chosen);
/* Same with the fingerprints in host memory (staged to the device by the engine). */
int shgpu_build_run_host(shgpu_engine *e, const shgpu_build_desc *desc, const uint64_t *h_hashes, uint64_t n,
                         shgpu_build **out);
void shgpu_build_free(shgpu_build *b);
uint64_t shgpu_build_nbuckets(const shgpu_build *b);
uint64_t shgpu_build_nnodes(const shgpu_build *b);
/* size_acc: nbuckets + 1 cumulative bucket sizes. */
int shgpu_build_bucket_sizes(const shgpu_build *b, uint64_t *size_acc);
/* nodes (may be NULL) and codes (code = found seed - start_seed), nnodes each, DFS pre-order. */
int shgpu_build_nodes(const shgpu_build *b, shgpu_node *nodes, uint64_t *codes);
/* Keys (fingerprint .second) in final leaf order, n of them (needs SHGPU_KEYS_HOST in desc->copy_keys). */
int shgpu_build_keys(const shgpu_build *b, uint64_t *keys);
/* The same keys in device memory owned by the build, valid until shgpu_build_free (needs SHGPU_KEYS_DEVICE).
 * The input of shgpu_leaf_orientation, which the reference fills on the host (leafKeys, ShockHash2.h:273;
 * ribbonInput, ShockHash.h:425-432): the classes hand it over without a round trip through host memory. */
int shgpu_build_device_keys(const shgpu_build *b, const uint64_t **d_keys);
/* Host-only: the forest shgpu_build_run would search for these bucket sizes (DFS pre-order). Call with
 * out = NULL to get the count. Needs no GPU. */

/* Retrieval input of leaves: replays each leaf's cuckoo insertions in leaf key order (the orientation that
 * fixes which of its two cells a key occupies) and writes the (key, bit) entries in the reference's order:
 * ShockHash / SIMDShockHash leaves cell by cell with bit = cell >= m/2 (include/ShockHash.h:425-432,
 * include/SIMDShockHash.hpp:536-540), ShockHash2 leaves key by key with bit = the key sits in its second cell
 * (include/shockhash/ShockHash2-internal.h:548-572). variant: SHGPU_*; d_keys: leaf-ordered keys; d_tasks: n
 * shgpu_orient_task; outputs sized by the tasks' out_offset + m. All device memory; SHGPU_ERR_ARG when a
 * task's keys cannot be placed with its seed. */
int shgpu_leaf_orientation(shgpu_engine *e, int variant, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                           uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits);
/* Fingerprints of string keys: SpookyHash V2 128-bit with seed 0 of each key, as {first, second} = {hash2, hash1}
 * (first_hash, include/ShockHash.h:129; the string constructors, include/ShockHash.h:159-185,
 * include/SIMDShockHash.hpp:138, include/ShockHash2.h:140). d_data: the keys' bytes, d_offsets: n + 1 offsets
 * (key i is d_data[d_offsets[i], d_offsets[i + 1])), d_hashes: 2n u64; all device memory. Runs on the engine's
 * stream (ordered before a following shgpu_build_run). */
int shgpu_hash_strings(shgpu_engine *e, const uint8_t *d_data, const uint64_t *d_offsets, uint64_t n, uint64_t *d_hashes);
/* One-bit retrieval structure of the leaf orientations, built on the device: a sharded ribbon. Replaces the
 * construction of the reference's SimpleRibbon<1> retrieval (include/ShockHash.h:657-658,
 * include/ShockHash2.h:491-497; the ribbon library itself is not vendored) at about 1.07 bits per key.
 * Layout (shockhash::RibbonRetrieval answers queries on it): S = shgpu_ribbon_shards(n) shards; key k belongs to
 * shard fastrange64(h, S) with h = remix(k + 0x632be59bd9b4e019). Shard s has n_s keys and
 * m_s = n_s + ceil(n_s / 20) one-bit slots (at least 64; none when empty) at slot offset offsets[s] (offsets[S] is
 * the total). With its seed byte d: g = remix(h + (d + 1) * 0x9E3779B97F4A7C15),
 * start = fastrange64(g, m_s - 63), row = remix(g ^ 0xd6e8feb86659fd93) | 1, and the key's bit is the parity of
 * row AND slots offsets[s] + start .. + 63 (bit j of row covers slot start + j; slot c is bit c % 64 of word c / 64).
 * d_keys: n keys, d_bits: n bytes whose bit 0 is the key's bit, both in device memory. Writes offsets (S + 1),
 * seeds (S) and the packed words (*words_used of them, including one word of padding, at most
 * shgpu_ribbon_max_words(n)) to host memory. SHGPU_ERR_SEARCH_LIMIT when a shard has no consistent seed byte
 * (repeated keys with different bits). */
uint64_t shgpu_ribbon_shards(uint64_t n);
uint64_t shgpu_ribbon_max_words(uint64_t n);
int shgpu_ribbon_build(shgpu_engine *e, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint32_t *h_offsets,
                       uint8_t *h_seeds, uint64_t *h_words, uint64_t words_capacity, uint64_t *words_used);

/* ShockHash2Flat's bucketing on the device, for n (< 2^31) keys' 64-bit hashes d_mhcs (device memory) and buckets of
 * k keys; mapping holds the `range` (2..64) threshold values of THRESHOLD_MAPPING, the last 2^32 - 1. It replaces the
 * host loops of ShockHash2Flat::ShockHash2Flat (include/ShockHash2Flat.h:70-124: the threshold layers with
 * bumping, flushBucket :204-233; :148-157 with evaluateKPerfect :280-302: every key's bucket).
 * shgpu_flat_layers: per bucket of the layers (the first layer_bases[*layers] of n / k buckets) its threshold index
 * (thresholds) and its free positions, k minus the keys it keeps (free_counts); the layer bases (*layers + 1 of at
 * most 3 entries); the keys bumped from every layer, as their last layer's hash (bumped_mhcs) and input index
 * (bumped_index), *n_bumped of them (capacity n each), in the reference's order. All outputs are host memory. The
 * caller places the bumped keys (the reference's fallback function over their decimal strings and free positions).
 * SHGPU_ERR_ARG when a bucket cannot take a threshold (the reference throws). */
int shgpu_flat_layers(shgpu_engine *e, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                      uint32_t range, uint8_t *thresholds, uint32_t *free_counts, uint64_t *layer_bases, uint32_t *layers,
                      uint64_t *bumped_mhcs, uint64_t *bumped_index, uint64_t *n_bumped);
/* shgpu_flat_bucket_keys: every key's bucket, the first layer whose stored threshold admits it, else
 * fb_bucket[j] for the key of input index fb_index[j] (values from (n / k) * k on, below (n / k) * k + n, name the
 * last, partial bucket, as evaluateKPerfect returns them); each of the n / k buckets must receive exactly k keys (SHGPU_ERR_ARG
 * otherwise), and bucket b's keys, in input order, go to d_bucket_keys[b * k .. (b + 1) * k) (device memory).
 * layer_bases, thresholds, fb_index and fb_bucket are host memory as shgpu_flat_layers returns them. */
int shgpu_flat_bucket_keys(shgpu_engine *e, const uint64_t *d_mhcs, uint64_t n, uint32_t k, const uint32_t *mapping,
                           uint32_t range, const uint64_t *layer_bases, uint32_t layers, const uint8_t *thresholds,
                           const uint64_t *fb_index, const uint64_t *fb_bucket, uint64_t n_fb, uint64_t *d_bucket_keys);

int shgpu_plan_forest(int variant, uint32_t leaf_size, const uint64_t *size_acc, uint64_t nbuckets, shgpu_node *out,
                      uint64_t capacity, uint64_t *count);
/* Milliseconds spent in the build's phases: [0] sort + bucketing, [1] tree setup, [2] seed search
 * (host clock, launch to results on the host), [3] the persistent search kernel alone (HIP events on
 * the engine's stream; 0 in plain mode). */
int shgpu_build_timings(const shgpu_build *b, double *ms4);

/* ---- Serving: continuous batching of construction requests.
 * A server owns `slots` concurrent builds (one stream, buffer set and host worker each; each slot's
 * persistent kernel uses 1/slots of the device). Requests are admitted in arrival order as soon as a
 * slot is free and their device footprint fits the HBM budget: a free slot takes every waiting request
 * that fits (up to the batch limits) and runs their trees in one persistent kernel, so work of requests
 * that arrive while others run overlaps on the device. Results equal a lone shgpu_build_run of the same
 * keys. */
typedef struct shgpu_server shgpu_server;
typedef struct shgpu_server_config {
    int variant;
    uint32_t leaf_size;
    uint64_t bucket_size;
    uint32_t slots;             /* concurrent builds, 1..16 */
    uint64_t hbm_budget_bytes;  /* device bytes the server may hold for running requests */
    int copy_keys;              /* results keep the leaf-ordered keys */
    uint64_t max_batch_keys;    /* a slot builds the waiting requests together up to this many keys (0: 8M) */
    uint32_t max_batch_requests;/* ... and up to this many requests (0: 256) */
} shgpu_server_config;

int shgpu_server_create(shgpu_engine *e, const shgpu_server_config *cfg, shgpu_server **out);
/* Waits for running requests; queued ones are dropped. Results not taken are freed. */
void shgpu_server_destroy(shgpu_server *s);
/* h_hashes (n fingerprints, host memory) must stay valid until the request is taken.
 * SHGPU_ERR_ADMISSION if the request alone exceeds the budget. */
int shgpu_server_submit(shgpu_server *s, const uint64_t *h_hashes, uint64_t n, uint64_t *id);
/* The same for fingerprints already resident in device memory (read only, valid until taken). */
int shgpu_server_submit_device(shgpu_server *s, const uint64_t *d_hashes, uint64_t n, uint64_t *id);
/* Blocks until the request finished (timeout_s < 0: no limit); status 2 = done, < 0 = -(error code). */
int shgpu_server_wait(shgpu_server *s, uint64_t id, double timeout_s, int *status);
/* Removes a finished request and hands over its result; latency = submit to completion, queue = submit
 * to admission. */
int shgpu_server_take(shgpu_server *s, uint64_t id, shgpu_build **out, double *latency_ms, double *queue_ms);
/* Device bytes the server reserves for a request of n keys while it runs. Besides these, a ShockHash2
 * server holds a fixed per-slot leaf scratch for its lifetime; it counts against the budget from
 * creation (shgpu_server_create fails with SHGPU_ERR_ADMISSION when it alone exceeds the budget). */
int shgpu_server_request_bytes(const shgpu_server *s, uint64_t n, uint64_t *bytes);
/* reserved_bytes = fixed scratch + reservations of running requests. */
int shgpu_server_info(shgpu_server *s, uint64_t *queued, uint64_t *running, uint64_t *reserved_bytes,
                      uint64_t *peak_reserved_bytes);

#ifdef __cplusplus
}
#endif
#endif
