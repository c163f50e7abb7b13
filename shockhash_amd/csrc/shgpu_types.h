/* Plain-C task records shared by the host side of libshockhash_amd.so and its kernels. */
#ifndef SHGPU_TYPES_H
#define SHGPU_TYPES_H
#include <stdint.h>

#define SHGPU_NOT_FOUND (~(uint64_t)0)

/* Layout constants of the sharded ribbon retrieval (shgpu_ribbon_build; queried by shockhash::RibbonRetrieval). */
#define SHGPU_RIBBON_WIDTH 64u                           /* slots one key's row covers */
#define SHGPU_RIBBON_SHARD_KEYS 2048u                    /* keys per shard on average */
#define SHGPU_RIBBON_SALT_SHARD 0x632be59bd9b4e019ull
#define SHGPU_RIBBON_SALT_ROW 0xd6e8feb86659fd93ull
#define SHGPU_RIBBON_SEED_STEP 0x9E3779B97F4A7C15ull

/* One leaf seed search: keys[offset .. offset+m), seeds start_seed, start_seed+1, ...
 * (at most max_trials of them). */
typedef struct shgpu_leaf_task {
    uint32_t offset;
    uint32_t m;
    uint64_t start_seed;
    uint64_t max_trials;
} shgpu_leaf_task;

/* Orientation of one leaf (retrieval input): keys[key_offset .. key_offset+m) in leaf order, the leaf's seed
 * (ShockHash1 / SIMDShockHash: code + start seed; ShockHash2: the pair code) and where its m retrieval entries
 * go in the output. */
typedef struct shgpu_orient_task {
    uint64_t out_offset;
    uint64_t seed;
    uint32_t key_offset;
    uint32_t m;
    uint32_t rotation;  /* ShockHash1 / SIMDShockHash rotation-fitted leaf: r + 1 (seed = x); 0: plain leaf */
    uint32_t pad;
} shgpu_orient_task;

/* Hash flavour of a split search. */
enum { SHGPU_SPLIT_SUX = 0, SHGPU_SPLIT_SIMD = 1 };

/* One split-seed search. kind 0: fanout 2, part = (h >= unit); kind 1/2: part = h / unit,
 * fanout = ceil(m / unit). */
typedef struct shgpu_split_task {
    uint32_t offset;
    uint32_t m;
    uint64_t start_seed;
    uint64_t max_trials;
    uint32_t unit;
    uint8_t kind;
    uint8_t hash;
    uint16_t pad;
} shgpu_split_task;

/* RecSplit tree node, one per node with m >= 2, numbered in depth-first pre-order inside each
 * bucket and buckets in order (the order the Golomb-Rice codes are written). */
enum { SHGPU_NODE_SPLIT2 = 0, SHGPU_NODE_AGGR_UPPER = 1, SHGPU_NODE_AGGR_LOWER = 2, SHGPU_NODE_LEAF = 3 };
typedef struct shgpu_node {
    uint32_t offset;      /* first key of the node in the bucket-ordered key array */
    uint16_t m;           /* keys in the node */
    uint8_t kind;         /* SHGPU_NODE_* */
    uint8_t n_children;   /* children that are nodes (size >= 2) */
    uint32_t unit;        /* split point (kind 0) or part size (kinds 1, 2) */
    uint32_t subtree;     /* nodes in this subtree including itself */
    uint64_t start_seed;  /* first seed tried; the stored code is seed - start_seed */
    uint64_t max_trials;
} shgpu_node;

#endif
