// Types, constants and launchers shared by the engine's host code (engine.hip) and its kernels, which are
// compiled in translation units of their own: the persistent build per variant (build_sh1.hip,
// build_simd.hip, build_sh2.hip) and the batched per-node searches (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/shockhash_amd.h"
#include "device_common.h"
#include "search_sh2.h"

namespace shgpu {

constexpr uint32_t EMPTY_SLOT = 0xFFFFFFFFu;

// Queue control words of a persistent build (BuildArgs::ctrl, written by the forest plan). Counters that
// many waves update at once sit in 128-byte lines of their own, the read-mostly words share one line.
constexpr uint32_t CTRL_LINE = 32;
constexpr uint32_t C_HEAD = 0;                   // queue head (over both parts of a two-part queue)
constexpr uint32_t C_SPLIT_TAIL = 1 * CTRL_LINE;  // split (or FIFO) tail
constexpr uint32_t C_FINISHED = 2 * CTRL_LINE;    // nodes finished
constexpr uint32_t C_LEAF_TAIL = 3 * CTRL_LINE;   // two-part queue: leaf tail
constexpr uint32_t C_STARTED = 4 * CTRL_LINE;     // windowed nodes started
constexpr uint32_t C_ERROR = 5 * CTRL_LINE;       // set when a wait is declared stuck
constexpr uint32_t C_NODES = C_ERROR + 1;         // nodes of the forest
constexpr uint32_t C_WINDOWED = C_ERROR + 2;      // windowed nodes of the forest
constexpr uint32_t CTRL_WORDS = 6 * CTRL_LINE;
constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / 64;

// Next task index for the whole wave (no divergent branch, see wave_fetch_add).
__device__ __forceinline__ uint32_t grab(uint32_t *counter) { return wave_fetch_add(counter, 1u); }

// Per-node search state of a build (windows of seeds shared by every wave that works on the node).
// gd: bit 63 set once a qualifying seed is known (GD_FOUND), bits 32..62 windows handed out, bits 0..31
// windows completed. One read-modify-write on it both completes a window and hands out the next.
constexpr unsigned long long GD_FOUND = 1ull << 63, GD_HANDED = 1ull << 32;
// Windows a node's search covers at most. The handed-out count has 31 bits; past the cap, every wave that joins
// a node still takes (and completes unsearched) one window, so the cap leaves 2^24 windows of margin.
constexpr uint64_t WINDOW_CAP = 0x7F000000ull;
__device__ __forceinline__ uint64_t gd_handed(unsigned long long gd) { return (gd >> 32) & 0x7FFFFFFFull; }
__device__ __forceinline__ uint64_t gd_completed(unsigned long long gd) { return gd & 0xFFFFFFFFull; }

struct NodeState {
    unsigned long long gd;    // GD_FOUND | windows handed out << 32 | windows completed
    unsigned long long best;  // smallest qualifying seed offset found so far (SHGPU_NOT_FOUND: none)
    uint32_t finalized;       // number of waves that claimed finalisation (the first one does it)
    uint32_t help_cap;        // waves that may help the node's owner at once (the plan's)
    uint32_t helpers;         // waves helping it now
    uint32_t help_all;        // windows completed after which any number of waves may help
};

struct BuildArgs {
    uint64_t *keys;
    uint64_t *tmp;
    const shgpu_node *nodes;
    uint64_t *codes;
    uint32_t *queue;
    uint32_t *ctrl;  // queue control words (C_*)
    NodeState *state;
    uint32_t *help;  // per wave slot: node whose windows the slot's wave searches, or EMPTY_SLOT
    uint32_t total;  // plain form: nodes (the persistent kernel reads ctrl[C_NODES])
    uint32_t nslots;
    uint8_t *scratch;          // ShockHash2 leaves: sh2_wave_bytes per wave slot
    uint64_t sh2_wave_bytes;
    Sh2Caps sh2_caps;          // table capacities of leaves above 64 keys
    int split_hash;
    uint32_t flags;     // shgpu_tuning.flags of the variant
    uint32_t rot_leaf;  // ROTATION_FITTING: leaves of exactly this many keys are rotation-fitted (0: none)
};

// A ShockHash1 / SIMDShockHash leaf searched by rotation fitting: its window results are (seed offset << 6) | r
// and its code is offset * m + r (include/ShockHash.h:401-402).
__host__ __device__ __forceinline__ bool rotation_leaf(const BuildArgs &a, uint32_t kind, uint32_t m) {
    return a.rot_leaf != 0 && kind == SHGPU_NODE_LEAF && m == a.rot_leaf && m >= 2;
}

// Queue layout: one FIFO over all nodes, or a split part [0, qsplit) and a leaf part [qsplit, total).
__host__ __device__ __forceinline__ bool two_part_queue(uint32_t flags) { return (flags & SHGPU_TUNE_SPLITS_FIRST) != 0; }

// Persistent build of one variant (k_build, grid x BLOCK threads) with sl / ss seeds per lane for leaf / split
// windows (1, 2 or 4; ShockHash2 leaves ignore sl), and the plain form's launch for node nid whose keys are in
// key buffer buf (k_node: 0 = keys, 1 = tmp; the buffer alternates with the node's depth). wide:
// ShockHash2 leaves of more than 64 keys occur (the kernels then carry the two-keys-per-lane search).
template <int VARIANT>
hipError_t launch_build_variant(uint32_t sl, uint32_t ss, bool wide, const BuildArgs &a, int grid, hipStream_t s);
template <int VARIANT>
hipError_t launch_node_variant(bool wide, const BuildArgs &a, uint32_t nid, uint32_t buf, uint32_t waves, hipStream_t s);

// Batched per-node searches (batch.hip); every call returns the launch's status.
hipError_t launch_sh1_batch(uint32_t seeds_per_lane, int grid, hipStream_t s, const uint64_t *keys,
                            const shgpu_leaf_task *tasks, uint32_t n, uint64_t *out, uint32_t *ctr);
// wide: some task has more than 64 keys; each wave's scratch is wave_bytes, caps sizes the wide tables
hipError_t launch_sh2_batch(bool wide, int grid, hipStream_t s, const uint64_t *keys, const shgpu_leaf_task *tasks, uint32_t n,
                            uint64_t *out, uint32_t *ctr, uint8_t *scratch, uint64_t wave_bytes, Sh2Caps caps,
                            uint32_t flags);
// check: validate every task on the device against n_keys (the no-allocation entry point)
hipError_t launch_split_batch(bool check, uint32_t seeds_per_lane, int grid, hipStream_t s, uint64_t *keys, uint64_t *tmp,
                              const shgpu_split_task *tasks, uint32_t n, uint64_t *out, uint32_t *ctr, uint64_t n_keys);

}  // namespace shgpu
