// Persistent construction kernel (k_build) and the plain one-launch-per-node form (k_node): the window
// protocol shared by split nodes and ShockHash1 leaves, the help registry, and ShockHash2 leaves searched by
// one wave. Included by one translation unit per variant, which instantiates that variant's launchers.
#pragma once
#include <type_traits>

#include "build_common.h"
#include "search_sh1.h"
#include "search_sh2.h"
#include "search_split.h"

namespace shgpu {

__device__ __forceinline__ unsigned long long wave_fetch_add64(unsigned long long *p, unsigned long long add) {
    const unsigned long long old = atomicAdd(p, __lane_id() == 0 ? add : 0ull);
    return uniform64(old);
}

// Keys live in two buffers of the same indexing, a.keys and a.tmp: a split node partitions its keys from its
// buffer into the other one, where its children find them, so no partition is copied back. Roots start in
// a.keys; a leaf found in a.tmp copies its keys to a.keys when it is finished, so the leaf-ordered keys end up
// in a.keys. Queue and help-registry entries carry a node id and its buffer (bit 31).
constexpr uint32_t ENTRY_BUF = 1u << 31;
// Queue slot of a child that its parent's finalising wave runs itself (continuation): taken and passed over.
constexpr uint32_t SKIP_SLOT = 0xFFFFFFFEu;
__device__ __forceinline__ uint64_t *key_buf(const BuildArgs &a, uint32_t buf) { return buf ? a.tmp : a.keys; }
__device__ __forceinline__ void copy_keys(const uint64_t *src, uint64_t *dst, uint32_t m) {
    for (uint32_t i = lane_id(); i < m; i += 64) dst[i] = src[i];
}

// The value that decides a node's code is read with a read-modify-write (atomicOr 0), which resolves at
// the coherence point.
__device__ __forceinline__ uint64_t read_best(NodeState *st) {
    return uniform64(atomicOr(&st->best, 0ull));
}


template <int SL, int SS>
__device__ __forceinline__ uint32_t window_seeds(uint32_t kind) {
    return kind == SHGPU_NODE_LEAF ? 64u * SL : kind == SHGPU_NODE_AGGR_LOWER ? 64u * SS : 64u;
}

// Stores the node's code; for a split node also partitions its keys by the found seed and (in the
// persistent form) publishes the children. An exhausted search still publishes the children so that
// the forest drains; the host then reports the error. Returns the queue entry of the child the calling wave
// runs next itself (its first, largest child; its queue slot gets SKIP_SLOT), or EMPTY_SLOT.
template <int VARIANT>
__device__ __forceinline__ uint32_t finalize_node(const BuildArgs &a, uint32_t nid, uint32_t buf, uint64_t off, bool persistent) {
    uint32_t cont = EMPTY_SLOT;
    const shgpu_node nd = a.nodes[nid];
    const uint32_t o = uniform(nd.offset), m = uniform(nd.m), kind = uniform(nd.kind);
    if (VARIANT != SHGPU_SHOCKHASH2 && rotation_leaf(a, kind, m) && off != SHGPU_NOT_FOUND) off = (off >> 6) * m + (off & 63);
    const uint64_t *src = key_buf(a, buf) + o;
    if (kind != SHGPU_NODE_LEAF) {
        const SplitShapeDev sh = make_split_shape(m, uniform(nd.unit), kind, a.split_hash);
        uint64_t *dst = key_buf(a, buf ^ 1u) + o;
        // an exhausted search leaves its children the node's keys as they are
        if (off != SHGPU_NOT_FOUND) split_partition_to(src, dst, sh, uniform64(nd.start_seed) + off);
        else copy_keys(src, dst, m);
        if (!buf) {
            // a part of one key is no node: its key is in leaf order already and goes to a.keys now
            const uint32_t p = lane_id(), unit = uniform(nd.unit);
            const uint32_t size = p + 1 < sh.fanout ? unit : m - (sh.fanout - 1) * unit;
            if (p < sh.fanout && size == 1) a.keys[o + p * unit] = dst[p * unit];
        }
        if (persistent) {
            __threadfence();  // children's keys visible device-wide before they are queued
            const uint32_t nc = uniform(nd.n_children);
            // every lane computes the same child ids and stores the same slots (no divergence); split
            // children go to the split part of the queue, leaves to the leaf part
            uint32_t child = nid + 1;
            for (uint32_t c = 0; c < nc; c++) {
                const shgpu_node cn = a.nodes[child];
                const bool to_leaf_part = two_part_queue(a.flags) && uniform(cn.kind) == SHGPU_NODE_LEAF;
                const uint32_t pos = wave_fetch_add(&a.ctrl[to_leaf_part ? C_LEAF_TAIL : C_SPLIT_TAIL], 1u);
                const uint32_t entry = child | (buf ? 0u : ENTRY_BUF);
                const bool keep = c == 0;
                if (keep) cont = entry;
                __hip_atomic_store(&a.queue[pos], keep ? SKIP_SLOT : entry, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                child = uniform(child + cn.subtree);
            }
        }
    } else if (buf) {
        copy_keys(src, a.keys + o, m);
    }
    a.codes[nid] = off;  // read by the host after the kernel ends
    if (persistent) wave_fetch_add(&a.ctrl[C_FINISHED], 1u);
    return cont;
}

// Searches windows of a split node or ShockHash1 leaf until its smallest qualifying seed is known.
// Windows are handed out in increasing order. A wave holds one window at a time; one atomic on the
// node's gd word completes it and, unless the search is over for this wave, hands out the next, and
// its return value says whether a qualifying seed is already known (then the window just taken is
// completed without being searched). A node is finalised by a wave whose completion leaves no window
// outstanding, with a best seed that lies inside the windows completed by then: the best seed is read
// after the count, so it may come from a later window whose predecessors are still being searched,
// and only a seed inside the complete windows is final (otherwise the wave that completes the last
// outstanding window sees the counts equal again and decides then). Every window below the winner has
// been searched, so the result is the sequential minimum.
// Keys are staged in the wave's LDS once per call (leaves, and split nodes of up to STAGE_KEYS keys).
// Returns the continuation of finalize_node when this wave finalised the node, else EMPTY_SLOT.
template <int VARIANT, int SL, int SS>
__device__ __forceinline__ uint32_t work_windows(const BuildArgs &a, uint32_t nid, uint32_t buf, PfLds &L, bool persistent,
                                                 uint64_t t_start) {
    uint32_t cont = EMPTY_SLOT;
    const shgpu_node nd = a.nodes[nid];
    const uint64_t *keys = key_buf(a, buf) + uniform(nd.offset);
    const uint32_t o = uniform(nd.offset), m = uniform(nd.m), kind = uniform(nd.kind);
    NodeState *st = &a.state[nid];
    const uint64_t W = window_seeds<SL, SS>(kind);
    // The window count has 31 bits: a bound beyond WINDOW_CAP windows ends the search at the cap, and
    // the finish rule below uses the same limit, so a node that fails is finalised as not found.
    const uint64_t start = uniform64(nd.start_seed), lim = min(uniform64(nd.max_trials), WINDOW_CAP * W);
    // rotation-fitted leaves report (seed offset << 6) | r: compare their seed offsets with the windows
    const bool rot = VARIANT != SHGPU_SHOCKHASH2 && rotation_leaf(a, kind, m);
    const uint32_t rsh = rot ? 6u : 0u;
    constexpr int SPLIT_HASH = VARIANT == SHGPU_SIMD_SHOCKHASH ? SHGPU_SPLIT_SIMD : SHGPU_SPLIT_SUX;
    SplitShapeDev sh;
    if (kind != SHGPU_NODE_LEAF) sh = make_split_shape(m, uniform(nd.unit), kind, a.split_hash);
    unsigned long long old = wave_fetch_add64(&st->gd, GD_HANDED);
    bool found = (old & GD_FOUND) != 0;
    uint64_t w = gd_handed(old);
    // stage the keys once (after the first window is taken: a node already solved costs no loads)
    uint64_t kv = 0;
    bool staged = false;
    if (!found) {
        if (kind == SHGPU_NODE_LEAF) {
            if constexpr (VARIANT != SHGPU_SHOCKHASH2)
                if (m >= 2) kv = sh1_stage(keys, m, L);
        } else if (m <= STAGE_KEYS) {
            split_stage(keys, m, L.stage);
            staged = true;
        }
    }
    // A split node's children wait for it: its windows run at the highest issue priority, so leaf searches on
    // the same SIMD do not slow it down (they rise with age, see age_priority).
    const bool split = kind != SHGPU_NODE_LEAF;
    for (;;) {
        if (split) __builtin_amdgcn_s_setprio(3);
        else age_priority(t_start);
        const uint64_t woff = w * W;
        const bool exhausted = woff >= lim;
        uint64_t r = SHGPU_NOT_FOUND;
        if (!exhausted && !found) {
            if (kind == SHGPU_NODE_LEAF) {
                if constexpr (VARIANT != SHGPU_SHOCKHASH2)  // ShockHash2 leaves never come here
                    r = m < 2 ? (woff == 0 ? 0 : SHGPU_NOT_FOUND)
                              : rot ? sh1_rot_window_staged<SL>(kv, m, start, woff, L) : sh1_window_staged<SL>(kv, m, start, woff, L);
            } else if (kind == SHGPU_NODE_AGGR_LOWER)
                r = split_window<SS, SPLIT_HASH>(keys, sh, start, woff, L.stage, staged);
            else
                r = split_window<1, SPLIT_HASH>(keys, sh, start, woff, L.stage, staged);
            if (r != SHGPU_NOT_FOUND && (r >> rsh) >= lim) r = SHGPU_NOT_FOUND;
            if (r != SHGPU_NOT_FOUND) {
                // same value in every lane; waiting for each returned value orders the best seed before the
                // found flag and the flag before the completion below (all resolve at the coherence point)
                const unsigned long long prev = atomicMin(&st->best, (unsigned long long)r);
                asm volatile("" ::"v"(prev));
                const unsigned long long prevf = atomicOr(&st->gd, GD_FOUND);
                asm volatile("" ::"v"(prevf));
            }
        }
        const bool more = r == SHGPU_NOT_FOUND && !exhausted && !found;
        old = wave_fetch_add64(&st->gd, 1ull + (more ? GD_HANDED : 0ull));
        const uint64_t handed = gd_handed(old);  // windows handed out before this operation
        if (gd_completed(old) + 1 == handed) {
            // windows [0, handed) are all complete; a seed found in one of them set GD_FOUND before its
            // completion, so without the flag none of them holds one
            const uint64_t b = (old & GD_FOUND) ? read_best(st) : SHGPU_NOT_FOUND;
            const uint64_t covered = handed * W;
            if ((b != SHGPU_NOT_FOUND && (b >> rsh) < covered) || (b == SHGPU_NOT_FOUND && covered >= lim)) {
                if (wave_fetch_add(&st->finalized, 1u) == 0) cont = finalize_node<VARIANT>(a, nid, buf, b, persistent);
            }
        }
        if (!more) {
            reset_priority();
            return cont;
        }
        w = handed;  // the window this operation handed out
        found = (old & GD_FOUND) != 0;
    }
}

// A node another wave is searching in windows that this wave may join: its queue entry, with this wave counted
// among the node's helpers, or EMPTY_SLOT. A node takes at most help_cap helpers at once (unless it has searched
// so long that every wave may help): a wave joins after raising the node's helper count below the cap (waves
// that scan the registry at the same moment would otherwise all join a newly registered node), and a node whose
// best seed is already known needs no help.
__device__ __forceinline__ uint32_t find_help(const BuildArgs &a, uint32_t slot) {
    const uint32_t lane = lane_id();
    const uint32_t first = (slot * 97u) % a.nslots;
    for (uint32_t base = 0; base < a.nslots; base += 64) {
        uint32_t i = first + base + lane;
        if (i >= a.nslots) i -= a.nslots;
        const uint32_t v = base + lane < a.nslots ? __hip_atomic_load(&a.help[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                                  : EMPTY_SLOT;
        bool open = false;
        if (v != EMPTY_SLOT) {
            const NodeState *st = &a.state[v & ~ENTRY_BUF];
            const unsigned long long gd = __hip_atomic_load(&st->gd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            open = !(gd & GD_FOUND) &&
                   (__hip_atomic_load(&st->helpers, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < st->help_cap ||
                    gd_completed(gd) >= st->help_all);
        }
        const uint64_t hit = ballot(open);
        if (hit) {
            const uint32_t entry = (uint32_t)__builtin_amdgcn_readlane((int)v, (int)__builtin_ctzll(hit));
            NodeState *st = &a.state[entry & ~ENTRY_BUF];
            if (wave_fetch_add(&st->helpers, 1u) >= uniform(st->help_cap) &&
                gd_completed(uniform64(__hip_atomic_load(&st->gd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) < uniform(st->help_all)) {
                wave_fetch_add(&st->helpers, 0xFFFFFFFFu);  // full meanwhile: leave the count as it was
                return EMPTY_SLOT;
            }
            return entry;
        }
    }
    return EMPTY_SLOT;
}

// A wave runs either a windowed search (pf) or a ShockHash2 leaf (sh2), never both at once.
template <int VARIANT, bool WIDE>
union BuildLds {
    PfLds pf;
    typename std::conditional<VARIANT == SHGPU_SHOCKHASH2, Sh2LdsAny<WIDE>, char>::type sh2;
};

constexpr uint32_t SPIN_LIMIT = 1u << 27;  // sleeps before a wait is declared stuck (well over ten seconds)

// Persistent construction kernel. Every wave takes queue slots in order; a finished split node publishes its
// children, except the first (largest) one, which the wave that finished the node runs next itself. With
// SHGPU_TUNE_SPLITS_FIRST the queue has a split part (slots [0, n_split)) and a leaf part (slots [n_split,
// total)) and children go to the part of their kind, so all split nodes are taken before any leaf; otherwise it
// is one FIFO. While a wave's slot is not yet published (its parent is still being split) the wave helps
// windowed searches. When every slot has been taken, waves keep helping until every node is finished.
// ShockHash2 leaves are searched by the wave that takes them; split nodes and ShockHash1 leaves by their owner
// in windows that idle waves may join (help registry). The loop has one place that searches windows and one
// that searches a ShockHash2 leaf, so each is compiled once.
template <int VARIANT, int SL, int SS, bool WIDE>
__global__ void __launch_bounds__(BLOCK, 6) k_build(BuildArgs a) {
    __shared__ BuildLds<VARIANT, WIDE> lds[WAVES_PER_BLOCK];
    const uint32_t w = threadIdx.x / 64;
    const uint32_t slot = blockIdx.x * WAVES_PER_BLOCK + w;
    BuildLds<VARIANT, WIDE> &L = lds[w];
    uint8_t *ws = VARIANT == SHGPU_SHOCKHASH2 ? a.scratch + (size_t)slot * a.sh2_wave_bytes : nullptr;
    const uint32_t total = uniform(a.ctrl[C_NODES]);  // written by the forest plan
    const uint32_t windowed = uniform(a.ctrl[C_WINDOWED]);
    uint32_t ticket = EMPTY_SLOT;  // queue slot this wave holds, not yet published
    uint32_t cont = EMPTY_SLOT;    // child this wave runs next
    bool taken_all = false;        // every queue slot has been handed out
    for (uint32_t spin = 0;;) {
        uint32_t entry = cont;
        bool own = true;
        cont = EMPTY_SLOT;
        if (entry == EMPTY_SLOT) {
            if (!taken_all && ticket == EMPTY_SLOT) {
                ticket = grab(&a.ctrl[C_HEAD]);
                if (ticket >= total) { taken_all = true; ticket = EMPTY_SLOT; }
            }
            if (ticket != EMPTY_SLOT) {
                const uint32_t v = uniform(__hip_atomic_load(&a.queue[ticket], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
                if (v == SKIP_SLOT) { ticket = EMPTY_SLOT; spin = 0; continue; }  // a child its parent's wave runs
                if (v != EMPTY_SLOT) { entry = v; ticket = EMPTY_SLOT; }
            } else if (uniform(__hip_atomic_load(&a.ctrl[C_FINISHED], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) >= total) {
                break;
            }
            if (entry == EMPTY_SLOT) {
                entry = find_help(a, slot);
                own = false;
                if (entry == EMPTY_SLOT) {
                    // Every slot taken and every windowed node started (C_STARTED reaches C_WINDOWED): no node
                    // will register for help any more, and a wave that finds none open leaves instead of polling
                    // beside the waves still searching (checked every 8th poll: ~6000 waves polling one word load
                    // a single L2 channel).
                    if (taken_all && (spin & 7) == 0 &&
                        uniform(__hip_atomic_load(&a.ctrl[C_STARTED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= windowed)
                        break;
                    if (++spin >= SPIN_LIMIT) { atomicExch(&a.ctrl[C_ERROR], 1u); break; }
                    if (taken_all) __builtin_amdgcn_s_sleep(8);
                    else __builtin_amdgcn_s_sleep(4);
                    continue;
                }
            }
            spin = 0;
        }
        __threadfence();
        // the same in every lane: keep the compiler's control flow scalar
        entry = uniform(entry);
        own = uniform(own ? 1u : 0u) != 0;
        const uint32_t nid = entry & ~ENTRY_BUF, buf = entry >> 31;
        const shgpu_node nd = a.nodes[nid];
        const uint64_t t_start = task_clock();
        if (VARIANT == SHGPU_SHOCKHASH2 && uniform(nd.kind) == SHGPU_NODE_LEAF) {
            if constexpr (VARIANT == SHGPU_SHOCKHASH2) {
                const uint32_t o = uniform(nd.offset), m = uniform(nd.m);
                const uint64_t code = sh2_leaf_search_any<WIDE>(key_buf(a, buf) + o, m, ws, L.sh2, a.sh2_caps, a.flags, t_start);
                if (buf) copy_keys(a.tmp + o, a.keys + o, m);
                a.codes[nid] = code;
                wave_fetch_add(&a.ctrl[C_FINISHED], 1u);
            }
            continue;
        }
        if (own) {
            wave_fetch_add(&a.ctrl[C_STARTED], 1u);  // windowed nodes started
            __hip_atomic_store(&a.help[slot], entry, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        cont = work_windows<VARIANT, SL, SS>(a, nid, buf, L.pf, true, t_start);
        if (own) __hip_atomic_store(&a.help[slot], EMPTY_SLOT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else wave_fetch_add(&a.state[nid].helpers, 0xFFFFFFFFu);
    }
}

// Plain form: one launch per node (the host walks the forest in DFS order on one stream). Every
// wave of the launch joins the node's windowed search; a ShockHash2 leaf is searched by wave 0.
template <int VARIANT, int SL, int SS, bool WIDE>
__global__ void __launch_bounds__(64) k_node(BuildArgs a, uint32_t nid, uint32_t buf) {
    __shared__ BuildLds<VARIANT, WIDE> L;
    const shgpu_node nd = a.nodes[nid];
    if (VARIANT == SHGPU_SHOCKHASH2 && uniform(nd.kind) == SHGPU_NODE_LEAF) {
        if constexpr (VARIANT == SHGPU_SHOCKHASH2) {
            if (blockIdx.x == 0) {
                const uint32_t o = uniform(nd.offset), m = uniform(nd.m);
                a.codes[nid] = sh2_leaf_search_any<WIDE>(key_buf(a, buf) + o, m, a.scratch, L.sh2, a.sh2_caps, a.flags,
                                                         task_clock());
                if (buf) copy_keys(a.tmp + o, a.keys + o, m);
            }
        }
        return;
    }
    work_windows<VARIANT, SL, SS>(a, nid, buf, L.pf, false, task_clock());
}


template <int VARIANT, int SL, int SS, bool WIDE>
hipError_t launch_build_t(const BuildArgs &a, int grid, hipStream_t s) {
    k_build<VARIANT, SL, SS, WIDE><<<grid, BLOCK, 0, s>>>(a);
    return hipGetLastError();
}

template <int VARIANT, bool WIDE>
hipError_t launch_build_w(uint32_t sl, uint32_t ss, const BuildArgs &a, int grid, hipStream_t s) {
    // ShockHash2 leaves do not use the ShockHash1 window: one SL instantiation serves every setting
#define SHGPU_CASE(A, B)                                                                                      \
    if constexpr (VARIANT != SHGPU_SHOCKHASH2 || A == 1) {                                                    \
        if ((VARIANT == SHGPU_SHOCKHASH2 || sl == A) && ss == B) return launch_build_t<VARIANT, A, B, WIDE>(a, grid, s); \
    }
    SHGPU_CASE(1, 1) SHGPU_CASE(1, 2) SHGPU_CASE(1, 4) SHGPU_CASE(2, 1) SHGPU_CASE(2, 2) SHGPU_CASE(2, 4)
    SHGPU_CASE(4, 1) SHGPU_CASE(4, 2) SHGPU_CASE(4, 4)
#undef SHGPU_CASE
    return hipErrorInvalidValue;  // seeds per lane outside {1, 2, 4}
}

template <int VARIANT>
hipError_t launch_build_variant(uint32_t sl, uint32_t ss, bool wide, const BuildArgs &a, int grid, hipStream_t s) {
    if constexpr (VARIANT == SHGPU_SHOCKHASH2) {
        if (wide) return launch_build_w<VARIANT, true>(sl, ss, a, grid, s);
    }
    return launch_build_w<VARIANT, false>(sl, ss, a, grid, s);
}

template <int VARIANT>
hipError_t launch_node_variant(bool wide, const BuildArgs &a, uint32_t nid, uint32_t buf, uint32_t waves, hipStream_t s) {
    constexpr int SL = VARIANT == SHGPU_SHOCKHASH2 ? 1 : 4;
    if constexpr (VARIANT == SHGPU_SHOCKHASH2) {
        if (wide) {
            k_node<VARIANT, SL, 4, true><<<waves, 64, 0, s>>>(a, nid, buf);
            return hipGetLastError();
        }
    }
    k_node<VARIANT, SL, 4, false><<<waves, 64, 0, s>>>(a, nid, buf);
    return hipGetLastError();
}

}  // namespace shgpu
