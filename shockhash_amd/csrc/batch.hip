// Batched per-node searches behind shgpu_sh1_leaf_search, shgpu_sh2_leaf_search, shgpu_split_search and
// shgpu_split_search_noalloc: each wave takes tasks from a device counter and searches one at a time.
#include "build_common.h"
#include "search_sh1.h"
#include "search_sh2.h"
#include "search_split.h"

namespace shgpu {

template <int S>
__global__ void __launch_bounds__(BLOCK) k_sh1_batch(const uint64_t *__restrict__ keys, const shgpu_leaf_task *__restrict__ tasks,
                                                     uint32_t n, uint64_t *__restrict__ out, uint32_t *counter) {
    __shared__ PfLds lds[WAVES_PER_BLOCK];
    PfLds &L = lds[threadIdx.x / 64];
    for (;;) {
        const uint32_t t = grab(counter);
        if (t >= n) break;
        const shgpu_leaf_task tk = tasks[t];
        const uint64_t x = sh1_leaf_search<S>(keys + uniform(tk.offset), uniform(tk.m), uniform64(tk.start_seed),
                                              uniform64(tk.max_trials), L);
        out[t] = x;  // wave-uniform value, stored by every lane
    }
}

template <bool WIDE>
__global__ void __launch_bounds__(BLOCK) k_sh2_batch(const uint64_t *__restrict__ keys, const shgpu_leaf_task *__restrict__ tasks,
                                                     uint32_t n, uint64_t *__restrict__ out, uint32_t *counter, uint8_t *scratch,
                                                     uint64_t wave_bytes, Sh2Caps caps, uint32_t flags) {
    __shared__ Sh2LdsAny<WIDE> lds[WAVES_PER_BLOCK];
    const uint32_t w = threadIdx.x / 64;
    Sh2LdsAny<WIDE> &L = lds[w];
    uint8_t *ws = scratch + (size_t)(blockIdx.x * WAVES_PER_BLOCK + w) * wave_bytes;
    for (;;) {
        const uint32_t t = grab(counter);
        if (t >= n) break;
        const shgpu_leaf_task tk = tasks[t];
        const uint64_t code = sh2_leaf_search_any<WIDE>(keys + uniform(tk.offset), uniform(tk.m), ws, L, caps, flags, task_clock());
        out[t] = code;
    }
}

// A split task the kernels can run on a key array of n_keys keys: 2 <= fanout <= 4, unit < m <= 65535, the
// node inside the array, a known kind and hash flavour (the checks shgpu_split_search makes on the host).
__device__ __forceinline__ bool split_task_valid(const shgpu_split_task &tk, uint64_t n_keys) {
    const uint32_t m = uniform(tk.m), unit = uniform(tk.unit), kind = uniform(tk.kind);
    if (m > 65535 || unit == 0 || unit >= m || kind > SHGPU_NODE_AGGR_LOWER || uniform(tk.hash) > SHGPU_SPLIT_SIMD)
        return false;
    const uint32_t fanout = kind == SHGPU_NODE_SPLIT2 ? 2u : (m + unit - 1) / unit;
    return fanout >= 2 && fanout <= 4 && (uint64_t)uniform(tk.offset) + m <= n_keys;
}

// CHECK: validate every task on the device (split_task_valid against n_keys); an invalid task leaves its keys
// untouched and reports SHGPU_NOT_FOUND.
template <int S, bool CHECK>
__global__ void __launch_bounds__(BLOCK) k_split_batch(uint64_t *keys, uint64_t *tmp, const shgpu_split_task *__restrict__ tasks,
                                                       uint32_t n, uint64_t *__restrict__ out, uint32_t *counter,
                                                       uint64_t n_keys) {
    __shared__ PfLds lds[WAVES_PER_BLOCK];
    uint64_t *stage = lds[threadIdx.x / 64].stage;
    for (;;) {
        const uint32_t t = grab(counter);
        if (t >= n) break;
        const shgpu_split_task tk = tasks[t];
        if (CHECK && !split_task_valid(tk, n_keys)) {
            out[t] = SHGPU_NOT_FOUND;
            continue;
        }
        const uint32_t off = uniform(tk.offset), m = uniform(tk.m);
        const SplitShapeDev sh = make_split_shape(m, uniform(tk.unit), uniform(tk.kind), uniform(tk.hash));
        const uint64_t x = split_search<S>(keys + off, sh, uniform64(tk.start_seed), uniform64(tk.max_trials), stage);
        if (x != SHGPU_NOT_FOUND) split_partition(keys + off, tmp + off, sh, x);
        out[t] = x;
    }
}


hipError_t launch_sh1_batch(uint32_t seeds_per_lane, int grid, hipStream_t s, const uint64_t *keys,
                            const shgpu_leaf_task *tasks, uint32_t n, uint64_t *out, uint32_t *ctr) {
    switch (seeds_per_lane) {
        case 1: k_sh1_batch<1><<<grid, BLOCK, 0, s>>>(keys, tasks, n, out, ctr); break;
        case 2: k_sh1_batch<2><<<grid, BLOCK, 0, s>>>(keys, tasks, n, out, ctr); break;
        default: k_sh1_batch<4><<<grid, BLOCK, 0, s>>>(keys, tasks, n, out, ctr); break;
    }
    return hipGetLastError();
}

hipError_t launch_sh2_batch(bool wide, int grid, hipStream_t s, const uint64_t *keys, const shgpu_leaf_task *tasks, uint32_t n,
                            uint64_t *out, uint32_t *ctr, uint8_t *scratch, uint64_t wave_bytes, Sh2Caps caps,
                            uint32_t flags) {
    if (wide) k_sh2_batch<true><<<grid, BLOCK, 0, s>>>(keys, tasks, n, out, ctr, scratch, wave_bytes, caps, flags);
    else k_sh2_batch<false><<<grid, BLOCK, 0, s>>>(keys, tasks, n, out, ctr, scratch, wave_bytes, caps, flags);
    return hipGetLastError();
}

template <bool CHECK>
static void launch_split_t(uint32_t seeds_per_lane, int grid, hipStream_t s, uint64_t *keys, uint64_t *tmp,
                           const shgpu_split_task *tasks, uint32_t n, uint64_t *out, uint32_t *ctr, uint64_t n_keys) {
    switch (seeds_per_lane) {
        case 1: k_split_batch<1, CHECK><<<grid, BLOCK, 0, s>>>(keys, tmp, tasks, n, out, ctr, n_keys); break;
        case 2: k_split_batch<2, CHECK><<<grid, BLOCK, 0, s>>>(keys, tmp, tasks, n, out, ctr, n_keys); break;
        default: k_split_batch<4, CHECK><<<grid, BLOCK, 0, s>>>(keys, tmp, tasks, n, out, ctr, n_keys); break;
    }
}

hipError_t launch_split_batch(bool check, uint32_t seeds_per_lane, int grid, hipStream_t s, uint64_t *keys, uint64_t *tmp,
                              const shgpu_split_task *tasks, uint32_t n, uint64_t *out, uint32_t *ctr, uint64_t n_keys) {
    if (check) launch_split_t<true>(seeds_per_lane, grid, s, keys, tmp, tasks, n, out, ctr, n_keys);
    else launch_split_t<false>(seeds_per_lane, grid, s, keys, tmp, tasks, n, out, ctr, n_keys);
    return hipGetLastError();
}

}  // namespace shgpu
