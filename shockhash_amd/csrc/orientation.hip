// Retrieval input of leaves on the device: each leaf's cuckoo insertions replayed in leaf key order, which fixes
// the cell (and so the orientation bit) of every key (include/ShockHash.h:425-432, include/SIMDShockHash.hpp:536-540,
// include/shockhash/ShockHash2-internal.h:548-572).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/shockhash_amd.h"
#include "device_common.h"
#include "orientation.h"

namespace {

// Exact decodes of the pairing functions (PairingFunction.h:52-56,83-91). The reference stores only codes
// whose double-precision decode is exact (isFloatAccurate*), so for those the integer-corrected estimate
// below equals the reference's decode.
__device__ __forceinline__ void unpair_triangular(uint64_t z, uint64_t &x, uint64_t &y) {
    uint64_t t = (uint64_t)floor(0.5 + sqrt(0.25 + 2.0 * (double)z));
    while (t > 1 && t * (t - 1) / 2 > z) t--;
    while ((t + 1) * t / 2 <= z) t++;
    x = t;
    y = z - t * (t - 1) / 2;
}
__device__ __forceinline__ void unpair_elegant(uint64_t z, uint64_t &x, uint64_t &y) {
    uint64_t f = (uint64_t)floor(sqrt((double)z));
    while (f > 0 && f * f > z) f--;
    while ((f + 1) * (f + 1) <= z) f++;
    const uint64_t r = z - f * f;
    if (r < f) { x = r; y = f; } else { x = f; y = r - f; }
}

// Per-thread tables of the orientation replay in LDS, one byte column per thread: byte x of thread t at
// x * BS + t for a block of BS threads. Cells and key indices of a leaf are below MAXM: 64 in blocks of 256
// threads, 128 (ShockHash2 leaves above 64 keys) in blocks of 128 threads, 48 KB of LDS either way.
constexpr uint32_t ORIENT_BLOCK = 256;
constexpr uint8_t CELL_EMPTY = 0xFF;

template <uint32_t BS>
struct OrientCol {
    uint8_t *p;
    __device__ uint8_t &operator[](uint32_t x) const { return p[x * BS]; }
};

// Cuckoo insertion of key k with cells c1, c2 into cell[] (CELL_EMPTY = free), entries identified by key
// index; the steps of TinyBinaryCuckooHashTable::insert (TinyBinaryCuckooHashTable.h:169-198). False when
// the key cannot be placed.
template <uint32_t BS>
__device__ bool orient_insert(OrientCol<BS> cell, OrientCol<BS> xr, uint32_t k, uint32_t c1, uint32_t c2) {
    uint32_t entry = k;
    xr[k] = (uint8_t)(c1 ^ c2);
    if (cell[c1] == CELL_EMPTY) { cell[c1] = (uint8_t)entry; return true; }
    if (cell[c2] == CELL_EMPTY) { cell[c2] = (uint8_t)entry; return true; }
    uint32_t cur = c1;
    for (uint32_t steps = 0; steps < 4096; steps++) {
        const uint32_t alt = xr[entry] ^ cur;
        const uint32_t evicted = cell[alt];
        cell[alt] = (uint8_t)entry;
        entry = evicted;
        if (entry == CELL_EMPTY) return true;
        cur = alt;
        if (cur == c1 && entry == k) return false;
    }
    return false;
}

// One thread per leaf: replay the leaf's insertions in leaf key order and write its retrieval input at
// out_offset. ShockHash1 / SIMDShockHash (BIPARTITE false): cells c1 < m/2 <= c2 of remix(key + seed) with
// seed = code + start seed, entries in cell order with bit = (cell >= m/2) (ShockHash.h:425-432,
// SIMDShockHash.hpp:536-540). ShockHash2: cells sh2(s1) + m/2 and sh2(s2) of the pair code's seeds, entries in
// key order with bit = (the key sits in its second cell) (ShockHash2-internal.h:548-572). The cell,
// eviction-XOR and second-cell tables (64 bytes each) live in LDS: the replay is a chain of dependent
// table reads.
// MAXM 64: leaves above 64 keys are left to the MAXM 128 launch (error bit 4 says there are some); MAXM 128: only
// the leaves above 64 keys.
template <bool BIPARTITE, uint32_t MAXM, uint32_t BS>
__global__ void __launch_bounds__(BS) k_leaf_orientation(const uint64_t *__restrict__ keys,
                                                         const shgpu_orient_task *__restrict__ tasks, uint64_t n,
                                                         uint64_t *out_keys, uint8_t *out_bits, uint32_t *error) {
    __shared__ uint8_t lds[3 * MAXM * BS];
    const OrientCol<BS> cell{lds + threadIdx.x}, xr{lds + MAXM * BS + threadIdx.x}, c2s{lds + 2 * MAXM * BS + threadIdx.x};
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const shgpu_orient_task tk = tasks[t];
        const uint64_t *lk = keys + tk.key_offset;
        const uint32_t m = tk.m;
        if (m < 1 || m > 128 || (!BIPARTITE && m > 64)) { atomicOr(error, 1u); continue; }
        if (MAXM == 64 && m > 64) { atomicOr(error, 4u); continue; }
        if (MAXM == 128 && m <= 64) continue;
        for (uint32_t c = 0; c < m; c++) cell[c] = CELL_EMPTY;
        bool ok = true;
        if (!BIPARTITE && tk.rotation == 0) {
            for (uint32_t k = 0; k < m && ok; k++) {
                const uint64_t r = shgpu::remix(lk[k] + tk.seed);
                const uint32_t c1 = (uint32_t)(((uint64_t)(uint32_t)r * (m / 2)) >> 32);
                const uint32_t c2 = (uint32_t)(((r >> 32) * ((m + 1) / 2)) >> 32) + m / 2;
                ok = orient_insert(cell, xr, k, c1, c2);
            }
            if (!ok) { atomicOr(error, 2u); continue; }
            for (uint32_t c = 0; c < m; c++) {
                out_keys[tk.out_offset + c] = lk[cell[c]];
                out_bits[tk.out_offset + c] = c < m / 2 ? 0 : 1;
            }
        } else if (!BIPARTITE) {
            // Rotation-fitted leaf (ShockHash.h:316-403): insertion order A keys (even) in leaf order, then B keys
            // (odd) in reverse leaf order; A cells of seed x & ~amask, B cells of seed x rotated by r. The c2s
            // column maps insertion positions to leaf indices. Entries in cell order, bit = (cell' >= m/2) with
            // cell' the cell before rotation (ShockHash.h:380-388).
            const uint32_t rot = tk.rotation - 1;
            const uint64_t amask = m > 32 ? 7ull : 0ull;
            if (rot >= m) { atomicOr(error, 1u); continue; }
            uint32_t na = 0, ib = m - 1;
            for (uint32_t k = 0; k < m; k++) {
                if ((lk[k] & 1) == 0) c2s[na++] = (uint8_t)k;
                else c2s[ib--] = (uint8_t)k;
            }
            for (uint32_t h = 0; h < m && ok; h++) {
                const uint64_t key = lk[c2s[h]];
                const bool odd = (key & 1) != 0;
                const uint64_t r = shgpu::remix(key + (odd ? tk.seed : (tk.seed & ~amask)));
                uint32_t c1 = (uint32_t)(((uint64_t)(uint32_t)r * (m / 2)) >> 32);
                uint32_t c2 = (uint32_t)(((r >> 32) * ((m + 1) / 2)) >> 32) + m / 2;
                if (odd) {
                    c1 = (c1 + rot) % m;
                    c2 = (c2 + rot) % m;
                }
                ok = orient_insert(cell, xr, h, c1, c2);
            }
            if (!ok) { atomicOr(error, 2u); continue; }
            for (uint32_t c = 0; c < m; c++) {
                const uint64_t key = lk[c2s[cell[c]]];
                const uint32_t own = (key & 1) ? (c + m - rot) % m : c;
                out_keys[tk.out_offset + c] = key;
                out_bits[tk.out_offset + c] = own < m / 2 ? 0 : 1;
            }
        } else {
            uint64_t s1, s2;
            unpair_triangular(tk.seed, s1, s2);
            uint64_t a1 = s1, b1 = s1, a2 = s2, b2 = s2;  // half-seeds of even / odd keys (m >= 10)
            if (m >= 10) {
                unpair_elegant(s1, a1, b1);
                unpair_elegant(s2, a2, b2);
            }
            const uint64_t H = (m + 1) / 2;
            for (uint32_t k = 0; k < m && ok; k++) {
                const uint64_t key = lk[k];
                const bool even = (key & 1) == 0;
                const uint32_t c1 = (uint32_t)__umul64hi(shgpu::remix(key + (even ? a1 : b1)), H) + m / 2;
                const uint32_t c2 = (uint32_t)__umul64hi(shgpu::remix(key + (even ? a2 : b2)), H);
                c2s[k] = (uint8_t)c2;
                ok = orient_insert(cell, xr, k, c1, c2);
            }
            if (!ok) { atomicOr(error, 2u); continue; }
            for (uint32_t k = 0; k < m; k++) {
                out_keys[tk.out_offset + k] = lk[k];
                out_bits[tk.out_offset + k] = cell[c2s[k]] == k ? 1 : 0;
            }
        }
    }
}

__global__ void k_zero_u32(uint32_t *p) { p[0] = 0; }

struct OrBuf {
    void *p = nullptr;
    ~OrBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int shgpu_leaf_orientation_impl(hipStream_t s, int bipartite, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                                uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits, std::string &err) {
    if (n == 0) return SHGPU_OK;
    OrBuf flag;
    if (!shgpu::device_malloc(&flag.p, 64)) { err = "orientation: hipMalloc failed"; return SHGPU_ERR_NOMEM; }
    uint32_t *d_err = (uint32_t *)flag.p;
    k_zero_u32<<<1, 1, 0, s>>>(d_err);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + ORIENT_BLOCK - 1) / ORIENT_BLOCK, 4096);
    if (bipartite) k_leaf_orientation<true, 64, ORIENT_BLOCK><<<grid, ORIENT_BLOCK, 0, s>>>(d_keys, d_tasks, n, d_out_keys, d_out_bits, d_err);
    else k_leaf_orientation<false, 64, ORIENT_BLOCK><<<grid, ORIENT_BLOCK, 0, s>>>(d_keys, d_tasks, n, d_out_keys, d_out_bits, d_err);
    uint32_t h = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d_err, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && bipartite && (h & 4)) {  // ShockHash2 leaves above 64 keys
        const uint32_t grid2 = (uint32_t)std::min<uint64_t>((n + 127) / 128, 4096);
        k_leaf_orientation<true, 128, 128><<<grid2, 128, 0, s>>>(d_keys, d_tasks, n, d_out_keys, d_out_bits, d_err);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h, d_err, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) { err = std::string("orientation: ") + hipGetErrorString(e); return SHGPU_ERR_HIP; }
    if (h & 1) { err = "orientation: a leaf has no keys or more than 64 (128 for ShockHash2)"; return SHGPU_ERR_ARG; }
    if (h & 2) { err = "orientation: a leaf's keys cannot all be placed with its seed (wrong code?)"; return SHGPU_ERR_ARG; }
    return SHGPU_OK;
}
