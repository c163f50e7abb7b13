// One-bit retrieval construction on the device: the structure that returns, for every key of the set,
// the orientation bit of the key's leaf cell. It takes the place of the reference's
// SimpleRibbon<1, ...> construction (include/ShockHash.h:152-153,657-658; include/ShockHash2.h:134,
// 491-497); query answers depend only on the bits it returns for the set's keys.
//
// Layout (shared with the host class shockhash::BitRetrieval, which answers the queries): 3 * seg cells of
// one bit, seg = floor(1.23 n / 3) + 8; key k uses one cell in each third, derived from remix of k and the
// seed; its bit is the XOR of its three cells. Construction is hypergraph peeling: repeatedly take a cell
// with exactly one key, remove that key; then assign bits in reverse peeling order. A peeling that does not
// take every key (rare, or duplicate keys) retries with the next seed.
//
// On the device the peeling runs in rounds. Round r: every cell that reached degree 1 claims its key
// (compare-and-swap on the key's round; a key reachable through two such cells is claimed once), then every
// claimed key leaves its three cells (one 64-bit atomic subtracts it from the cell's degree and index sum),
// and cells that drop to degree 1 form the next round's frontier. Claims only read cells, removals only
// update them, so each round's two kernels see consistent cells. Keys claimed in one round have distinct cells of
// degree 1 and none lies on another's cells, so the bits of one round are assigned in parallel, rounds in
// reverse order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/shockhash_amd.h"
#include "device_common.h"
#include "retrieval.h"

namespace {

constexpr uint32_t UNPEELED = 0xFFFFFFFFu;
constexpr int RT_BLOCK = 256, RT_GRID = 2048;
constexpr int ROUNDS_PER_CHECK = 32;
constexpr int MAX_ROUNDS = 1 << 16;

__device__ __forceinline__ void rt_cells(uint64_t key, uint64_t seed, uint64_t seg, uint64_t c[3]) {
    const uint64_t h0 = shgpu::remix(key + seed * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull);
    const uint64_t h1 = shgpu::remix(h0 ^ 0xd6e8feb86659fd93ull);
    c[0] = __umul64hi(h0, seg);
    c[1] = seg + __umul64hi(h1, seg);
    c[2] = 2 * seg + __umul64hi(shgpu::remix(h1 + 0x8cb92ba72f3d8dd7ull), seg);
}

#define GRID_STRIDE(i, n) for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// A cell's word: its degree (number of keys not yet peeled) in bits 40..63 and the sum of those keys'
// indices in bits 0..39. One 64-bit atomic add or subtract updates both; a cell of degree 1 holds its key's
// index. Sums fit while a cell holds fewer than 256 keys (a degree of 255 or more is reported as an error:
// at 1.23 cells per key only repeated keys get there).
constexpr int CELL_DEG_SHIFT = 40;
constexpr uint64_t CELL_ONE = 1ull << CELL_DEG_SHIFT, CELL_SUM_MASK = CELL_ONE - 1;
constexpr uint64_t CELL_MAX_DEG = 255;
__device__ __forceinline__ uint64_t cell_deg(uint64_t w) { return w >> CELL_DEG_SHIFT; }

__global__ void k_rt_clear(uint64_t *cw, uint8_t *bits, uint64_t ncell, uint32_t *round_of, uint64_t n,
                           uint32_t *counters, uint32_t *round_begin) {
    GRID_STRIDE(v, ncell) {
        cw[v] = 0;
        bits[v] = 0;
    }
    GRID_STRIDE(i, n) round_of[i] = UNPEELED;
    if (blockIdx.x == 0 && threadIdx.x < 8) counters[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) round_begin[0] = 0;
}

// counters[3] is set when a cell reaches CELL_MAX_DEG keys.
__global__ void k_rt_degrees(const uint64_t *__restrict__ keys, uint64_t n, uint64_t seed, uint64_t seg, uint64_t *cw,
                             uint32_t *counters) {
    GRID_STRIDE(i, n) {
        uint64_t c[3];
        rt_cells(keys[i], seed, seg, c);
        for (int k = 0; k < 3; k++) {
            const uint64_t old = atomicAdd((unsigned long long *)&cw[c[k]], (unsigned long long)(CELL_ONE + i));
            if (cell_deg(old) + 1 >= CELL_MAX_DEG) counters[3] = 1;
        }
    }
}

// Appends to a global list (out, length *len) staged per block in LDS: a block reserves its range with one
// global atomic when it ends, instead of one per wave and append. Appends to one counter serialise in the
// memory system (about 10 ns each), which made the frontier and peeled lists of a 10M-key build cost
// milliseconds. Entries past the stage's capacity go to the list directly. Every thread of the block must
// call flush.
constexpr uint32_t STAGE_CAP = 4096;
struct BlockAppend {
    uint32_t *buf, *n, *base;  // LDS: staged entries, their count, the reserved global position
    uint32_t *out, *len;
    __device__ void init() {
        if (threadIdx.x == 0) *n = 0;
        __syncthreads();
    }
    __device__ void push(uint32_t v) {
        const uint32_t slot = atomicAdd(n, 1u);
        if (slot < STAGE_CAP) buf[slot] = v;
        else out[atomicAdd(len, 1u)] = v;
    }
    __device__ void flush() {
        __syncthreads();
        const uint32_t cnt = *n < STAGE_CAP ? *n : STAGE_CAP;
        if (threadIdx.x == 0) *base = cnt ? atomicAdd(len, cnt) : 0u;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) out[*base + j] = buf[j];
    }
};
#define BLOCK_APPEND(name, out_, len_)                                   \
    __shared__ uint32_t name##_buf[STAGE_CAP], name##_n, name##_base;    \
    BlockAppend name{name##_buf, &name##_n, &name##_base, (out_), (len_)}; \
    name.init()

// counters: [0] frontier length of the current round, [1] next frontier length, [2] keys peeled
__global__ void __launch_bounds__(RT_BLOCK) k_rt_frontier0(const uint64_t *cw, uint64_t ncell, uint32_t *frontier,
                                                           uint32_t *counters) {
    BLOCK_APPEND(app, frontier, &counters[0]);
    GRID_STRIDE(v, ncell) {
        if (cell_deg(cw[v]) == 1) app.push((uint32_t)v);
    }
    app.flush();
}

// Round r, claims: a frontier cell still of degree 1 claims its only key.
// A cell whose index sum has overflowed (degree 255 or more, reported through counters[3]) can show a degree of
// 1 with any sum: indices at or past n are not claimed.
__global__ void __launch_bounds__(RT_BLOCK) k_rt_claim(const uint32_t *frontier, uint32_t *counters, const uint64_t *cw,
                                                       uint64_t n, uint32_t *round_of, uint32_t *assigned, uint32_t *peeled,
                                                       uint32_t r) {
    BLOCK_APPEND(app, peeled, &counters[2]);
    const uint32_t len = counters[0];
    GRID_STRIDE(f, len) {
        const uint32_t v = frontier[f];
        const uint64_t w = cw[v];
        if (cell_deg(w) != 1 || (w & CELL_SUM_MASK) >= n) continue;
        const uint32_t i = (uint32_t)(w & CELL_SUM_MASK);
        if (atomicCAS(&round_of[i], UNPEELED, r) == UNPEELED) {
            assigned[i] = v;
            app.push(i);
        }
    }
    app.flush();
}

// Round r, removals: the keys claimed in round r (peeled[begin, counters[2])) leave their cells.
__global__ void __launch_bounds__(RT_BLOCK) k_rt_remove(const uint64_t *__restrict__ keys, uint64_t seed, uint64_t seg,
                                                        const uint32_t *peeled, const uint32_t *round_begin, uint32_t r,
                                                        const uint32_t *counters, uint64_t *cw, uint32_t *next_frontier,
                                                        uint32_t *next_len) {
    BLOCK_APPEND(app, next_frontier, next_len);
    const uint32_t begin = round_begin[r], end = counters[2];
    GRID_STRIDE(f, (uint64_t)(end - begin)) {
        const uint32_t i = peeled[begin + f];
        uint64_t c[3];
        rt_cells(keys[i], seed, seg, c);
        for (int k = 0; k < 3; k++) {
            const uint64_t old = atomicAdd((unsigned long long *)&cw[c[k]], (unsigned long long)(0 - (CELL_ONE + i)));
            if (cell_deg(old) == 2) app.push((uint32_t)c[k]);
        }
    }
    app.flush();
}

// Between rounds (one thread): the next frontier becomes the current one, the next round starts where the
// peeled list ends now.
__global__ void k_rt_advance(uint32_t *counters, uint32_t *round_begin, uint32_t r) {
    counters[0] = counters[1];
    counters[1] = 0;
    round_begin[r + 1] = counters[2];
}

// Bits of the keys peeled in round r, given every later round's.
__global__ void k_rt_assign(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ want, uint64_t seed, uint64_t seg,
                            const uint32_t *peeled, const uint32_t *round_begin, uint32_t r, const uint32_t *assigned,
                            uint8_t *bits) {
    const uint32_t begin = round_begin[r], end = round_begin[r + 1];
    GRID_STRIDE(f, (uint64_t)(end - begin)) {
        const uint32_t i = peeled[begin + f];
        uint64_t c[3];
        rt_cells(keys[i], seed, seg, c);
        const uint32_t v = assigned[i];
        uint32_t x = want[i] & 1u;
        for (int k = 0; k < 3; k++)
            if (c[k] != v) x ^= bits[c[k]];
        bits[v] = (uint8_t)x;
    }
}

__global__ void k_rt_pack(const uint8_t *__restrict__ bits, uint64_t ncell, uint64_t *words) {
    GRID_STRIDE(w, (ncell + 63) / 64) {
        uint64_t x = 0;
        const uint64_t base = w * 64;
        for (uint32_t b = 0; b < 64 && base + b < ncell; b++) x |= (uint64_t)(bits[base + b] & 1u) << b;
        words[w] = x;
    }
}

// ------------------------------------------------------------------ leaf orientation (retrieval input)
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
        if (!BIPARTITE) {
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

struct RtBuf {
    void *p = nullptr;
    ~RtBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

uint64_t shgpu_retrieval_cells_per_third(uint64_t n) { return (uint64_t)(n * 1.23 / 3.0) + 8; }

int shgpu_retrieval_build_impl(hipStream_t s, const uint64_t *d_keys, const uint8_t *d_bits, uint64_t n, uint64_t *seed_out,
                               uint64_t *seg_out, uint64_t *h_words, uint64_t words_capacity, std::string &err,
                               void **cache, size_t *cache_bytes) {
    const uint64_t seg = shgpu_retrieval_cells_per_third(n);
    const uint64_t ncell = 3 * seg, nwords = (ncell + 63) / 64;
    if (n >= UNPEELED || ncell >= UNPEELED) { err = "retrieval: more than 2^32 - 1 keys or cells"; return SHGPU_ERR_UNSUPPORTED; }
    if (words_capacity < nwords) { err = "retrieval: output capacity below " + std::to_string(nwords) + " words"; return SHGPU_ERR_ARG; }
    // one allocation: cell words (ncell u64), frontier, next frontier (ncell u32 each), bits (ncell u8),
    // round_of, assigned, peeled (n u32 each), round starts, counters, packed words; every part starts on a
    // 256-byte boundary
    const size_t parts[] = {ncell * 8, ncell * 4, ncell * 4, ncell, (n + 1) * 4, (n + 1) * 4, (n + 1) * 4,
                            (size_t)(MAX_ROUNDS + 2) * 4, 64, nwords * 8};
    size_t bytes = 0;
    for (size_t b : parts) bytes += (b + 255) / 256 * 256;
    RtBuf buf;
    void *base = nullptr;
    if (cache && cache_bytes) {
        if (*cache_bytes < bytes) {
            if (*cache) (void)hipFree(*cache);
            *cache = nullptr;
            *cache_bytes = 0;
            if (!shgpu::device_malloc(cache, bytes)) {
                *cache = nullptr;
                err = "retrieval: hipMalloc of " + std::to_string(bytes) + " bytes failed";
                return SHGPU_ERR_NOMEM;
            }
            *cache_bytes = bytes;
        }
        base = *cache;
    } else {
        if (!shgpu::device_malloc(&buf.p, bytes)) { err = "retrieval: hipMalloc of " + std::to_string(bytes) + " bytes failed"; return SHGPU_ERR_NOMEM; }
        base = buf.p;
    }
    char *q = (char *)base;
    auto take = [&](size_t b) { char *r = q; q += (b + 255) / 256 * 256; return r; };
    uint64_t *cw = (uint64_t *)take(ncell * 8);
    uint32_t *fr[2] = {(uint32_t *)take(ncell * 4), (uint32_t *)take(ncell * 4)};
    uint8_t *bits = (uint8_t *)take(ncell);
    uint32_t *round_of = (uint32_t *)take((n + 1) * 4), *assigned = (uint32_t *)take((n + 1) * 4), *peeled = (uint32_t *)take((n + 1) * 4);
    uint32_t *round_begin = (uint32_t *)take((size_t)(MAX_ROUNDS + 2) * 4);
    uint32_t *counters = (uint32_t *)take(64);
    uint64_t *words = (uint64_t *)take(nwords * 8);
    if ((size_t)(q - (char *)base) > bytes) { err = "retrieval: buffer layout"; return SHGPU_ERR_HIP; }
#define RT_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) { err = std::string(#call) + ": " + hipGetErrorString(e_); return SHGPU_ERR_HIP; } \
    } while (0)
    for (uint64_t seed = 0; seed < 64; seed++) {
        k_rt_clear<<<RT_GRID, RT_BLOCK, 0, s>>>(cw, bits, ncell, round_of, n, counters, round_begin);
        RT_HIP(hipGetLastError());
        k_rt_degrees<<<RT_GRID, RT_BLOCK, 0, s>>>(d_keys, n, seed, seg, cw, counters);
        k_rt_frontier0<<<RT_GRID, RT_BLOCK, 0, s>>>(cw, ncell, fr[0], counters);
        RT_HIP(hipGetLastError());
        // round r reads frontier fr[r & 1] (length counters[0]) and fills fr[(r + 1) & 1] (length counters[1])
        uint32_t rounds = 0, host_ctr[4] = {0, 0, 0, 0};
        for (;;) {
            for (int k = 0; k < ROUNDS_PER_CHECK; k++, rounds++) {
                if (rounds >= (uint32_t)MAX_ROUNDS) { err = "retrieval: peeling did not converge"; return SHGPU_ERR_HIP; }
                k_rt_claim<<<RT_GRID, RT_BLOCK, 0, s>>>(fr[rounds & 1], counters, cw, n, round_of, assigned, peeled, rounds);
                k_rt_remove<<<RT_GRID, RT_BLOCK, 0, s>>>(d_keys, seed, seg, peeled, round_begin, rounds, counters, cw,
                                                          fr[(rounds + 1) & 1], &counters[1]);
                k_rt_advance<<<1, 1, 0, s>>>(counters, round_begin, rounds);
            }
            RT_HIP(hipGetLastError());
            RT_HIP(hipMemcpyAsync(host_ctr, counters, sizeof(host_ctr), hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
            if (host_ctr[3]) { err = "retrieval: a cell holds 255 or more keys (repeated keys?)"; return SHGPU_ERR_ARG; }
            if (host_ctr[0] == 0) break;  // empty frontier: no further round can peel
        }
        if (host_ctr[2] != n) continue;  // not every key peeled: next seed
        // round_begin holds the start of every round and the end of the last one; rounds past the last
        // productive one are empty
        for (uint32_t r = rounds; r-- > 0;)
            k_rt_assign<<<RT_GRID, RT_BLOCK, 0, s>>>(d_keys, d_bits, seed, seg, peeled, round_begin, r, assigned, bits);
        k_rt_pack<<<RT_GRID, RT_BLOCK, 0, s>>>(bits, ncell, words);
        RT_HIP(hipGetLastError());
        RT_HIP(hipMemcpyAsync(h_words, words, nwords * 8, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        *seed_out = seed;
        *seg_out = seg;
        return SHGPU_OK;
    }
#undef RT_HIP
    err = "retrieval: peeling failed for 64 seeds (duplicate keys?)";
    return SHGPU_ERR_SEARCH_LIMIT;
}

int shgpu_leaf_orientation_impl(hipStream_t s, int bipartite, const uint64_t *d_keys, const shgpu_orient_task *d_tasks,
                                uint64_t n, uint64_t *d_out_keys, uint8_t *d_out_bits, std::string &err) {
    if (n == 0) return SHGPU_OK;
    RtBuf flag;
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
