// Device-side primitives shared by every ShockHash seed-search kernel (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shgpu_types.h"

namespace shgpu {

// hipMalloc that leaves no error behind when it fails: a failed allocation sets the runtime's last error, and
// the next hipGetLastError (after a later kernel launch) would report it as that launch's failure.
inline bool device_malloc(void **p, size_t bytes) {
    if (hipMalloc(p, bytes) == hipSuccess) return true;
    *p = nullptr;
    (void)hipGetLastError();
    return false;
}
inline bool host_malloc_pinned(void **p, size_t bytes) {
    if (hipHostMalloc(p, bytes, hipHostMallocDefault) == hipSuccess) return true;
    *p = nullptr;
    (void)hipGetLastError();
    return false;
}

// Stafford Mix13, the 64-bit finaliser the reference calls `remix`
// (in-tree text: include/shockhash/SimdUtils.h:42-46).
__device__ __forceinline__ uint64_t remix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// MurmurHash3 fmix32 (include/SIMDShockHash.hpp:76-83).
__device__ __forceinline__ uint32_t remix32(uint32_t z) {
    z ^= z >> 16; z *= 0x85ebca6bu; z ^= z >> 13; z *= 0xc2b2ae35u; z ^= z >> 16;
    return z;
}

// floor(x * n / 2^64) for n < 2^32 (fastrange64 at ShockHash2-internal.h:84,423,442).
__device__ __forceinline__ uint32_t fastrange64(uint64_t x, uint32_t n) {
    return (uint32_t)__umul64hi(x, (uint64_t)n);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Lane l's 64-bit value (l wave-uniform). The builtin returns int: cast through uint32_t so the
// low word is not sign-extended into the high word.
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)l);
    return ((uint64_t)hi << 32) | lo;
}

// Takes the next index of a device counter for the whole wave without a divergent branch: lane 0
// adds 1, the other lanes add 0, and every lane uses lane 0's old value.
__device__ __forceinline__ uint32_t wave_fetch_add(uint32_t *counter, uint32_t add) {
    const uint32_t old = atomicAdd(counter, __lane_id() == 0 ? add : 0u);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)old);
}

// Orders LDS traffic between the lanes of one wavefront (a wave executes in lock step, LDS
// operations of one wave complete in issue order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 24-bit multiply-add d = (a & 0xFFFFFF) * (b & 0xFFFFFF) + c (mod 2^32), one full-rate VALU op. Exact mod 2^24
// for any a, b, which is all the callers use. The compiler splits mul24 + add into a 32-bit multiply and an
// add3, so the instruction is written out.
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}

// Wave issue priority by how long the wave has held its current task. The SIMD's instruction arbiter
// favours the waves launched first, so in a saturated persistent kernel the last-launched waves progress
// many times slower than the first (measured on a 10M-key build: 0.7 ms mean per 256-key node for the
// first-launched quarter of the grid, 14 ms for the last) and the nodes they hold hold back their
// subtrees. Raising a wave's priority as its task ages makes tasks finish roughly in the order they were
// taken. task_clock() counts at 100 MHz (s_memrealtime).
__device__ __forceinline__ uint64_t task_clock() { return __builtin_amdgcn_s_memrealtime(); }
__device__ __forceinline__ void age_priority(uint64_t t_start) {
    const uint64_t held = task_clock() - t_start;
    if (held > 400000) __builtin_amdgcn_s_setprio(3);       // 4 ms
    else if (held > 200000) __builtin_amdgcn_s_setprio(2);  // 2 ms
    else if (held > 100000) __builtin_amdgcn_s_setprio(1);  // 1 ms
}
__device__ __forceinline__ void reset_priority() { __builtin_amdgcn_s_setprio(0); }

// Keys a wave stages in LDS and reads back as wave-wide broadcasts (every lane reads the same address:
// one LDS instruction, no bank conflicts, no VALU work), so the hash loops spend their VALU issue on
// hashing only.
constexpr uint32_t STAGE_KEYS = 256;

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

}  // namespace shgpu
