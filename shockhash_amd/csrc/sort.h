// Stable LSD radix sort of n (key, value) u64 pairs by the key bits [0, end_bit) on `stream`. With
// tmp == nullptr only tmp_bytes is set (the scratch size for n pairs). (rocPRIM's sort with a nonzero
// begin_bit and end_bit 64 does not order the keys on this ROCm release: 100k random keys sorted on bits
// [8, 64), [24, 64), [40, 64) or [48, 64) come out with about half of all neighbours inverted, while [0, 64)
// is exact; so callers rotate the bits they sort on down to the bottom.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

hipError_t shgpu_sort_pairs_u64(void *tmp, size_t &tmp_bytes, const uint64_t *keys_in, uint64_t *keys_out,
                                const uint64_t *values_in, uint64_t *values_out, int n, int end_bit,
                                hipStream_t stream);
