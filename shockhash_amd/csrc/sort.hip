// Radix sort of (first, second) pairs by first, the device counterpart of parallelPartition's sort
// (include/shockhash/Sorter.hpp:15-36, src/Sorter.cpp:6-8). Kept in its own translation unit: the rocPRIM
// instantiations dominate the library's compile time.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "sort.h"

hipError_t shgpu_sort_pairs_u64(void *tmp, size_t &tmp_bytes, const uint64_t *keys_in, uint64_t *keys_out,
                                const uint64_t *values_in, uint64_t *values_out, int n, int end_bit, hipStream_t stream) {
    return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, values_in, values_out, (size_t)n, 0u,
                                     (unsigned)end_bit, stream);
}
