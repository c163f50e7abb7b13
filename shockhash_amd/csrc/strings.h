// Device fingerprints of string keys (strings.hip), called by the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

// SpookyHash V2 128-bit (seed 0) of n keys (bytes d_data[d_offsets[i], d_offsets[i + 1])) into d_hashes
// (n (first, second) pairs) on stream s. Returns an SHGPU status, message in err.
int shgpu_hash_strings_impl(hipStream_t s, const uint8_t *d_data, const uint64_t *d_offsets, uint64_t n, uint64_t *d_hashes,
                            std::string &err);
