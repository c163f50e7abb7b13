// Persistent-build kernels of SIMDShockHash (fmix32 split hash, start seeds 17-19), compiled on their own so the
// variants build in parallel.
#include "build_kernels.h"

namespace shgpu {
template hipError_t launch_build_variant<SHGPU_SIMD_SHOCKHASH>(uint32_t, uint32_t, bool, const BuildArgs &, int, hipStream_t);
template hipError_t launch_node_variant<SHGPU_SIMD_SHOCKHASH>(bool, const BuildArgs &, uint32_t, uint32_t, uint32_t, hipStream_t);
}  // namespace shgpu
