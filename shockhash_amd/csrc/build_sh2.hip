// Persistent-build kernels of ShockHash2 (bipartite leaf search), compiled on their own so the
// variants build in parallel.
#include "build_kernels.h"

namespace shgpu {
template hipError_t launch_build_variant<SHGPU_SHOCKHASH2>(uint32_t, uint32_t, bool, const BuildArgs &, int, hipStream_t);
template hipError_t launch_node_variant<SHGPU_SHOCKHASH2>(bool, const BuildArgs &, uint32_t, uint32_t, uint32_t, hipStream_t);
}  // namespace shgpu
