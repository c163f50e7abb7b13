// libShockHash2Precompiled.so: the reference's compiled ShockHash2 seam (src/ShockHash2-precompiled.cpp:6-49) on
// the MI355X engine. The reference dispatches param to a template instance per leaf size 2..128 and runs
// BijectionsShockHash2::findSeed and constructRetrieval on the CPU; here the search runs on the GPU
// (shgpu_sh2_leaf_search through the process's engine) and the retrieval input is the host replay of the leaf's
// cuckoo insertions.
#include "shockhash/ShockHash2-precompiled.h"

#include <stdexcept>
#include <string>

#include "shockhash/RecSplitHost.h"

namespace shockhash {

namespace {
constexpr size_t MAX_SEAM_LEAF = 128;  // the reference's dispatchLeafSize<128>
}

size_t shockhash2query(size_t param, size_t seed, uint64_t key, size_t retrieved) {
    if (param <= 1 || param > MAX_SEAM_LEAF) return 0;  // dispatchLeafSizeQ<I <= 1> and past the largest instance
    return (size_t)sh2Cell((uint32_t)param, seed, key, retrieved);
}

size_t shockhash2construct(size_t param, std::vector<uint64_t> &leafKeys,
                           std::vector<std::pair<uint64_t, uint8_t>> &ribbonInput) {
    if (param <= 1 || param > MAX_SEAM_LEAF) return 0;
    if (leafKeys.size() != param)
        throw std::invalid_argument("shockhash2construct: " + std::to_string(leafKeys.size()) + " keys for a leaf of " +
                                    std::to_string(param));
    const uint64_t code = gpu::Engine::instance().sh2LeafCode(leafKeys.data(), (uint32_t)param);
    if (!sh2LeafRetrieval(leafKeys.data(), (uint32_t)param, code, ribbonInput))
        throw std::logic_error("Should be possible to construct");  // the reference's message (ShockHash2-internal.h:561)
    return code;
}

}  // namespace shockhash
