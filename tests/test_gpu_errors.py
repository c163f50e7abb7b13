"""Error paths of the engine and the class templates on the GPU: keys that repeat end with a status code
within the node's seed bound (the reference loops forever on them, include/SIMDShockHash.hpp:131,171), and an
allocation that cannot succeed reports SHGPU_ERR_NOMEM and leaves the engine usable. Every kernel in these
tests exits normally; the failing searches stop at their bounds."""
import numpy as np
import pytest

from shockhash_amd import host, workload
from shockhash_amd.engine import ERR_NOMEM, ERR_SEARCH_LIMIT, DeviceArray, ShgpuError

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def with_repeats(n, copies, seed):
    """n fingerprints where fingerprint 5 is repeated `copies` times in total (spread over the array)."""
    h = workload.fingerprints(n, seed=seed)
    for i in range(1, copies):
        h[100 * i] = h[5]
    return h


@pytest.mark.parametrize("variant,leaf", [("sh1", 8), ("simd", 8), ("sh2", 16), ("sh1", 30)])
def test_class_three_equal_keys_report_search_limit(variant, leaf):
    """Three equal fingerprints land in one leaf (splits send equal keys to the same part): three edges between
    the same two cells are never a pseudoforest, so the leaf search runs to its bound, 2^(k(m) + 24) seeds for a
    ShockHash leaf of m keys with Golomb parameter k(m), or the ShockHash2 half-seed/candidate tables, and the
    constructor throws with SHGPU_ERR_SEARCH_LIMIT."""
    log(f"three equal keys, {variant} L={leaf}")
    h = with_repeats(5000, 3, seed=61)
    with pytest.raises(host.ConstructionError) as ei:
        host.MPHF(variant, leaf, h, 2000)
    assert ei.value.status == ERR_SEARCH_LIMIT, str(ei.value)


@pytest.mark.parametrize("variant,leaf", [("sh1", 30), ("sh2", 64)])
def test_class_two_equal_keys_are_reported(variant, leaf):
    """Two equal fingerprints may pass the leaf search (two parallel edges form one cycle) but can never be told
    apart: the retrieval structure never peels them, and the constructor throws with SHGPU_ERR_SEARCH_LIMIT."""
    log(f"two equal keys, {variant} L={leaf}")
    h = with_repeats(5000, 2, seed=62)
    with pytest.raises(host.ConstructionError) as ei:
        host.MPHF(variant, leaf, h, 2000)
    assert ei.value.status == ERR_SEARCH_LIMIT, str(ei.value)


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_engine_small_explicit_bound(engine, mode):
    """With a small explicit max_trials (2^22 seeds) the ShockHash leaf 30 build stops there, in both forms."""
    log(f"explicit bound, {mode}")
    h = with_repeats(5000, 3, seed=63)
    engine.set_variant_tuning("sh1", max_trials=1 << 22)
    try:
        with pytest.raises(ShgpuError) as ei:
            engine.build(h, "sh1", 30, 2000, mode=mode)
    finally:
        engine.set_variant_tuning("sh1", max_trials=1 << 40)
    assert ei.value.status == ERR_SEARCH_LIMIT
    # the engine still builds distinct keys afterwards
    b = engine.build(workload.fingerprints(5000, seed=64), "sh1", 30, 2000, mode=mode)
    assert not np.any(b.codes == np.uint64((1 << 64) - 1))


def test_impossible_allocation_reports_nomem(engine):
    """A device allocation larger than the device's memory returns SHGPU_ERR_NOMEM, frees nothing it did not
    own and leaves the device's free memory as it was; later calls work."""
    free0, total = engine.memory()
    assert 0 < free0 <= total
    with pytest.raises(ShgpuError) as ei:
        DeviceArray(engine, 1 << 50, np.uint8)
    assert ei.value.status == ERR_NOMEM
    free1, _ = engine.memory()
    assert abs(free1 - free0) < (64 << 20)
    b = engine.build(workload.fingerprints(3000, seed=65), "sh2", 64, 2000)
    assert len(b.codes) > 0
