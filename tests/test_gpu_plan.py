"""Forest planned on the device (persistent builds): the library compares every node with the host's plan
(SHGPU_TUNE_CHECK_PLAN), the codes equal the oracle's, and a set holding a bucket above 2999 keys is
reported without being searched."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload
from shockhash_amd.engine import ERR_BUCKET_TOO_LARGE, ShgpuError, plan_forest

pytestmark = pytest.mark.gpu

SPLITS_FIRST, CHECK_PLAN = 2, 4


def log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("flags", [CHECK_PLAN, CHECK_PLAN | SPLITS_FIRST])
@pytest.mark.parametrize("variant,leaf,bucket,n", [("sh2", 64, 2000, 100000), ("sh1", 30, 2000, 40000),
                                                   ("simd", 24, 1000, 20000), ("sh1", 2, 100, 3000),
                                                   ("sh2", 2, 100, 3000), ("sh1", 1, 100, 2000),
                                                   ("simd", 3, 100, 3000), ("sh2", 16, 500, 20000)])
def test_device_plan_equals_host_plan(engine, variant, leaf, bucket, n, flags):
    log(f"device plan {variant} L={leaf} b={bucket} n={n} flags {flags}")
    engine.set_tuning(max_trials=1 << 26, flags=flags)
    try:
        h = workload.fingerprints(n, seed=31)
        b = engine.build(h, variant, leaf, bucket)
    finally:
        engine.set_tuning(flags=0)
    ob = O.OracleBuild(variant, leaf, h, bucket, threads=8)
    assert np.array_equal(b.size_acc, ob.size_acc)
    assert np.array_equal(b.nodes["m"], ob.node_m)
    assert np.array_equal(b.codes, ob.node_code)
    host = plan_forest(variant, leaf, b.size_acc)
    for f in ("offset", "m", "kind", "n_children", "unit", "subtree", "start_seed"):
        assert np.array_equal(b.nodes[f], host[f]), f


def one_bucket_overfull(n, seed):
    """n fingerprints whose .first all lie in the lower half: with bucket size 2000 and n <= 4000 the first
    of the two buckets holds every key."""
    h = workload.fingerprints(n, seed=seed)
    h[:, 0] >>= np.uint64(1)
    return h


def test_bucket_too_large_is_reported(engine):
    log("bucket too large")
    h = one_bucket_overfull(3500, seed=32)
    with pytest.raises(ShgpuError) as ei:
        engine.build(h, "sh2", 64, 2000)
    assert ei.value.status == ERR_BUCKET_TOO_LARGE
    ok = workload.fingerprints(5000, seed=33)
    b = engine.build(ok, "sh2", 64, 2000)
    assert np.array_equal(b.codes, O.OracleBuild("sh2", 64, ok, 2000, threads=8).node_code)


@pytest.mark.parametrize("groups", [5000, 500])
def test_sort_runs_sharing_the_sorted_bits(engine, groups):
    """The radix sort orders on the top 24 key bits at n = 100,000 and a fix-up pass orders the keys that share
    them. Fingerprints whose .first take only `groups` distinct top-24-bit values (spread over the range) form
    runs of ~20 keys (fix-up) or ~200 keys (longer than the fix-up handles: the build sorts on all 64 bits).
    Bucket sizes, leaf-ordered keys and codes equal the oracle's."""
    log(f"sort runs, {groups} groups")
    n = 100_000
    rng = np.random.default_rng(34 + groups)
    top = rng.integers(0, groups, n, dtype=np.uint64) * np.uint64((1 << 24) // groups)
    h = workload.fingerprints(n, seed=35)
    h[:, 0] = (top << np.uint64(40)) | (h[:, 0] >> np.uint64(24))
    assert len(np.unique(h[:, 0])) == n
    b = engine.build(h, "sh2", 64, 2000)
    ob = O.OracleBuild("sh2", 64, h, 2000, threads=8)
    assert np.array_equal(b.size_acc, ob.size_acc)
    assert np.array_equal(b.codes, ob.node_code)
    # ShockHash2 hands the keys of every leaf to the retrieval structure in final leaf order (a part of one key
    # is no leaf)
    leaf = b.nodes["kind"] == 3
    idx = np.concatenate([np.arange(o, o + m) for o, m in zip(b.nodes["offset"][leaf], b.nodes["m"][leaf])])
    assert np.array_equal(b.keys(n)[idx], ob.ribbon_keys[: len(idx)])
