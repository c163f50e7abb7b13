"""ROTATION_FITTING on the GPU (include/ShockHash.h:316-403,251-265; include/SIMDShockHash.hpp:419-506,826-840):
whole builds with rotation-fitted full leaves against the oracle's restatement (node codes x * leaf + r, bucket
sizes), both scheduling forms, and the class templates ShockHash<L, true> / SIMDShockHash<L, true>: descriptor
words, retrieval input, queries, save/load."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import host, workload

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


# leaf sizes on both sides of 32 (group A's seed mask is 7 above 32 keys); the plain form repeats the persistent
# search node by node, so it runs on the shapes up to 40 keys only (one rotation leaf size above 32 suffices)
SHAPES = [("sh1", 8, 6000, 500), ("sh1", 16, 12000, 1000), ("sh1", 30, 20000, 2000), ("sh1", 40, 20000, 2000),
          ("simd", 16, 12000, 1000), ("simd", 58, 12000, 2000)]


@pytest.mark.parametrize("mode,variant,leaf,n,bucket",
                         [("persistent",) + s for s in SHAPES] + [("plain",) + s for s in SHAPES if s[1] <= 40])
def test_rotation_build_matches_oracle(engine, variant, leaf, n, bucket, mode):
    log(f"rotation build {variant} L={leaf} n={n} bucket={bucket} mode={mode}")
    engine.set_tuning(max_trials=1 << 26)
    h = workload.fingerprints(n, seed=41 + leaf)
    ob = O.OracleBuild(variant, leaf, h, bucket, threads=8, rotation=True)
    log("  oracle done")
    b = engine.build(h, variant, leaf, bucket, mode=mode, rotation=True)
    assert np.array_equal(b.size_acc, ob.size_acc)
    assert np.array_equal(b.nodes["m"], ob.node_m)
    full = (b.nodes["kind"] == 3) & (b.nodes["m"] == leaf)
    assert full.sum() > 0
    bad = np.nonzero(b.codes != ob.node_code)[0]
    assert len(bad) == 0, f"{len(bad)} nodes differ, first {bad[:5]}: got {b.codes[bad[:5]]} want {ob.node_code[bad[:5]]}"
    log(f"  ok, {int(full.sum())} rotation leaves, timings {b.timings_ms}")


@pytest.mark.parametrize("variant,leaf,n", [("sh1", 30, 30000), ("sh1", 40, 20000), ("sh1", 8, 9000),
                                            ("simd", 16, 20000), ("simd", 58, 12000)])
def test_rotation_class_matches_oracle(variant, leaf, n, tmp_path):
    log(f"rotation class {variant} L={leaf} n={n}")
    h = workload.fingerprints(n, seed=51 + leaf)
    f = host.MPHF(variant, leaf, h, 2000, threads=4, keep_retrieval_input=True, rotation=True)
    ob = O.OracleBuild(variant, leaf, h, 2000, threads=8, rotation=True)
    assert np.array_equal(f.codes, ob.node_code)
    assert np.array_equal(f.descriptor_words, ob.words)
    k, b = f.ribbon_input
    assert np.array_equal(k, ob.ribbon_keys) and np.array_equal(b, ob.ribbon_bits)
    q = f(h)
    assert np.array_equal(q, ob.query(h))
    assert len(np.unique(q)) == n and int(q.max()) < n
    path = tmp_path / "rot.shf"
    f.save(path)
    g = host.MPHF.load(variant, leaf, path, rotation=True)
    assert np.array_equal(g(h), q)
    with pytest.raises(RuntimeError):
        host.MPHF.load(variant, leaf, path, rotation=False)  # another class: ROTATION_FITTING differs
    log(f"  ok {f.timings_ms}")


def test_rotation_string_keys():
    """String constructor of ShockHash<30, true>: SpookyV2 fingerprints on the device, queries by string."""
    keys = workload.random_strings(20000, seed=6)
    f = host.MPHF("sh1", 30, keys, 2000, threads=4, rotation=True)
    q = f(keys)
    assert len(np.unique(q)) == len(keys) and int(q.max()) < len(keys)
    hashes = np.array([O.spooky128(k, 0) for k in keys], dtype=np.uint64)
    ob = O.OracleBuild("sh1", 30, hashes, 2000, threads=8, rotation=True)
    assert np.array_equal(q, ob.query(hashes))
