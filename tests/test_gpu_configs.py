"""The workload configurations other than the bench's, at their stated sizes where the CPU oracle can check
them in seconds, otherwise on a stated subset: GPU engine / class templates against the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import host, workload

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def log(msg):
    print(msg, flush=True)


def test_config1_shockhash30_strings_1m():
    """ShockHash<30,false>, bucket 2000, 1,000,000 distinct random ASCII strings of length 10..50 (seed 3),
    fingerprinted with SpookyV2 as the string constructor does: node codes, descriptor words, retrieval
    input and every one of the 1,000,000 queries equal the oracle's; the queries form a bijection."""
    log("config 1: generating 1M strings")
    strings = workload.random_strings(1_000_000, seed=3)
    log("  building on the GPU (class template)")
    f = host.MPHF("sh1", 30, strings, 2000, threads=8, keep_retrieval_input=True)
    log("  oracle")
    h = O.spooky128_batch(strings, threads=16)
    ob = O.OracleBuild("sh1", 30, h, 2000, threads=16)
    assert np.array_equal(f.codes, ob.node_code)
    assert np.array_equal(f.descriptor_words, ob.words)
    k, b = f.ribbon_input
    assert np.array_equal(k, ob.ribbon_keys[: len(k)]) and np.array_equal(b, ob.ribbon_bits[: len(b)])
    log("  queries")
    q = f(strings)
    assert np.array_equal(q, ob.query(h))
    assert len(np.unique(q)) == len(strings) and int(q.max()) < len(strings)


def test_config2_isolated_leaves_200k(engine):
    """200,000 isolated ShockHash1 leaves of 40 keys (the .second of uniform fingerprints, seed 11), seeds from
    0 (benchmark/bijections/ShockHash1.h:13-35): every GPU seed equals the oracle's minimal seed, all 200,000 of
    them (tests/golden/config2_seeds.npz, written by tests/golden/make_config2_seeds.py)."""
    log("config 2: 200k leaves of 40")
    gold = np.load(os.path.join(GOLDEN, "config2_seeds.npz"))
    n, m = int(gold["n_leaves"]), int(gold["m"])
    keys = workload.isolated_leaves(n, m, seed=int(gold["seed"]))
    engine.set_tuning(max_trials=1 << 40)
    got = engine.sh1_leaf_search(keys.reshape(-1), np.arange(n) * m, np.full(n, m), np.zeros(n, np.uint64))
    bad = np.nonzero(got != gold["seeds"])[0]
    assert len(bad) == 0, f"{len(bad)} leaves differ, first {bad[:5]}: got {got[bad[:5]]} want {gold['seeds'][bad[:5]]}"


def test_config3_simd58_two_buckets():
    """SIMDShockHash<58,false>, bucket 2000: the first 4,000 fingerprints of the 10M workload's generator
    (SplitMix64 seed 1), i.e. two buckets of leaves of up to 58 keys (~10^9 hash evaluations each on the CPU);
    codes, descriptor words, retrieval input and queries equal the oracle's."""
    log("config 3 subset: SIMDShockHash L58, 4000 keys")
    h = workload.fingerprints(4000, seed=1)
    f = host.MPHF("simd", 58, h, 2000, threads=4)
    ob = O.OracleBuild("simd", 58, h, 2000, threads=16)
    assert np.array_equal(f.codes, ob.node_code)
    assert np.array_equal(f.descriptor_words, ob.words)
    q = f(h)
    assert np.array_equal(q, ob.query(h))
    assert len(np.unique(q)) == len(h)


@pytest.mark.timeout(900)
def test_config3_simd58_10m_full_build():
    """SIMDShockHash<58,false>, bucket 2000, all 10,000,000 fingerprints of the config-3 workload (SplitMix64
    seed 1), ~10^14 hash evaluations on the GPU: the node codes of 32 buckets spread over the partition equal
    the oracle's (tests/golden/config3_simd58_buckets.npz, written by tests/golden/make_config3_buckets.py, a few
    core-minutes per bucket), and the 10,000,000 queries are distinct and below n."""
    from shockhash_amd.engine import plan_forest
    gold = np.load(os.path.join(GOLDEN, "config3_simd58_buckets.npz"))
    n, leaf, bsize = int(gold["n"]), int(gold["leaf"]), int(gold["bucket_size"])
    log(f"config 3: SIMDShockHash L{leaf}, {n} keys (about a minute on the GPU)")
    h = workload.fingerprints(n, seed=int(gold["seed"]))
    f = host.MPHF("simd", leaf, h, bsize, threads=1)
    log(f"  built in {f.timings_ms['constructor'] / 1e3:.1f} s; comparing {len(gold['buckets'])} buckets")
    acc = workload.bucket_sizes(h, bsize)
    nodes = plan_forest("simd", leaf, acc)
    codes = f.codes
    assert len(codes) == len(nodes)
    root = {}
    i = 0
    for b in range(len(acc) - 1):
        if acc[b + 1] - acc[b] > 1:
            root[b] = (i, int(nodes["subtree"][i]))
            i += int(nodes["subtree"][i])
    assert i == len(nodes)
    for b in gold["buckets"]:
        sel = gold["node_bucket"] == b
        start, cnt = root[int(b)]
        assert cnt == int(sel.sum()), f"bucket {b}: node count"
        assert np.array_equal(nodes["m"][start:start + cnt], gold["node_m"][sel]), f"bucket {b}: tree shape"
        assert np.array_equal(codes[start:start + cnt], gold["node_code"][sel]), f"bucket {b}: codes"
    log("  queries")
    q = f(h)
    assert int(q.max()) < n and len(np.unique(q)) == n

