"""The workload configurations other than the bench's, at their stated sizes where the CPU oracle can check
them in seconds, otherwise on a stated subset: GPU engine / class templates against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import host, workload

pytestmark = pytest.mark.gpu


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
    0 (benchmark/bijections/ShockHash1.h): every GPU seed satisfies the leaf predicate, and on a sample of
    2,000 leaves it is the oracle's minimal seed."""
    log("config 2: 200k leaves of 40")
    n, m = 200_000, 40
    keys = workload.isolated_leaves(n, m, seed=11)
    engine.set_tuning(max_trials=1 << 40)
    got = engine.sh1_leaf_search(keys.reshape(-1), np.arange(n) * m, np.full(n, m), np.zeros(n, np.uint64))
    log("  validity of every seed")
    assert O.sh1_valid_batch(keys, got, threads=16).all()
    pick = np.random.default_rng(12).choice(n, 2000, replace=False)
    log("  oracle on 2000 leaves")
    want = O.sh1_leaf_batch(keys[pick], start=0, threads=16)
    assert np.array_equal(got[pick], want)


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
