"""The drop-in class templates (C++, libshockhash_host.so) built on the GPU engine, against the oracle:
descriptor words bit-exact, retrieval input identical, every query identical, MPHF property."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import host, workload

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_readme_known_answer():
    """README.md:42-46: ShockHash<30,false> on {abc, def, 123, 456}, bucket 2000 -> 1 3 2 0."""
    log("README KAT")
    f = host.MPHF("sh1", 30, [b"abc", b"def", b"123", b"456"], 2000)
    assert list(f([b"abc", b"def", b"123", b"456"])) == [1, 3, 2, 0]


@pytest.mark.parametrize("variant,leaf,n", [("sh2", 64, 40000), ("sh1", 30, 30000), ("simd", 24, 30000),
                                            ("sh2", 16, 25000), ("sh1", 8, 9000), ("simd", 8, 9000),
                                            ("sh2", 100, 30000), ("sh2", 128, 30000)])
def test_class_matches_oracle(variant, leaf, n):
    log(f"class {variant} L={leaf} n={n}")
    h = workload.fingerprints(n, seed=31)
    f = host.MPHF(variant, leaf, h, 2000, threads=4, keep_retrieval_input=True)
    ob = O.OracleBuild(variant, leaf, h, 2000, threads=8)
    assert np.array_equal(f.codes, ob.node_code)
    assert np.array_equal(f.descriptor_words, ob.words)
    k, b = f.ribbon_input
    assert np.array_equal(k, ob.ribbon_keys[: len(k)]) and np.array_equal(b, ob.ribbon_bits[: len(b)])
    q = f(h)
    assert np.array_equal(q, ob.query(h))
    assert len(np.unique(q)) == n and int(q.max()) < n
    log(f"  ok {f.timings_ms}")


@pytest.mark.parametrize("variant,leaf", [("sh2", 64), ("sh1", 30), ("simd", 16)])
@pytest.mark.parametrize("n,bucket", [(1, 2000), (2, 2000), (3, 2000), (65, 2000), (500, 2000), (1000, 8),
                                      (3000, 100)])
def test_class_small_and_ragged(variant, leaf, n, bucket):
    """Tiny key sets and small buckets (many buckets of 0, 1 or 2 keys): codes, descriptor words, retrieval
    input and queries equal the oracle's."""
    log(f"class {variant} L={leaf} n={n} bucket={bucket}")
    h = workload.fingerprints(n, seed=32)
    f = host.MPHF(variant, leaf, h, bucket, threads=1, keep_retrieval_input=True)
    ob = O.OracleBuild(variant, leaf, h, bucket, threads=1)
    assert np.array_equal(f.codes, ob.node_code)
    assert np.array_equal(f.descriptor_words, ob.words)
    k, b = f.ribbon_input
    assert np.array_equal(k, ob.ribbon_keys[: len(k)]) and np.array_equal(b, ob.ribbon_bits[: len(b)])
    q = f(h)
    assert np.array_equal(q, ob.query(h))
    assert len(np.unique(q)) == n and int(q.max()) < n


def test_string_keys_shockhash2():
    log("strings sh2")
    keys = workload.random_strings(20000, seed=4)
    f = host.MPHF("sh2", 24, keys, 2000, threads=4)
    q = f(keys)
    assert len(np.unique(q)) == len(keys) and int(q.max()) < len(keys)
    hashes = np.array([O.spooky128(k, 0) for k in keys], dtype=np.uint64)
    ob = O.OracleBuild("sh2", 24, hashes, 2000, threads=8)
    assert np.array_equal(q, ob.query(hashes))


def test_string_keys_buffers_kept_between_builds():
    """String constructions one after another reuse (and grow) the engine's page-locked and device buffers: a
    small set, a larger one, then smaller ones again (one with empty and long keys, one of a single key), each
    equal to the oracle's build over its SpookyV2 fingerprints."""
    rng = np.random.default_rng(41)
    sets = [workload.random_strings(3000, seed=42), workload.random_strings(60000, seed=43),
            [b""] + [rng.integers(0, 256, size=int(L), dtype=np.uint8).tobytes() for L in rng.integers(1, 700, 2000)],
            [b"only"], workload.random_strings(5000, seed=44)]
    for keys in sets:
        keys = list(dict.fromkeys(keys))  # distinct keys
        log(f"strings sh1 n={len(keys)}")
        f = host.MPHF("sh1", 30, keys, 2000)
        hashes = np.array([O.spooky128(k, 0) for k in keys], dtype=np.uint64).reshape(-1, 2)
        ob = O.OracleBuild("sh1", 30, hashes, 2000, threads=8)
        assert np.array_equal(f.codes, ob.node_code)
        assert np.array_equal(f(keys), ob.query(hashes))
        f.close()


@pytest.mark.parametrize("variant,leaf", [("sh2", 64), ("sh1", 30), ("simd", 24), ("sh2", 100)])
def test_save_load_round_trip(tmp_path, variant, leaf):
    """save() then load() into a fresh object of the class: every query of the reloaded function equals the
    original's (and the oracle's), and the file holds the function's bits."""
    log(f"save/load {variant} L={leaf}")
    h = workload.fingerprints(30000, seed=33)
    f = host.MPHF(variant, leaf, h, 2000)
    path = tmp_path / "f.bin"
    f.save(path)
    g = host.MPHF.load(variant, leaf, path)
    q = f(h)
    assert np.array_equal(g(h), q)
    assert np.array_equal(q, O.OracleBuild(variant, leaf, h, 2000, threads=8).query(h))
    assert path.stat().st_size * 8 <= f.bits + 1024
    with pytest.raises(RuntimeError):
        host.MPHF.load(variant, 16 if leaf != 16 else 8, path)
