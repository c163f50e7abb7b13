"""One-bit retrieval structure built on the device (shgpu_ribbon_build, a sharded ribbon): every key of the set
retrieves its bit. Queries are answered here from the documented layout with numpy (independent of the library's
code); the class templates' queries over the same structure are checked against the oracle in test_gpu_classes.py
and test_gpu_configs.py."""
import numpy as np
import pytest

from shockhash_amd import workload
from shockhash_amd.engine import ERR_SEARCH_LIMIT, ShgpuError, load

pytestmark = pytest.mark.gpu

SALT_SHARD, SALT_ROW, STEP = np.uint64(0x632BE59BD9B4E019), np.uint64(0xD6E8FEB86659FD93), np.uint64(0x9E3779B97F4A7C15)


def log(msg):
    print(msg, flush=True)


def remix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def fastrange64(h, n):
    """floor(h * n / 2^64) for n < 2^32 (n may be an array)."""
    hi, lo = h >> np.uint64(32), h & np.uint64(0xFFFFFFFF)
    n = np.asarray(n, dtype=np.uint64)
    return (hi * n + ((lo * n) >> np.uint64(32))) >> np.uint64(32)


def parity(x):
    for s in (32, 16, 8, 4, 2, 1):
        x = x ^ (x >> np.uint64(s))
    return x & np.uint64(1)


def retrieve(keys, offsets, seeds, words):
    with np.errstate(over="ignore"):
        h = remix(keys + SALT_SHARD)
        s = fastrange64(h, len(seeds))
        base = offsets[s].astype(np.uint64)
        m = offsets[s + 1].astype(np.uint64) - base
        g = remix(h + (seeds[s].astype(np.uint64) + np.uint64(1)) * STEP)
        p = base + fastrange64(g, m - np.uint64(63))
        row = remix(g ^ SALT_ROW) | np.uint64(1)
    w, b = p >> np.uint64(6), p & np.uint64(63)
    lo = words[w] >> b
    hi = np.where(b > 0, words[w + np.uint64(1)] << ((np.uint64(64) - b) & np.uint64(63)), np.uint64(0))
    return parity(row & (lo | hi)).astype(np.uint8)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, 100_000, 2_000_000])
def test_every_key_retrieves_its_bit(engine, n):
    log(f"ribbon n={n}")
    keys = workload.fingerprints(max(n, 1), seed=40)[:n, 1].copy()
    assert len(np.unique(keys)) == n
    bits = np.random.default_rng(41).integers(0, 2, n, dtype=np.uint8)
    offsets, seeds, words = engine.ribbon_build(keys, bits)
    lib = load()
    S = lib.shgpu_ribbon_shards(n)
    assert len(seeds) == S and len(offsets) == S + 1
    m = np.diff(offsets.astype(np.int64))
    counts = np.bincount(fastrange64(remix(keys + SALT_SHARD), S).astype(np.int64), minlength=S) if n else np.zeros(S, int)
    want_m = np.where(counts == 0, 0, np.maximum(64, counts + (counts + 19) // 20))
    assert np.array_equal(m, want_m)  # slots per shard follow the documented rule
    assert len(words) == (int(offsets[-1]) + 63) // 64 + 1
    if n:
        assert np.array_equal(retrieve(keys, offsets, seeds, words), bits)
    if n >= 100_000:
        bits_per_key = (64 * len(words) + 32 * len(offsets) + 8 * len(seeds)) / n
        log(f"  {bits_per_key:.4f} bits/key, seeds used up to {int(seeds.max())}")
        assert bits_per_key < 1.08


def test_all_zero_and_all_one_bits(engine):
    keys = workload.fingerprints(50_000, seed=42)[:, 1].copy()
    for b in (0, 1):
        bits = np.full(len(keys), b, dtype=np.uint8)
        offsets, seeds, words = engine.ribbon_build(keys, bits)
        assert np.array_equal(retrieve(keys, offsets, seeds, words), bits)


def test_repeated_key_with_different_bits_is_reported(engine):
    """A key present twice with different bits makes its shard inconsistent for every seed byte."""
    keys = workload.fingerprints(1000, seed=43)[:, 1].copy()
    keys[7] = keys[3]
    bits = np.zeros(1000, dtype=np.uint8)
    bits[7] = 1
    with pytest.raises(ShgpuError) as ei:
        engine.ribbon_build(keys, bits)
    assert ei.value.status == ERR_SEARCH_LIMIT


def test_repeated_key_with_equal_bits_retrieves_it(engine):
    """A repeated key with the same bit is a redundant equation: the structure still returns every bit."""
    keys = workload.fingerprints(1000, seed=44)[:, 1].copy()
    keys[:300] = keys[500]
    bits = np.random.default_rng(45).integers(0, 2, 1000, dtype=np.uint8)
    bits[:300] = bits[500]
    offsets, seeds, words = engine.ribbon_build(keys, bits)
    assert np.array_equal(retrieve(keys, offsets, seeds, words), bits)
