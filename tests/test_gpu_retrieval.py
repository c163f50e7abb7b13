"""One-bit retrieval structure built on the device (shgpu_retrieval_build): every key of the set retrieves its
bit. Queries are answered here from the documented layout with numpy (independent of the library's code);
the class templates' queries over the same structure are checked against the oracle in test_gpu_classes.py and
test_gpu_configs.py."""
import numpy as np
import pytest

from shockhash_amd import workload
from shockhash_amd.engine import ERR_ARG, ERR_SEARCH_LIMIT, ShgpuError, load

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def log(msg):
    print(msg, flush=True)


def remix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def fastrange64(h, n):
    """floor(h * n / 2^64) for n < 2^32."""
    hi, lo = h >> np.uint64(32), h & np.uint64(0xFFFFFFFF)
    n = np.uint64(n)
    return (hi * n + ((lo * n) >> np.uint64(32))) >> np.uint64(32)


def retrieve(keys, seed, seg, words):
    with np.errstate(over="ignore"):
        h0 = remix(keys + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019))
        h1 = remix(h0 ^ np.uint64(0xD6E8FEB86659FD93))
        h2 = remix(h1 + np.uint64(0x8CB92BA72F3D8DD7))
    seg = np.uint64(seg)
    cells = [fastrange64(h0, seg), seg + fastrange64(h1, seg), np.uint64(2) * seg + fastrange64(h2, seg)]
    out = np.zeros(len(keys), dtype=np.uint64)
    for c in cells:
        out ^= (words[c >> np.uint64(6)] >> (c & np.uint64(63))) & np.uint64(1)
    return out.astype(np.uint8)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, 100_000, 2_000_000])
def test_every_key_retrieves_its_bit(engine, n):
    log(f"retrieval n={n}")
    keys = workload.fingerprints(max(n, 1), seed=40)[:n, 1].copy()
    assert len(np.unique(keys)) == n
    bits = np.random.default_rng(41).integers(0, 2, n, dtype=np.uint8)
    seed, seg, words = engine.retrieval_build(keys, bits)
    lib = load()
    assert seg == lib.shgpu_retrieval_cells_per_third(n)
    assert len(words) == (3 * seg + 63) // 64
    if 3 * seg % 64:
        assert int(words[-1]) >> (3 * seg % 64) == 0  # cells past the last are zero
    assert np.array_equal(retrieve(keys, seed, seg, words), bits)


def test_all_zero_and_all_one_bits(engine):
    keys = workload.fingerprints(50_000, seed=42)[:, 1].copy()
    for b in (0, 1):
        bits = np.full(len(keys), b, dtype=np.uint8)
        seed, seg, words = engine.retrieval_build(keys, bits)
        assert np.array_equal(retrieve(keys, seed, seg, words), bits)


def test_duplicate_keys_are_reported(engine):
    keys = workload.fingerprints(1000, seed=43)[:, 1].copy()
    keys[7] = keys[3]
    with pytest.raises(ShgpuError) as ei:
        engine.retrieval_build(keys, np.zeros(1000, dtype=np.uint8))
    assert ei.value.status == ERR_SEARCH_LIMIT


def test_key_repeated_past_the_cell_limit_is_reported(engine):
    """A key repeated 300 times puts 300 keys in its cells, past the 255 that a cell's index sum can count:
    reported as an argument error at once, not searched over every seed."""
    keys = workload.fingerprints(1000, seed=44)[:, 1].copy()
    keys[:300] = keys[500]
    with pytest.raises(ShgpuError) as ei:
        engine.retrieval_build(keys, np.zeros(1000, dtype=np.uint8))
    assert ei.value.status == ERR_ARG
