"""String fingerprints on the device (shgpu_hash_strings, SpookyHash V2 128-bit, seed 0) against the oracle's
SpookyHash on every length 0..400 (the short path below 192 bytes with each tail size, the 96-byte-block path
above) and on random keys; string builds through the class templates are checked in test_gpu_configs.py."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload

pytestmark = pytest.mark.gpu


def test_every_length_0_to_400(engine):
    rng = np.random.default_rng(50)
    strings = [rng.integers(0, 256, size=L, dtype=np.uint8).tobytes() for L in range(401) for _ in range(3)]
    assert np.array_equal(engine.hash_strings(strings), O.spooky128_batch(strings))


def test_random_ascii_keys(engine):
    strings = workload.random_strings(200_000, seed=51)
    assert np.array_equal(engine.hash_strings(strings), O.spooky128_batch(strings, threads=16))


def test_empty_batch(engine):
    assert engine.hash_strings([]).shape == (0, 2)
