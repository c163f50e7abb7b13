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


def test_page_locked_host_buffer_round_trip(engine):
    """shgpu_host_alloc memory is a valid source and target of the engine's copies (the class templates pack
    string keys into it)."""
    import ctypes
    from shockhash_amd import engine as eng
    lib = eng.load()
    p = ctypes.c_void_p()
    n = 1 << 20
    assert lib.shgpu_host_alloc(engine.h, n * 8, ctypes.byref(p)) == 0 and p.value
    try:
        src = np.random.default_rng(52).integers(0, 2**63, size=n, dtype=np.uint64)
        ctypes.memmove(p.value, src.ctypes.data, n * 8)
        d = eng.DeviceArray(engine, n, np.uint64)
        assert lib.shgpu_copy_to_device(engine.h, d.ptr, p.value, n * 8) == 0
        assert np.array_equal(d.download(), src)
        back = np.frombuffer((ctypes.c_uint64 * n).from_address(p.value), dtype=np.uint64)
        assert lib.shgpu_copy_to_host(engine.h, p.value, d.ptr, 8 * 1000) == 0
        assert np.array_equal(back[:1000], src[:1000])
        d.free()
    finally:
        assert lib.shgpu_host_free(engine.h, p.value) == 0
