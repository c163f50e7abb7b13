"""The engine's host-to-device copies: large copies from pageable memory go through page-locked chunks filled by
host threads (sizes at, below and above the staging threshold, not multiples of the chunk, one after another so
that the chunks are reused), and must arrive byte for byte."""
import numpy as np
import pytest

from shockhash_amd import engine as eng

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.mark.parametrize("nbytes", [8, 16 * MIB - 8, 16 * MIB, 16 * MIB + 8, 37 * MIB + 24, 200 * MIB + 4096 + 16])
def test_upload_sizes(engine, nbytes):
    print(f"upload {nbytes} bytes", flush=True)
    rng = np.random.default_rng(nbytes % 1000)
    for rep in range(2):
        src = rng.integers(0, 2**63, size=nbytes // 8, dtype=np.uint64)
        d = eng.DeviceArray.from_host(engine, src)
        assert np.array_equal(d.download(), src)
        d.free()


def test_upload_from_an_offset_view(engine):
    """A source that does not start on a page or chunk boundary."""
    rng = np.random.default_rng(7)
    big = rng.integers(0, 2**63, size=(40 * MIB) // 8 + 3, dtype=np.uint64)
    src = big[3:]
    d = eng.DeviceArray.from_host(engine, src)
    assert np.array_equal(d.download(), src)
    d.free()
