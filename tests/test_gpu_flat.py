"""ShockHash2Flat<k> on the GPU (include/ShockHash2Flat.h:37-302): every bucket's base case searched in one batched
ShockHash2 leaf search, its orientation and the retrieval structure built on the device; the function's value of
every key equals the oracle's restatement, for sizes that are and are not multiples of k."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import host, workload

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("k,n", [(16, 16 * 600), (32, 20011), (64, 64 * 1500), (100, 100037), (128, 128 * 120),
                                 (16, 7), (16, 17), (32, 31), (64, 64 * 2 + 5), (64, 64 * 11)])
def test_flat_matches_oracle(k, n):
    log(f"flat k={k} n={n}")
    mh = workload.fingerprints(n, seed=300 + k)[:, 1].copy()
    of = O.OracleFlat(k, mh)
    log(f"  oracle: {of.nbuckets} buckets, {of.bumped} keys in the fallback, {of.seed_fallbacks} seed fallbacks")
    f = host.FlatMPHF(k, mh)
    info = f.info
    assert info["nbuckets"] == of.nbuckets and info["layers"] == of.layers
    assert info["seed_fallbacks"] == of.seed_fallbacks
    q = f(mh)
    want = of.query(mh)
    bad = np.nonzero(q != want)[0]
    assert len(bad) == 0, f"{len(bad)} keys differ, e.g. {q[bad[:5]]} vs {want[bad[:5]]}"
    assert len(np.unique(q)) == n
    assert len(f.codes) == of.nbuckets
    log(f"  phases {f.phases_ms}")
    log(f"  ok {info}")


def test_flat_base_case_codes_equal_leaf_search():
    """The bucket codes are ShockHash2's leaf codes of the buckets' keys (BijectionsShockHash2::findSeed): checked
    on the oracle's leaf search for the first buckets via the keys' bucket assignment."""
    k, n = 64, 64 * 300
    mh = workload.fingerprints(n, seed=17)[:, 1].copy()
    of = O.OracleFlat(k, mh)
    f = host.FlatMPHF(k, mh)
    bucket, _ = of.evaluate(mh)
    codes = f.codes
    for b in range(20):
        keys = mh[bucket == b]  # input order, as the reference fills its buckets
        assert len(keys) == k
        assert codes[b] == O.sh2_find_seed(keys)


def test_flat_layers_rejects_bad_mappings(engine):
    """shgpu_flat_layers checks the threshold mapping before any work: 2..64 non-decreasing values ending at
    2^32 - 1 (the reference's mapping has that shape, include/ShockHash2Flat.h:17-35)."""
    import ctypes
    from shockhash_amd import engine as eng
    lib = eng.load()
    n, k = 64 * 10, 64
    d = eng.DeviceArray.from_host(engine, workload.fingerprints(n, seed=5)[:, 1].copy())
    thr, free = np.zeros(n, np.uint8), np.zeros(n, np.uint32)
    bases, layers, nb = np.zeros(4, np.uint64), ctypes.c_uint32(), ctypes.c_uint64()
    bm, bi = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def run(mapping):
        m = np.asarray(mapping, dtype=np.uint32)
        return lib.shgpu_flat_layers(engine.h, d.ptr, n, k, P(m), len(m), P(thr), P(free), P(bases), ctypes.byref(layers),
                                     P(bm), P(bi), ctypes.byref(nb))
    assert run([(i + 1) * (1 << 28) - 1 for i in range(16)]) == 0
    for bad in ([0xFFFFFFFF], [5, 0xFFFFFFFE], [9, 3, 0xFFFFFFFF], [0] * 64 + [0xFFFFFFFF]):
        assert run(bad) == eng.ERR_ARG, bad
    d.free()
