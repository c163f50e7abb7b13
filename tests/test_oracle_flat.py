"""The oracle's ShockHash2Flat restatement (include/ShockHash2Flat.h:37-302) on the CPU: its threshold mapping,
and the invariants its construction must meet whatever a slip in the bucketing would break (every bucket holds
exactly k keys, which the build checks; the queries are injective, and a bijection onto [0, N) when N is a multiple
of k; a key's bucket is the first layer whose threshold admits it)."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload

ONE_THIRD = 0xFFFFFFFF // 3


def test_threshold_mapping_shapes():
    """_fill_mapping (ShockHash2Flat.h:17-35): 0 and one third at the bottom, an interpolation from two thirds, all
    thresholds at the top."""
    assert list(O.flat_mapping(2)) == [0, 0xFFFFFFFF]
    assert list(O.flat_mapping(3)) == [0, ONE_THIRD, 0xFFFFFFFF]
    m = O.flat_mapping(32)
    assert m[0] == 0 and m[1] == ONE_THIRD and m[2] == 2 * ONE_THIRD and m[31] == 0xFFFFFFFF
    assert m[30] == 2 * ONE_THIRD + ONE_THIRD * 28 // 29
    assert all(int(a) < int(b) for a, b in zip(m, m[1:]))


@pytest.mark.parametrize("k,n", [(16, 16 * 400), (32, 32 * 500), (64, 64 * 1000), (100, 100 * 300)])
def test_flat_is_a_bijection_for_multiples_of_k(k, n):
    mh = workload.fingerprints(n, seed=100 + k)[:, 1]
    f = O.OracleFlat(k, mh)
    q = f.query(mh)
    assert len(np.unique(q)) == n and int(q.max()) == n - 1
    assert f.layers == 2 and f.nbuckets == n // k and f.bumped > 0


@pytest.mark.parametrize("k,n", [(16, 5003), (64, 100037), (100, 100037), (32, 31), (16, 7)])
def test_flat_ragged_sizes_are_injective(k, n):
    """N not a multiple of k: the reference places the last partial bucket's keys past the buckets that no layer
    handled (ShockHash2Flat.h:130-137), so the values stay distinct but can exceed N - 1."""
    mh = workload.fingerprints(n, seed=200 + k)[:, 1]
    f = O.OracleFlat(k, mh)
    q = f.query(mh)
    assert len(np.unique(q)) == n
    full = f.nbuckets * k
    assert int((q < full).sum()) == full  # every bucket position is taken by exactly one key


def test_flat_layers_decide_first():
    """evaluateKPerfect (ShockHash2Flat.h:270-299): a key sits in its layer-0 bucket or, bumped there, in its
    layer-1 bucket (hash remix(mhc)), unless the fallback placed it; the fallback places exactly the keys both
    layers bumped, each into a free position of a bucket, and every bucket ends with k keys."""
    k, n = 64, 64 * 800
    mh = workload.fingerprints(n, seed=7)[:, 1]
    f = O.OracleFlat(k, mh)
    bucket, _ = f.evaluate(mh)
    b1, b2 = f.layer_bases[1], f.layer_bases[2]
    lo32 = np.uint64(0xFFFFFFFF)
    layer0 = (mh & lo32) * np.uint64(b1) >> np.uint64(32)
    layer1 = np.uint64(b1) + (((workload.mix13(mh) & lo32) * np.uint64(b2 - b1)) >> np.uint64(32))
    elsewhere = (bucket != layer0) & (bucket != layer1)
    assert elsewhere.sum() <= f.bumped and elsewhere.sum() >= 0.95 * f.bumped
    counts = np.bincount(bucket[bucket < f.nbuckets].astype(np.int64), minlength=f.nbuckets)
    assert (counts == k).all()


@pytest.mark.parametrize("k,n", [(16, 16 * 500 + 7), (64, 64 * 900), (100, 100 * 400 + 33)])
def test_flat_thresholds_by_brute_force(k, n):
    """flushBucket (ShockHash2Flat.h:204-233) stated over every threshold index instead of by its two compares: a
    bucket of at most k keys keeps the top index; a larger one keeps the largest index t, no larger than the index
    its k-th smallest threshold value maps to (compact_threshold, :194-202), whose threshold value admits at most k of
    the bucket's keys. Checked on every first-layer bucket against the indices the oracle stores; a slip in the
    rule (the k-th or (k+1)-th key, < for <=, a missing or extra decrement) changes some of them."""
    mh = workload.fingerprints(n, seed=400 + k)[:, 1]
    f = O.OracleFlat(k, mh)
    trange = 1 << (int(np.floor(np.log2(k))) - 1)
    mapping = O.flat_mapping(trange).astype(np.uint64)
    thr = f.thresholds()
    b0 = f.layer_bases[1]
    bucket = ((mh & np.uint64(0xFFFFFFFF)) * np.uint64(b0)) >> np.uint64(32)
    value = mh >> np.uint64(32)
    order = np.lexsort((value, bucket))
    bs, vs = bucket[order], value[order]
    starts = np.searchsorted(bs, np.arange(b0 + 1, dtype=np.uint64))
    decrements = 0
    for b in range(b0):
        v = vs[starts[b]:starts[b + 1]]
        if len(v) <= k:
            want = trange - 1
        else:
            kept = np.searchsorted(v, mapping, side="right")  # keys admitted by each index's threshold
            top = int(np.searchsorted(mapping, v[k - 1], side="left"))
            want = max(t for t in range(top + 1) if kept[t] <= k)
            decrements += want < top
        assert int(thr[b]) == want, (b, len(v))
    assert decrements > 0  # the decrementing branch was exercised


def _flush_free(bucket, value, nb, k, mapping, thr):
    """Per bucket of one layer (keys given by bucket and threshold value): the free positions flushBucket lists
    (k minus the bucket's size, or the bumped keys among its k smallest values) and which keys it bumps."""
    order = np.lexsort((value, bucket))
    bs, vs = bucket[order], value[order]
    starts = np.searchsorted(bs, np.arange(nb + 1, dtype=np.uint64))
    free, bumped = [], np.zeros(len(bucket), dtype=bool)
    for b in range(nb):
        v = vs[starts[b]:starts[b + 1]]
        if len(v) <= k:
            free += [b] * (k - len(v))
        else:
            over = v > mapping[int(thr[b])]
            free += [b] * int(over[:k].sum())
            bumped[order[starts[b]:starts[b + 1]][over]] = True
    return free, bumped


@pytest.mark.parametrize("k,n", [(16, 16 * 500 + 7), (64, 64 * 900 + 13), (100, 100 * 400)])
def test_flat_free_positions_by_brute_force(k, n):
    """freePositions (ShockHash2Flat.h:124-147) rebuilt from the keys: every layer bucket's free places in bucket
    order (layer 0, then layer 1 over the keys layer 0 bumped, rehashed by remix), then one place per missing key
    of the buckets no layer handled (k each), then the partial bucket's. The fallback's value of a bumped key
    indexes this list, so together with the pinned thresholds and ShockHash2 leaf search it fixes every value."""
    mh = workload.fingerprints(n, seed=500 + k)[:, 1]
    f = O.OracleFlat(k, mh)
    trange = 1 << (int(np.floor(np.log2(k))) - 1)
    mapping = O.flat_mapping(trange).astype(np.uint64)
    thr = f.thresholds()
    lo32 = np.uint64(0xFFFFFFFF)
    b0 = f.layer_bases[1]
    bucket = ((mh & lo32) * np.uint64(b0)) >> np.uint64(32)
    want, bumped = _flush_free(bucket, mh >> np.uint64(32), b0, k, mapping, thr[:b0])
    keys = mh[bumped]
    if f.layers == 2:
        b1 = f.layer_bases[2] - b0
        m1 = workload.mix13(keys)
        free1, bumped1 = _flush_free(((m1 & lo32) * np.uint64(b1)) >> np.uint64(32), m1 >> np.uint64(32), b1, k,
                                     mapping, thr[b0:])
        want += [b0 + b for b in free1]
        keys = keys[bumped1]
    assert len(keys) == f.bumped
    handled, in_end = f.layer_bases[f.layers], n - f.nbuckets * k
    additional = f.bumped - len(want)
    want += [handled + i // k for i in range(additional - in_end)] + [f.nbuckets + i for i in range(additional - in_end, additional)]
    got = f.free_positions()
    assert len(got) == len(want) and np.array_equal(got, np.array(want, dtype=np.uint64))
