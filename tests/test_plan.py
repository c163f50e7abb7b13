"""Host-side tree planning of the engine (no GPU) against the oracle's recursion."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import engine as E
from shockhash_amd import workload


@pytest.mark.parametrize("variant,leaf", [("sh2", 64), ("sh1", 30), ("simd", 58), ("sh2", 16), ("sh1", 8), ("simd", 24)])
def test_forest_matches_oracle_recursion(variant, leaf):
    h = workload.fingerprints(30000, seed=3)
    ob = O.OracleBuild(variant, leaf, h, 2000, threads=8) if leaf <= 30 else None
    if ob is None:
        # large leaves make the oracle's seed search slow; compare shapes only via bucket sizes
        nb = max(1, (len(h) + 1999) // 2000)
        first = np.sort(h[:, 0])
        b = ((first.astype(object) * nb) >> 64).astype(np.uint64)
        size_acc = np.searchsorted(b, np.arange(nb + 1), side="left").astype(np.uint64)
    else:
        size_acc = ob.size_acc
    nodes = E.plan_forest(variant, leaf, size_acc)
    # every bucket's subtree: root offset = bucket start, node sizes sum consistency
    roots = [i for i in range(len(nodes))]
    assert nodes["subtree"].sum() > 0
    if ob is not None:
        assert np.array_equal(nodes["m"].astype(np.uint32), ob.node_m)
        kinds = np.where(nodes["kind"] == 3, 3, nodes["kind"])
        assert np.array_equal(kinds, ob.node_kind)


def test_forest_subtree_counts():
    """subtree sizes are consistent with DFS pre-order: children follow their parent contiguously."""
    size_acc = np.array([0, 2000, 3999, 6000, 6001, 6003], dtype=np.uint64)
    nodes = E.plan_forest("sh2", 64, size_acc)
    i = 0
    while i < len(nodes):
        assert nodes["subtree"][i] >= 1
        i += int(nodes["subtree"][i])  # jump over a whole bucket tree
    assert i == len(nodes)


def test_oversized_bucket_rejected():
    with pytest.raises(E.ShgpuError) as ei:
        E.plan_forest("sh2", 64, np.array([0, 3000], dtype=np.uint64))
    assert ei.value.status == E.ERR_BUCKET_TOO_LARGE


@pytest.mark.parametrize("variant,leaf", [("sh2", 64), ("sh1", 30), ("simd", 58), ("sh1", 8), ("sh2", 16)])
def test_node_seed_bounds_follow_the_golomb_parameter(variant, leaf):
    """Every planned node's max_trials is min(2^40, 2^(k(m) + 24)), k(m) the Golomb-Rice parameter of a node of
    m keys in the class templates' compile-time memo (the engine evaluates the same formula at run time)."""
    from shockhash_amd import host as H
    size_acc = np.array([0, 2000, 3999, 6000, 6100, 6102, 8999], dtype=np.uint64)
    nodes = E.plan_forest(variant, leaf, size_acc)
    memo = H.memo(variant, leaf)
    k = (memo[nodes["m"].astype(np.int64)] >> np.uint64(27)).astype(np.int64)
    want = np.minimum(np.uint64(1 << 40), np.left_shift(np.uint64(1), (k + 24).astype(np.uint64)))
    assert np.array_equal(nodes["max_trials"], want)
