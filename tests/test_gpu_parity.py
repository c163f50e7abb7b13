"""GPU engine vs the CPU oracle, bit-exact, through the C ABI (libshockhash_amd.so)."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload
from shockhash_amd.engine import KEYS_DEVICE, KEYS_HOST, NODES_HOST, SPLIT_DTYPE, DeviceArray, plan_forest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_sh1_tiny_leaf_first(engine):
    """One small leaf first: a minimal end-to-end call of the kernel."""
    log("sh1 tiny leaf")
    keys = workload.isolated_leaves(1, 8, seed=10)
    want = O.sh1_leaf_batch(keys, start=0, threads=1)
    got = engine.sh1_leaf_search(keys.reshape(-1), [0], [8], np.zeros(1, np.uint64), max_trials=1 << 20)
    log(f"  got {got} want {want}")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("S", [1, 2, 4])
def test_sh1_isolated_leaves_n40(engine, S):
    """Isolated base case (bijections/ShockHash1.h semantics: seed from 0), n = 40."""
    log(f"sh1 isolated n=40 S={S}")
    engine.set_tuning(sh1_seeds_per_lane=S)
    keys = workload.isolated_leaves(48, 40, seed=11)
    want = O.sh1_leaf_batch(keys, start=0, threads=8)
    log("  oracle done")
    got = engine.sh1_leaf_search(keys.reshape(-1), np.arange(48) * 40, np.full(48, 40), np.zeros(48, np.uint64),
                                 max_trials=1 << 26)
    log("  gpu done")
    engine.set_tuning(sh1_seeds_per_lane=4)
    assert np.array_equal(got, want)


def test_sh1_leaves_all_sizes(engine):
    """Leaves of every size 1..60 with the per-level start seeds used inside the trees."""
    log("sh1 all sizes")
    rng = np.random.default_rng(5)
    sizes = np.array([m for m in range(1, 61) for _ in range(2)], dtype=np.uint32)
    sizes = sizes[sizes <= 34]  # larger leaves are covered by the n=40 batch and the builds
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint32)
    keys = workload.fingerprints(int(sizes.sum()), seed=12)[:, 1]
    starts = np.array([O.lib().orc_start_seed(int(rng.integers(0, 20))) for _ in sizes], dtype=np.uint64)
    got = engine.sh1_leaf_search(keys, offsets, sizes, starts, max_trials=1 << 26)
    want = np.array([O.sh1_leaf(keys[o:o + m], int(s)) if m >= 2 else int(s)
                     for o, m, s in zip(offsets, sizes, starts)], dtype=np.uint64)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("flags", [0, 1, 8])
def test_sh2_leaves_all_sizes(engine, flags):
    """ShockHash2 leaves of every size 2..128 (Basic finder below 10, QuadSplit from 10; up to MAX_LEAF_SIZE2, two
    keys per lane above 64), on the register path (flags 0), on the scratch-memory path alone (flags 1,
    SHGPU_TUNE_SH2_GENERAL_ONLY), and with tables too small for most leaves above 64 keys, which the second pass
    then finishes (flags 8, SHGPU_TUNE_SH2_TIGHT_TABLES)."""
    log(f"sh2 all sizes, flags {flags}")
    sizes = np.array(list(range(2, 129)) * 2 + [64] * 32 + [128] * 16, dtype=np.uint32)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint32)
    keys = workload.fingerprints(int(sizes.sum()), seed=13)[:, 1]
    want = O.sh2_leaf_batch(keys, offsets, sizes, threads=8)
    engine.set_tuning(flags=flags)
    try:
        got = engine.sh2_leaf_search(keys, offsets, sizes)
    finally:
        engine.set_tuning(flags=0)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"mismatch at sizes {sizes[bad][:10]} got {got[bad][:5]} want {want[bad][:5]}"


def test_sh2_leaves_m64_many(engine):
    """2000 leaves of 64 keys: both paths agree with each other on every leaf and with the oracle on a
    sample (the register path hands the rare leaves that outgrow its tables to the scratch path)."""
    log("sh2 2000 leaves m=64")
    n = 2000
    sizes = np.full(n, 64, dtype=np.uint32)
    offsets = (np.arange(n) * 64).astype(np.uint32)
    keys = workload.fingerprints(n * 64, seed=14)[:, 1]
    fast = engine.sh2_leaf_search(keys, offsets, sizes)
    engine.set_tuning(flags=1)
    try:
        general = engine.sh2_leaf_search(keys, offsets, sizes)
    finally:
        engine.set_tuning(flags=0)
    assert np.array_equal(fast, general)
    pick = np.arange(0, n, 10)
    want = O.sh2_leaf_batch(keys, offsets[pick], sizes[pick], threads=8)
    assert np.array_equal(fast[pick], want)


def _remix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return z ^ (z >> np.uint64(31))


def _sh2_candidates_through(keys, code):
    """Number of ShockHash2 candidates (QuadSplitCandidateFinder order, m = 64) up to the end of the half-seed
    that holds the winning pair's larger candidate: how many candidates a search of this leaf tabulates."""
    cj = int(code)
    x = int(np.floor((1 + np.sqrt(1 + 8.0 * cj)) / 2))  # unpairTriangular: cj = x(x-1)/2 + y, y < x
    while x * (x - 1) // 2 > cj:
        x -= 1
    while (x + 1) * x // 2 <= cj:
        x += 1
    c_j = x
    s_star = int(np.floor(np.sqrt(c_j)))
    S = s_star + 1
    ev, od = keys[(keys & np.uint64(1)) == 0], keys[(keys & np.uint64(1)) == 1]
    s = np.arange(S, dtype=np.uint64)
    with np.errstate(over="ignore"):
        ma = np.bitwise_or.reduce(np.uint64(1) << (_remix(ev[None, :] + s[:, None]) >> np.uint64(59)), axis=1)
        mb = np.bitwise_or.reduce(np.uint64(1) << (_remix(od[None, :] + s[:, None]) >> np.uint64(59)), axis=1)
    ok = (ma[:, None] | mb[None, :]) == np.uint64(0xFFFFFFFF)  # ok[a, b]: (a, b) covers both halves
    return int(np.triu(ok).sum() + np.tril(ok, -1).sum())


def test_sh2_leaves_past_register_tables(engine):
    """m = 64 leaves that need more than the register path's 128 candidates: the register path reports them
    with the first untested j and the scratch path continues from there. Picked on the CPU from 30000 random
    leaves (about 0.14% of leaves), searched together with ordinary leaves, compared with the oracle."""
    log("sh2 leaves past the register tables")
    n = 30000
    keys = workload.fingerprints(n * 64, seed=31)[:, 1].reshape(n, 64).copy()
    sizes = np.full(n, 64, dtype=np.uint32)
    offsets = (np.arange(n) * 64).astype(np.uint32)
    codes = O.sh2_leaf_batch(keys.reshape(-1), offsets, sizes, threads=8)
    top = np.argsort(codes)[-300:]
    big = [int(l) for l in top if _sh2_candidates_through(keys[l], codes[l]) > 128]
    log(f"  {len(big)} leaves need more than 128 candidates")
    assert len(big) >= 10
    pick = np.array(sorted(set(big) | set(range(0, n, 97))), dtype=np.int64)
    sub = keys[pick].reshape(-1).copy()
    got = engine.sh2_leaf_search(sub, (np.arange(len(pick)) * 64).astype(np.uint32),
                                 np.full(len(pick), 64, dtype=np.uint32))
    bad = np.nonzero(got != codes[pick])[0]
    assert len(bad) == 0, f"mismatch at leaves {pick[bad][:10]} got {got[bad][:5]} want {codes[pick][bad][:5]}"


@pytest.mark.parametrize("variant,leaf", [(0, 64), (1, 58), (0, 30), (0, 16), (1, 16), (1, 24)])
def test_split_search(engine, variant, leaf):
    """Split nodes of every kind: seeds and stable partitions equal the oracle's. The sizes include
    exact power-of-two fanouts (m = 2 * unit, m = 4 * unit), which the kernel counts from the top
    bits of the hash, and near misses, which it counts with thresholds."""
    log(f"split variant={variant} leaf={leaf}")
    rng = np.random.default_rng(7 + variant)
    # (m, level) covering fanout-2, upper and lower aggregation nodes of the leaf size.
    sizes = [2000, 1536, 768, 700, 464, 512, 256, 232, 230, 200, 130, 128, 120, 116, 100, 96, 70, 64, 60, 48,
             33, 32]
    tasks, chunks, expect_codes, expect_keys = [], [], [], []
    off = 0
    for m in sizes:
        level = int(rng.integers(0, 5))
        if m <= leaf:
            continue  # a leaf, not a split node
        keys = workload.fingerprints(m, seed=100 + m)[:, 1]
        st_variant = 2 if (variant == 0 and leaf == 64) else variant
        # oracle split: variant 0/2 -> sux flavour (ShockHash, ShockHash2), 1 -> SIMD flavour
        code, perm, kind, fanout, unit, start = O.split(st_variant, leaf, level, keys)
        if fanout > 4 or m <= leaf:
            continue
        t = np.zeros(1, dtype=SPLIT_DTYPE)
        t["offset"], t["m"], t["start_seed"], t["max_trials"] = off, m, start, 1 << 40
        t["unit"], t["kind"], t["hash"] = unit, kind, 1 if variant == 1 else 0
        tasks.append(t)
        chunks.append(keys)
        expect_codes.append(code)
        expect_keys.append(perm)
        off += m
    tasks = np.concatenate(tasks)
    tasks["max_trials"] = 1 << 26
    seeds, perm = engine.split_search(np.concatenate(chunks), tasks)
    assert np.array_equal(seeds - tasks["start_seed"], np.array(expect_codes, dtype=np.uint64))
    assert np.array_equal(perm, np.concatenate(expect_keys))


@pytest.mark.parametrize("variant,leaf,n", [("sh2", 64, 30000), ("sh1", 30, 20000), ("simd", 16, 20000),
                                            ("sh2", 16, 20000), ("sh1", 8, 5000), ("sh2", 100, 30000),
                                            ("sh2", 128, 20000)])
@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_build_matches_oracle(engine, variant, leaf, n, mode):
    """Whole construction: bucket sizes, every node code (DFS order) and the final key order."""
    log(f"build {variant} L={leaf} n={n} mode={mode}")
    engine.set_tuning(max_trials=1 << 26)
    h = workload.fingerprints(n, seed=21)
    ob = O.OracleBuild(variant, leaf, h, 2000, threads=8)
    b = engine.build(h, variant, leaf, 2000, mode=mode)
    assert np.array_equal(b.size_acc, ob.size_acc)
    assert np.array_equal(b.nodes["m"], ob.node_m)
    assert np.array_equal(b.codes, ob.node_code)
    log(f"  ok, timings {b.timings_ms}")


@pytest.mark.parametrize("waves", [24, 12, 4])
def test_build_fixed_grid(engine, waves):
    """The ShockHash2 schedule sizes its grid by batch (waves_per_cu 0: 4 waves per CU at this size); the
    other grids, 24 waves per CU as at 10M keys included, give the same codes."""
    log(f"sh2 build with {waves} waves per CU")
    n = 30000
    h = workload.fingerprints(n, seed=24)
    ob = O.OracleBuild("sh2", 64, h, 2000, threads=8)
    engine.set_variant_tuning("sh2", waves_per_cu=waves)
    try:
        b = engine.build(h, "sh2", 64, 2000)
    finally:
        engine.set_variant_tuning("sh2", waves_per_cu=0)
    assert np.array_equal(b.codes, ob.node_code)


@pytest.mark.parametrize("variant,leaf,n", [("sh2", 64, 60000), ("sh1", 30, 30000), ("simd", 16, 20000)])
def test_build_nodes_from_device(engine, variant, leaf, n):
    """SHGPU_NODES_HOST: the nodes come back from the device plan and equal the host plan of the same
    bucket sizes (offsets relative to the build's keys)."""
    log(f"nodes from the device plan, {variant} L={leaf} n={n}")
    h = workload.fingerprints(n, seed=25)
    d = DeviceArray.from_host(engine, h.reshape(-1))
    dev = engine.build_dev(d, n, variant, leaf, 2000, copy_keys=NODES_HOST)
    host = engine.build_dev(d, n, variant, leaf, 2000, copy_keys=0)
    d.free()
    assert np.array_equal(dev.codes, host.codes)
    assert np.array_equal(dev.nodes, host.nodes)
    plan = plan_forest(variant, leaf, dev.size_acc)
    for field in ("offset", "m", "kind", "n_children", "unit", "subtree", "start_seed"):  # max_trials: the schedule's
        assert np.array_equal(dev.nodes[field], plan[field]), field


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_build_device_keys(engine, mode):
    """SHGPU_KEYS_DEVICE: the build keeps its leaf-ordered keys in device memory of its own (what the classes
    hand to the orientation pass); they equal the host copy, and a later build does not overwrite them."""
    log(f"build device keys, mode {mode}")
    n = 30000
    h = workload.fingerprints(n, seed=22)
    b = engine.build(h, "sh2", 64, 2000, mode=mode, copy_keys=KEYS_HOST | KEYS_DEVICE)
    host = b.keys(n)
    assert np.array_equal(np.sort(host), np.sort(h[:, 1]))
    other = engine.build(workload.fingerprints(n, seed=23), "sh2", 64, 2000, mode=mode, copy_keys=KEYS_DEVICE)
    assert np.array_equal(b.device_keys(engine, n), host)
    assert np.array_equal(np.sort(other.device_keys(engine, n)), np.sort(workload.fingerprints(n, seed=23)[:, 1]))


@pytest.mark.parametrize("flags", [0, 2])
@pytest.mark.parametrize("variant,leaf,bucket,n,reps", [("simd", 24, 1000, 45047, 100), ("sh1", 30, 1000, 40000, 50)])
def test_cooperative_windows_repeated(engine, variant, leaf, bucket, n, reps, flags):
    """Many waves join the windowed searches of the last open nodes (leaves of ~500-2000 seeds are searched
    by up to a dozen waves at once). Every one of `reps` builds must return the oracle's codes: a node may be
    finalised only with a seed inside the windows completed when the window counts meet, never with a
    seed a later window found while an earlier one was still being searched."""
    log(f"cooperative windows {variant} L={leaf} b={bucket} n={n} x{reps}, flags {flags}")
    engine.set_tuning(max_trials=1 << 26, flags=flags)
    h = workload.fingerprints(n, seed=45)
    ob = O.OracleBuild(variant, leaf, h, bucket, threads=8)
    d = DeviceArray.from_host(engine, h.reshape(-1))
    bad = []
    for rep in range(reps):
        b = engine.build_dev(d, n, variant, leaf, bucket, copy_keys=False)
        diff = np.nonzero(b.codes != ob.node_code)[0]
        if len(diff):
            i = diff[0]
            bad.append(f"rep {rep}: {len(diff)} nodes differ, node {i} kind {b.nodes['kind'][i]} m {b.nodes['m'][i]} "
                       f"got {b.codes[i]} want {ob.node_code[i]}")
    d.free()
    engine.set_tuning(flags=0)
    assert not bad, "; ".join(bad[:5])


@pytest.mark.parametrize("leaf", [100, 128])
def test_wide_leaf_build_second_pass(engine, leaf):
    """ShockHash2 leaf 100 / 128 builds with tables sized for a quarter of the expected half-seeds
    (SHGPU_TUNE_SH2_TIGHT_TABLES): most leaves outgrow them in the persistent kernel and the host's second pass with
    larger tables finishes them; the codes equal the oracle's."""
    log(f"sh2 L={leaf} build, tight tables")
    h = workload.fingerprints(20000, seed=27)
    ob = O.OracleBuild("sh2", leaf, h, 2000, threads=8)
    engine.set_variant_tuning("sh2", flags=8)
    try:
        b = engine.build(h, "sh2", leaf, 2000)
    finally:
        engine.set_variant_tuning("sh2", flags=0)
    assert np.array_equal(b.codes, ob.node_code)

