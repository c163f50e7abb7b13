"""Pins the oracle's seed searches against brute force over seeds with independent predicates:
exhaustive perfect-matching search instead of cuckoo walks, explicit candidate enumeration instead of
the incremental candidate lists, plain per-part counts for splits (no GPU)."""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload

M64 = (1 << 64) - 1
L = O.lib()


def cells_sh1(key, seed, m):
    """TinyBinaryCuckooHashTable::getCandidateCells (TinyBinaryCuckooHashTable.h:138-149) composed from
    the pinned primitives."""
    r = L.orc_remix((key + seed) & M64)
    return L.orc_fastrange32(r & 0xFFFFFFFF, m // 2), L.orc_fastrange32(r >> 32, (m + 1) // 2) + m // 2


def has_perfect_matching(edges, ncells):
    """Keys -> distinct cells, each key in one of its two cells (Kuhn's augmenting paths)."""
    owner = [-1] * ncells

    def try_key(k, seen):
        for c in edges[k]:
            if c in seen:
                continue
            seen.add(c)
            if owner[c] < 0 or try_key(owner[c], seen):
                owner[c] = k
                return True
        return False
    return all(try_key(k, set()) for k in range(len(edges)))


@pytest.mark.parametrize("m", [2, 3, 4, 5, 6, 7, 8, 9, 10, 12])
def test_sh1_leaf_minimal_seed(m):
    rng = np.random.default_rng(m)
    for trial in range(4):
        keys = [int(k) for k in workload.fingerprints(m, seed=1000 + 10 * m + trial)[:, 1]]
        start = int(rng.integers(0, 2**63)) * 2 + 1
        x = start
        while True:
            cc = [cells_sh1(k, x, m) for k in keys]
            covered = set(c for pair in cc for c in pair)
            if len(covered) == m and has_perfect_matching(cc, m):
                break
            x = (x + 1) & M64
        got, by_cell = O.sh1_leaf(np.array(keys, dtype=np.uint64), start, with_cells=True)
        assert got == x
        # orientation: a permutation of the keys, each key in one of its own cells
        assert sorted(by_cell) == list(range(m))
        for cell, k in enumerate(by_cell):
            assert cell in cells_sh1(keys[k], x, m)


def half_hash(key, seed, H):
    return L.orc_fastrange64(L.orc_remix((key + seed) & M64), H)


def sh2_bruteforce(keys):
    """BijectionsShockHash2::findSeed semantics (ShockHash2-internal.h:501-546) from first principles."""
    m = len(keys)
    H = (m + 1) // 2
    full = set(range(H))
    S = 1
    while True:  # grow the half-seed bound until a success is found below it
        S *= 2
        if m < 10:
            cands = []
            for s in range(S * S):
                h = [half_hash(k, s, H) for k in keys]
                if set(h) == full:
                    cands.append((s, h))
        else:
            cands = []
            for a in range(S):
                for b in range(S):
                    h = [half_hash(k, a if k % 2 == 0 else b, H) for k in keys]
                    if set(h) == full:
                        cands.append((L.orc_pair_elegant(a, b), h))
            cands.sort()
        iso = []
        for code, h in cands:
            iso.append({i for i in range(m) if h.count(h[i]) == 1})
        for j in range(1, len(cands)):
            for i in range(j):
                if iso[j] & iso[i]:
                    continue
                edges = [(cands[j][1][k] + m // 2, cands[i][1][k]) for k in range(m)]
                if has_perfect_matching(edges, m):
                    return L.orc_pair_triangular(cands[j][0], cands[i][0])


@pytest.mark.parametrize("m", [2, 3, 4, 5, 7, 8, 9, 10, 11, 12, 13])
def test_sh2_leaf_bruteforce(m):
    for trial in range(2):
        keys = [int(k) for k in workload.fingerprints(m, seed=2000 + 10 * m + trial)[:, 1]]
        want = sh2_bruteforce(keys)
        assert O.sh2_find_seed(np.array(keys, dtype=np.uint64)) == want


def test_sh2_retrieval_is_a_placement():
    """constructRetrieval's bits place every key in a distinct cell (ShockHash2-internal.h:586-605)."""
    for m in [2, 5, 9, 10, 17, 32, 33, 64]:
        keys = workload.fingerprints(m, seed=3000 + m)[:, 1]
        code = O.sh2_find_seed(keys)
        bits = O.sh2_retrieval(keys, code)
        cells = [L.orc_sh2_query_cell(m, code, int(k), int(b)) for k, b in zip(keys, bits)]
        assert sorted(cells) == list(range(m))


@pytest.mark.parametrize("variant,leaf,m", [(0, 4, 20), (0, 4, 13), (0, 4, 7), (1, 4, 20), (1, 4, 13), (1, 4, 7),
                                            (2, 4, 50), (2, 4, 30), (2, 4, 14), (0, 30, 100)])
def test_split_minimal_seed_and_stable_partition(variant, leaf, m):
    keys = [int(k) for k in workload.fingerprints(m, seed=4000 + m + variant)[:, 1]]
    level = 2
    code, perm, kind, fanout, unit, start = O.split(variant, leaf, level, np.array(keys, dtype=np.uint64))

    def h(key, x):
        if variant == 1:
            return L.orc_remix_and_remap((key + x) & M64, m)
        return L.orc_remap16(L.orc_remix((key + x) & M64), m)

    def part(v):
        return int(v >= unit) if kind == 0 else v // unit
    x = start
    while True:
        cnt = [0] * fanout
        for k in keys:
            cnt[part(h(k, x))] += 1
        if all(c == unit for c in cnt[: (1 if kind == 0 else fanout - 1)]):
            break
        x = (x + 1) & M64
    assert code == (x - start) & M64
    want = [k for p in range(fanout) for k in keys if part(h(k, x)) == p]
    assert list(perm) == want


@pytest.mark.parametrize("split,fanout", [(232, 3), (240, 3), (58, 4), (60, 4), (30, 4), (120, 3), (24, 2), (48, 2)])
def test_simd_packed_counters_equal_plain_counts(split, fanout):
    """SIMDShockHash's aggrLevel compares 8-bit packed counters mod 2^32 (SIMDShockHash.hpp:284-305);
    for every (SPLIT, fanout) aggrLevel is instantiated with (SPLIT = leaf or lower_aggr < 255), packed equality holds exactly for
    count vectors with every non-last part equal to SPLIT. Checked over all count vectors."""
    for m in {split * (fanout - 1) + 1, split * fanout - 1, split * fanout}:
        # all compositions (c_0..c_{f-1}) of m
        if fanout == 2:
            cs = np.array([[c, m - c] for c in range(m + 1)], dtype=np.int64)
        elif fanout == 3:
            cs = np.array([[a, b, m - a - b] for a in range(m + 1) for b in range(m + 1 - a)], dtype=np.int64)
        else:
            a, b, c = np.meshgrid(np.arange(m + 1), np.arange(m + 1), np.arange(m + 1), indexing="ij")
            ok = a + b + c <= m
            cs = np.stack([a[ok], b[ok], c[ok], m - (a + b + c)[ok]], axis=1).astype(np.int64)
        packed = np.zeros(len(cs), dtype=np.uint64)
        for p in range(min(fanout, 4)):
            packed = (packed + (cs[:, p].astype(np.uint64) << np.uint64(8 * p))) & np.uint64(0xFFFFFFFF)
        found = split
        for i in range(1, fanout - 1):
            found |= split << (8 * i)
        found |= ((m - (fanout - 1) * split) << (8 * (fanout - 1))) & 0xFFFFFFFF
        accept_packed = packed == np.uint64(found & 0xFFFFFFFF)
        accept_plain = np.all(cs[:, : fanout - 1] == split, axis=1)
        assert np.array_equal(accept_packed, accept_plain)


def test_spooky_batch_equals_single():
    strings = [b"abc", b"", b"x" * 31, b"y" * 32, b"z" * 191] + workload.random_strings(200, seed=9)
    got = O.spooky128_batch(strings, threads=3)
    want = np.array([O.spooky128(x, 0) for x in strings], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_sh1_valid_batch():
    """The predicate check accepts the oracle's minimal seeds and rejects every smaller seed (brute force)."""
    keys = workload.isolated_leaves(40, 12, seed=31)
    seeds = O.sh1_leaf_batch(keys, start=0, threads=4)
    assert O.sh1_valid_batch(keys, seeds).all()
    for s in range(0, 64):
        trial = np.where(seeds > s, np.uint64(s), seeds + np.uint64(1 << 40))
        ok = O.sh1_valid_batch(keys, trial)
        assert not ok[seeds > s].any()


def rot_bits(v, r, m):
    """Rotation of an m-bit mask by r (ShockHash.h:297-299)."""
    allm = (1 << m) - 1
    return ((v << r) | (v >> (m - r))) & allm if r else v & allm


@pytest.mark.parametrize("m", [3, 4, 6, 8, 10, 12, 33])
def test_sh1_rotation_fitting_minimal(m):
    """ROTATION_FITTING leaves (ShockHash.h:316-403): the first (x, r) in order whose A cells (seed x & ~mask, mask 7
    above 32 keys) and B cells (seed x, rotated by r) admit a perfect matching of keys to cells; the oracle's
    placement puts every key in one of its own (rotated) cells."""
    mask = 7 if m > 32 else 0
    for trial in range(3 if m < 20 else 1):
        keys = [int(k) for k in workload.fingerprints(m, seed=5000 + 10 * m + trial)[:, 1]]
        start = [0, 12345, 2**40 + 7][trial]
        x = start
        found = None
        while found is None:
            ca = {i: cells_sh1(k, x & ~mask, m) for i, k in enumerate(keys) if k % 2 == 0}
            cb = {i: cells_sh1(k, x, m) for i, k in enumerate(keys) if k % 2 == 1}
            a = b = 0
            for p in ca.values():
                a |= (1 << p[0]) | (1 << p[1])
            for p in cb.values():
                b |= (1 << p[0]) | (1 << p[1])
            for r in range(m):
                if a | rot_bits(b, r, m) != (1 << m) - 1:
                    continue
                edges = [ca[i] if i in ca else tuple((c + r) % m for c in cb[i]) for i in range(m)]
                if has_perfect_matching(edges, m):
                    found = (x, r, edges)
                    break
            x += 1
        gx, gr, by_cell = O.sh1_leaf_rot(np.array(keys, dtype=np.uint64), start)
        assert (gx, gr) == found[:2]
        assert sorted(by_cell) == list(range(m))
        for cell, i in enumerate(by_cell):
            assert cell in found[2][i]
        # the query's cell of every key equals its placement, given the bit the construction stores
        code = (gx - start) * m + gr
        for cell, i in enumerate(by_cell):
            k = keys[i]
            own = (cell - gr) % m if k & 1 else cell
            bit = 0 if own < m // 2 else 1
            assert L.orc_sh1_rot_query_cell(m, code, start, k, bit) == cell


@pytest.mark.parametrize("variant,leaf", [("sh1", 8), ("sh1", 16), ("simd", 12), ("sh1", 34)])
def test_rotation_build_is_a_bijection(variant, leaf):
    """A whole ROTATION_FITTING build (ShockHash.h:316-403; SIMDShockHash.hpp:419-506): full leaves carry codes
    x * leaf + r, shorter leaves the plain search, and the oracle's queries map the keys onto 0..n-1."""
    n = 3000 if leaf > 20 else 6000
    h = workload.fingerprints(n, seed=77 + leaf)
    b = O.OracleBuild(variant, leaf, h, 100 if leaf > 20 else 60, rotation=True)
    q = b.query(h)
    assert len(set(q.tolist())) == n and int(q.max()) == n - 1
    full = (b.node_kind == 3) & (b.node_m == leaf)
    assert full.any()
    plain = O.OracleBuild(variant, leaf, h, 100 if leaf > 20 else 60)
    # same tree shapes; rotation codes differ from the plain leaves' seeds only on full leaves
    assert np.array_equal(b.node_m, plain.node_m) and np.array_equal(b.node_kind, plain.node_kind)
    assert np.array_equal(b.node_code[~full], plain.node_code[~full])
