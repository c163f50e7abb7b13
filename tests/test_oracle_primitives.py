"""Pins the oracle's primitives with published vectors and closed forms (no GPU)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))
M64 = (1 << 64) - 1


def test_remix_is_splitmix64_finaliser():
    L = O.lib()
    s = 1234567
    for want in GOLD["splitmix64"]["seed_1234567"]:
        s = (s + 0x9E3779B97F4A7C15) & M64
        assert L.orc_remix(s) == int(want)
    assert L.orc_remix(0x9E3779B97F4A7C15) == int(GOLD["splitmix64"]["seed_0_first"], 16)


def murmur3_32(data: bytes, seed: int, fmix) -> int:
    """MurmurHash3_x86_32 body; its finaliser is the oracle function under test."""
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h = seed & 0xFFFFFFFF
    n = len(data) // 4
    rotl = lambda x, r: ((x << r) | (x >> (32 - r))) & 0xFFFFFFFF
    for i in range(n):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * c1) & 0xFFFFFFFF
        k = rotl(k, 15)
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
        h = rotl(h, 13)
        h = (h * 5 + 0xE6546B64) & 0xFFFFFFFF
    tail = data[4 * n:]
    k = 0
    if len(tail) >= 3:
        k ^= tail[2] << 16
    if len(tail) >= 2:
        k ^= tail[1] << 8
    if len(tail) >= 1:
        k ^= tail[0]
        k = (k * c1) & 0xFFFFFFFF
        k = rotl(k, 15)
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
    h ^= len(data)
    return fmix(h)


def test_remix32_is_murmur3_fmix32():
    L = O.lib()
    for c in GOLD["murmur3_x86_32"]["cases"]:
        assert murmur3_32(bytes.fromhex(c["hex"]), c["seed"], L.orc_remix32) == int(c["out"], 16), c


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_range_reductions_closed_forms(seed):
    L = O.lib()
    rng = np.random.default_rng(seed)
    xs = [int(v) for v in rng.integers(0, 2**63, size=200, dtype=np.uint64)] + [0, M64, 1 << 63, (1 << 48) - 1]
    ns = [1, 2, 3, 7, 32, 64, 1000, 2999, 3000]
    for x in xs:
        x2 = (x * 2 + 1) & M64
        for n in ns:
            assert L.orc_fastrange64(x2, n) == (x2 * n) >> 64
            assert L.orc_remap128(x2, n) == (x2 * n) >> 64
            assert L.orc_remap16(x2, n) == ((x2 & ((1 << 48) - 1)) * n) >> 48 < max(n, 1)
            lo = x2 & 0xFFFFFFFF
            assert L.orc_fastrange32(lo, n) == (lo * n) >> 32
            r32 = L.orc_remix32((x2 >> 32) ^ lo)
            assert L.orc_remix_and_remap(x2, n) == ((r32 & 0xFFFF) * n) >> 16


def test_pairing_tables_and_round_trips():
    L = O.lib()
    for x, y, z in GOLD["pairing_tables"]["triangular"]:
        assert L.orc_pair_triangular(x, y) == z
    for x, y, z in GOLD["pairing_tables"]["elegant"]:
        assert L.orc_pair_elegant(x, y) == z
    import ctypes
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    for x in range(1, 300):
        for y in range(0, x):
            z = L.orc_pair_triangular(x, y)
            L.orc_unpair_triangular(z, ctypes.byref(a), ctypes.byref(b))
            assert (a.value, b.value) == (x, y)
    for x in range(0, 200):
        for y in range(0, 200):
            z = L.orc_pair_elegant(x, y)
            L.orc_unpair_elegant(z, ctypes.byref(a), ctypes.byref(b))
            assert (a.value, b.value) == (x, y)
    # triangular codes are lexicographic in (x, y): the first success in (j, i) order is the minimum
    codes = [L.orc_pair_triangular(x, y) for x in range(1, 60) for y in range(x)]
    assert codes == sorted(codes) and len(set(codes)) == len(codes)
    # large values stay exact where the reference's float checks would pass
    assert L.orc_float_accurate_triangular(1 << 24, 12345) and L.orc_float_accurate_elegant(1 << 20, 77)


def test_spooky_readme_kat_end_to_end():
    """README known answer through the whole oracle (SpookyV2 + bucketing + ShockHash1 leaf + orientation)."""
    kat = GOLD["readme_kat"]
    h = np.array([O.spooky128(k.encode(), 0) for k in kat["keys"]], dtype=np.uint64)
    b = O.OracleBuild("sh1", kat["leaf_size"], h, kat["bucket_size"])
    assert list(b.query(h)) == kat["values"]


def test_power_of_two_split_parts_from_top_bits():
    """The split kernel's shortcut for m = F * unit, F in {2, 4}: the part of a key is bits 46..47 of
    remix(z) (sux flavour), and those equal bits 14..15 of mulhi(lo, C_lo) + lo * C_hi + hi * C_lo
    (mod 2^16) where (hi, lo) is Mix13's state before the last multiply. Checked against the oracle's
    remix and against floor(h / unit) with h = remap16(remix(z), m)."""
    rng = np.random.default_rng(3)
    z = rng.integers(0, 2**64 - 1, size=20000, dtype=np.uint64)
    M = 0xFFFFFFFFFFFFFFFF
    for zz in z[:4000]:
        x = int(zz)
        x = ((x ^ (x >> 30)) * 0xbf58476d1ce4e5b9) & M
        x ^= x >> 27
        lo, hi = x & 0xFFFFFFFF, x >> 32
        top = ((lo * 0x133111eb) >> 32) + (lo & 0xFFFFFF) * (0x94d049bb & 0xFFFFFF) + \
            (hi & 0xFFFFFF) * (0x133111eb & 0xFFFFFF)
        r = O.lib().orc_remix(int(zz))
        assert (top >> 14) & 3 == (r >> 46) & 3
        for m, unit in ((256, 64), (120, 30), (1536, 768), (128, 64)):
            h = ((r & (2**48 - 1)) * m) >> 48
            F = m // unit
            assert h // unit == ((r >> 46) & 3) >> (2 - F.bit_length() + 1)


@pytest.mark.parametrize("unit", [8, 30, 58, 64, 85])
def test_packed_four_way_counts_are_exact(unit):
    """F = 4: sum of 1 << (8 * part) equals unit * 0x01010101 (mod 2^32) exactly when every part holds
    unit keys, over every count vector (c0, c1, c2, c3) with c0 + c1 + c2 + c3 = 4 * unit."""
    m = 4 * unit
    c0, c1, c2 = np.meshgrid(np.arange(m + 1), np.arange(m + 1), np.arange(m + 1), indexing="ij")
    c3 = m - c0 - c1 - c2
    ok = c3 >= 0
    c0, c1, c2, c3 = c0[ok], c1[ok], c2[ok], c3[ok]
    packed = (c0 + (c1 << 8) + (c2 << 16) + (c3 << 24)) & 0xFFFFFFFF
    hit = packed == unit * 0x01010101
    assert np.array_equal(hit, (c0 == unit) & (c1 == unit) & (c2 == unit))


def test_spookyv2_self_test_table():
    """SpookyV2's own self-test values (Hash32 of bytes (i + 128) & 255, lengths 0..63) pin the oracle's short
    path, which fingerprints every string key of the configurations (10..50 bytes)."""
    want = [int(x, 16) for x in GOLD["spookyv2_hash32"]["expected"]]
    buf = bytes(((j + 128) & 255) for j in range(len(want)))
    h = O.spooky128_batch([buf[:i] for i in range(len(want))])
    assert [int(x) & 0xFFFFFFFF for x in h[:, 1]] == want


def test_spookyv2_long_path_blocks():
    """Long messages (>= 192 bytes) are not pinned by a vector; check what the algorithm fixes: the hash of a
    message depends on every byte, including the last block's zero padding being distinct from data zeros."""
    base = bytes(range(256)) * 2
    hs = O.spooky128_batch([base[:L] for L in range(190, 400)])
    assert len({tuple(r) for r in hs}) == len(hs)  # every length distinct
    flipped = [bytearray(base[:300]) for _ in range(300)]
    for i, b in enumerate(flipped):
        b[i] ^= 1
    hf = O.spooky128_batch([bytes(b) for b in flipped])
    assert len({tuple(r) for r in hf}) == 300 and tuple(O.spooky128_batch([base[:300]])[0]) not in {tuple(r) for r in hf}
