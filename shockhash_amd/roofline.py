"""Integer-VALU roofline of the seed search: hash evaluations derived from the outputs, priced in VALU ops.

The search is bound by VALU issue: there is no MFMA work and the HBM traffic of a build is a few GB against
tens of milliseconds of hashing. The unit of work is one hash evaluation (one key hashed with one seed:
remix of key + seed, range reduction, mask or count update). Evaluations are counted from the outputs alone,
as the reference's sequential loops make them (every key for every seed up to the stored one, no early
exit, include/ShockHash.h:414-418,457-459):

* split node, ShockHash1 leaf: m * (code + 1);
* ROTATION_FITTING leaf (m = leaf size, code = x * m + r): per x the odd keys (group B, ~m/2) are hashed, the even
  ones (group A) only when x & mask is 0 (mask 7 above 32 keys, 0 below; include/ShockHash.h:340-352), so
  m/2 * (x + 1) + m/2 * ceil((x + 1) / (mask + 1)); the m rotation checks per x are bit operations, not counted;
* ShockHash2 leaf, m >= 10: m * (s* + 1) half-seed evaluations, where (c_j, c_i) = unpairTriangular(code),
  (a, b) = unpairElegant(c_j) and s* = max(a, b) is the last half-seed the candidate search needs
  (include/shockhash/ShockHash2-internal.h:435-470); m < 10: m * (c_j + 1).

Each evaluation is priced at the VALU instructions the kernel's inner loop issues for it, read off the
gfx950 ISA of libshockhash_amd.so (loop-body VALU count / evaluations per body; the keys arrive by LDS
broadcast reads, which are not VALU). The price differs by counting mode: a split whose parts are exact
quarters or halves counts the top bits of a partial hash, other shapes compare the 48-bit (sux) or 16-bit
(SIMDShockHash) hash with one threshold per part boundary. Speculative seeds past the winner, pair tests
of ShockHash2 leaves, partitions and scheduling are not counted: the fraction of peak is the share of the
chip's VALU issue capacity spent on the evaluations the reference's algorithm needs.

The price table below is the cost convention of this roofline and stays fixed: it is not re-priced when a
kernel changes, so a kernel that needs fewer instructions for the same evaluations reports a higher fraction.
It replaces a flat 32 ops per 64-bit-mix evaluation (16 for the 32-bit SIMDShockHash split hash), an estimate
that put the measured 10M-key ShockHash2 build at 1.31 of peak: 32 ops per evaluation would need 2.41e10
wave-level VALU instructions, and rocprofv3 counted 1.69e10 (SQ_INSTS_VALU) for the whole kernel, so the flat
price overcounts the work. evals_per_s is reported beside the fraction.
"""
import numpy as np

MI355X_CUS, SIMDS_PER_CU, LANES_PER_SIMD, CLOCK_GHZ = 256, 4, 16, 2.4
INT32_PEAK_TOPS = MI355X_CUS * SIMDS_PER_CU * LANES_PER_SIMD * CLOCK_GHZ / 1e3  # 39.3 T lane-ops/s

KIND_LEAF = 3
# VALU instructions per evaluation, per split counting mode and hash flavour
SPLIT_OPS = {
    "sux": {"F2": 15.5, "F4": 16.5, "NT1": 20.0, "NT2": 22.0, "NT3": 24.0},
    "simd": {"F2": 10.5, "F4": 10.5, "NT1": 11.5, "NT2": 13.0, "NT3": 15.0},
}
SH1_LEAF_OPS = 22.25  # remix (18) + two range reductions + two mask updates, plus loop overhead
SH2_HALF_SEED_OPS = 21.75  # remix (18) + fastrange64 (2) + mask update, plus loop overhead


def split_modes(nodes):
    """Counting mode of every split node, the rule of make_split_shape (search_split.h): 'F2' when a
    fanout-2 node halves exactly, 'F4' when m = 4 * unit with 3 * unit < 256, else 'NT<k>' with k part
    boundaries (k = fanout - 1). Leaves get ''."""
    m = nodes["m"].astype(np.int64)
    unit = np.maximum(nodes["unit"].astype(np.int64), 1)
    kind = nodes["kind"]
    fan = np.where(kind == 0, 2, (m + unit - 1) // unit)
    modes = np.array(["NT" + str(int(f) - 1) for f in fan], dtype=object)
    modes[(fan == 2) & (m == 2 * unit)] = "F2"
    modes[(fan == 4) & (m == 4 * unit) & (3 * unit < 256)] = "F4"
    modes[kind == KIND_LEAF] = ""
    return modes


def unpair_triangular(c):
    """(x, y) with c = x (x - 1) / 2 + y, y < x (PairingFunction.h:52-56), exact for c < 2^52."""
    c = np.asarray(c, dtype=np.float64)
    x = np.floor(0.5 + np.sqrt(0.25 + 2.0 * c))
    tri = x * (x - 1) / 2
    x = np.where(tri > c, x - 1, x)
    return x, c - x * (x - 1) / 2


def unpair_elegant(z):
    """(a, b) with z = pairElegant(a, b) (PairingFunction.h:83-91)."""
    z = np.asarray(z, dtype=np.float64)
    f = np.floor(np.sqrt(z))
    f = np.where(f * f > z, f - 1, f)
    r = z - f * f
    return np.where(r < f, r, f), np.where(r < f, f, r - f)


def evaluations(nodes, codes, variant, rotation_leaf=0):
    """Per-node hash evaluations a sequential search makes to reach the stored codes (float64); rotation_leaf: the
    leaf size of a ROTATION_FITTING build (0: none)."""
    m = nodes["m"].astype(np.float64)
    kind = nodes["kind"]
    codes = codes.astype(np.float64)
    ev = np.where(kind != KIND_LEAF, m * (codes + 1), 0.0)
    leaf = kind == KIND_LEAF
    if variant == "sh2":
        cj, _ = unpair_triangular(codes[leaf])
        a, b = unpair_elegant(cj)
        s_star = np.maximum(a, b)
        ml = m[leaf]
        ev[leaf] = np.where(ml >= 10, ml * (s_star + 1), ml * (cj + 1))
    else:
        ev[leaf] = np.where(m[leaf] >= 2, m[leaf] * (codes[leaf] + 1), 0.0)
        if rotation_leaf:
            rot = leaf & (nodes["m"] == rotation_leaf)
            x1 = np.floor(codes[rot] / rotation_leaf) + 1
            per = 8.0 if rotation_leaf > 32 else 1.0
            ev[rot] = rotation_leaf / 2 * x1 + rotation_leaf / 2 * np.ceil(x1 / per)
    return ev


def valu_ops(nodes, codes, variant, rotation_leaf=0):
    """(evaluations, VALU ops, evaluations in leaves) of a build."""
    ev = evaluations(nodes, codes, variant, rotation_leaf)
    modes = split_modes(nodes)
    flavour = "simd" if variant == "simd" else "sux"
    price = np.array([SPLIT_OPS[flavour][md] if md else 0.0 for md in modes])
    leaf = nodes["kind"] == KIND_LEAF
    price[leaf] = SH2_HALF_SEED_OPS if variant == "sh2" else SH1_LEAF_OPS
    return float(ev.sum()), float((ev * price).sum()), float(ev[leaf].sum())


def roofline(nodes, codes, variant, kernel_ms, traffic_bytes=None, rotation_leaf=0):
    """The bench's roofline record for one search kernel launch of kernel_ms milliseconds."""
    evals, ops, evals_leaf = valu_ops(nodes, codes, variant, rotation_leaf)
    achieved = ops / (kernel_ms / 1e3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": INT32_PEAK_TOPS, "unit": "TOP/s",
            "frac": achieved / INT32_PEAK_TOPS, "traffic": traffic_bytes,
            "ops_per_eval": ops / evals if evals else 0.0, "hash_evals": evals, "hash_evals_leaves": evals_leaf,
            "evals_per_s": evals / (kernel_ms / 1e3)}
