"""ctypes bindings for the CPU oracle (oracle/_build/liboracle.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

U64 = ctypes.c_uint64
U32 = ctypes.c_uint32
P = ctypes.c_void_p

_SIGS = {
    "orc_remix": (U64, [U64]),
    "orc_fastrange32": (U32, [U32, U32]),
    "orc_fastrange64": (U64, [U64, U64]),
    "orc_remap16": (U64, [U64, U64]),
    "orc_remap128": (U64, [U64, U64]),
    "orc_remix32": (U32, [U32]),
    "orc_remix_and_remap": (U32, [U64, U32]),
    "orc_start_seed": (U64, [ctypes.c_int]),
    "orc_pair_triangular": (U64, [U64, U64]),
    "orc_unpair_triangular": (None, [U64, P, P]),
    "orc_pair_elegant": (U64, [U64, U64]),
    "orc_unpair_elegant": (None, [U64, P, P]),
    "orc_float_accurate_triangular": (ctypes.c_int, [U64, U64]),
    "orc_float_accurate_elegant": (ctypes.c_int, [U64, U64]),
    "orc_spooky128": (None, [P, ctypes.c_size_t, U64, P, P]),
    "orc_spooky128_batch": (None, [P, P, ctypes.c_size_t, U64, P, ctypes.c_int]),
    "orc_sh1_valid_batch": (None, [P, ctypes.c_size_t, U32, P, P, ctypes.c_int]),
    "orc_pseudoforest_prefix": (ctypes.c_size_t, [ctypes.c_int, P, P, ctypes.c_size_t, ctypes.c_size_t]),
    "orc_sh1_cells": (None, [U64, U64, U32, P, P]),
    "orc_sh1_leaf": (U64, [P, U32, U64, P]),
    "orc_sh1_leaf_batch": (None, [P, ctypes.c_size_t, U32, U64, P, ctypes.c_int]),
    "orc_sh2_find_seed": (U64, [P, U32]),
    "orc_sh2_leaf_batch": (None, [P, P, P, ctypes.c_size_t, P, ctypes.c_int]),
    "orc_sh2_retrieval": (None, [P, U32, U64, P]),
    "orc_sh2_query_cell": (U64, [U32, U64, U64, U64]),
    "orc_sh2_hash": (U64, [U32, U64, U64]),
    "orc_split": (U64, [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, U32, P, P, P, P]),
    "orc_split_ok": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, U32, U64]),
    "orc_golomb_memo": (None, [ctypes.c_int, ctypes.c_int, P]),
    "orc_build": (P, [ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]),
    "orc_build_rot": (P, [ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]),
    "orc_flat_build": (P, [U32, P, ctypes.c_size_t, ctypes.c_int]),
    "orc_flat_free": (None, [P]),
    "orc_flat_query": (None, [P, P, ctypes.c_size_t, P]),
    "orc_flat_info": (None, [P, P]),
    "orc_flat_evaluate": (None, [P, P, ctypes.c_size_t, P, P]),
    "orc_flat_mapping": (None, [U32, P]),
    "orc_flat_thresholds": (None, [P, P]),
    "orc_flat_free_positions": (None, [P, P]),
    "orc_sh1_leaf_rot": (None, [P, U32, U64, ctypes.POINTER(U64), ctypes.POINTER(U32), P]),
    "orc_sh1_rot_query_cell": (U64, [U32, U64, U64, U64, U64]),
    "orc_build_buckets": (P, [ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, ctypes.c_size_t, P, ctypes.c_size_t,
                              ctypes.c_int]),
    "orc_build_free": (None, [P]),
    "orc_build_nbuckets": (ctypes.c_size_t, [P]),
    "orc_build_nwords": (ctypes.c_size_t, [P]),
    "orc_build_bits": (U64, [P]),
    "orc_build_words": (None, [P, P]),
    "orc_build_acc": (None, [P, P, P]),
    "orc_build_ribbon": (None, [P, P, P]),
    "orc_build_nnodes": (ctypes.c_size_t, [P]),
    "orc_build_nodes": (None, [P, P, P, P, P, P]),
    "orc_build_query": (None, [P, P, ctypes.c_size_t, P]),
}

VARIANT = {"sh1": 0, "simd": 1, "sh2": 2}

_lib = None
_path = LIB_PATH


def use_library(path):
    """Loads the oracle from `path` (another build of oracle/shockhash_oracle.cpp) instead of the default build;
    only before the first call."""
    global _path
    if _lib is not None and path != _path:
        raise RuntimeError("the oracle library is already loaded")
    _path = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_path):
            raise RuntimeError(f"oracle not built: {_path} (run __graft_entry__.build())")
        L = ctypes.CDLL(_path)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def sh1_leaf(keys, start=0, with_cells=False):
    keys = u64(keys)
    cells = np.zeros(len(keys), dtype=np.int32)
    x = lib().orc_sh1_leaf(ptr(keys), len(keys), start, ptr(cells))
    return (x, cells) if with_cells else x


def sh1_leaf_rot(keys, start=0):
    """Rotation-fitted full leaf (ShockHash.h:316-403): (absolute x, rotation r, leaf index of the key in each
    cell)."""
    keys = u64(keys)
    cells = np.zeros(len(keys), dtype=np.int32)
    x, r = ctypes.c_uint64(), ctypes.c_uint32()
    lib().orc_sh1_leaf_rot(ptr(keys), len(keys), start, ctypes.byref(x), ctypes.byref(r), ptr(cells))
    return x.value, r.value, cells


def sh1_leaf_batch(keys2d, start=0, threads=8):
    keys2d = u64(keys2d)
    n, m = keys2d.shape
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_sh1_leaf_batch(ptr(keys2d), n, m, start, ptr(out), threads)
    return out


def sh2_find_seed(keys):
    keys = u64(keys)
    return lib().orc_sh2_find_seed(ptr(keys), len(keys))


def sh2_leaf_batch(keys, offsets, sizes, threads=8):
    keys = u64(keys)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    out = np.zeros(len(sizes), dtype=np.uint64)
    lib().orc_sh2_leaf_batch(ptr(keys), ptr(offsets), ptr(sizes), len(sizes), ptr(out), threads)
    return out


def sh2_retrieval(keys, code):
    keys = u64(keys)
    bits = np.zeros(len(keys), dtype=np.uint8)
    lib().orc_sh2_retrieval(ptr(keys), len(keys), code, ptr(bits))
    return bits


def split(variant, leaf, level, keys):
    """Returns (code, permuted keys, kind, fanout, unit, start)."""
    keys = u64(keys).copy()
    kind = ctypes.c_int()
    fanout = ctypes.c_uint32()
    unit = ctypes.c_uint32()
    start = ctypes.c_uint64()
    code = lib().orc_split(variant, leaf, level, ptr(keys), len(keys), ctypes.byref(kind),
                           ctypes.byref(fanout), ctypes.byref(unit), ctypes.byref(start))
    return code, keys, kind.value, fanout.value, unit.value, start.value


def golomb_memo(variant, leaf):
    out = np.zeros(3000, dtype=np.uint64)
    lib().orc_golomb_memo(VARIANT[variant] if isinstance(variant, str) else variant, leaf, ptr(out))
    return out


def spooky128(data: bytes, seed=0):
    a = ctypes.c_uint64()
    b = ctypes.c_uint64()
    buf = ctypes.create_string_buffer(data, len(data))
    lib().orc_spooky128(buf, len(data), seed, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def spooky128_batch(strings, seed=0, threads=8):
    """(n, 2) fingerprints (first, second) of byte strings."""
    data = np.frombuffer(b"".join(strings), dtype=np.uint8)
    offs = np.zeros(len(strings) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(x) for x in strings])
    out = np.zeros((len(strings), 2), dtype=np.uint64)
    if len(data) == 0:
        data = np.zeros(1, dtype=np.uint8)
    lib().orc_spooky128_batch(ptr(data), ptr(offs), len(strings), seed, ptr(out), threads)
    return out


def sh1_valid_batch(keys2d, seeds, threads=8):
    """Bool per leaf: seeds[l] satisfies the ShockHash1 leaf predicate for keys2d[l]."""
    keys2d = u64(keys2d)
    seeds = u64(seeds)
    n, m = keys2d.shape
    ok = np.zeros(n, dtype=np.uint8)
    lib().orc_sh1_valid_batch(ptr(keys2d), n, m, ptr(seeds), ptr(ok), threads)
    return ok.astype(bool)


class OracleBuild:
    """Full RecSplit construction by the oracle: descriptors, EF inputs, ribbon input, node codes."""

    def __init__(self, variant, leaf, hashes, bucket_size, threads=1, rotation=False):
        hashes = u64(hashes).reshape(-1, 2)
        self.n = hashes.shape[0]
        L = lib()
        make = L.orc_build_rot if rotation else L.orc_build
        self.h = make(VARIANT[variant], leaf, ptr(hashes), self.n, bucket_size, threads)
        if not self.h:
            raise RuntimeError("oracle build failed")
        self.nbuckets = L.orc_build_nbuckets(self.h)
        self.words = np.zeros(L.orc_build_nwords(self.h), dtype=np.uint64)
        L.orc_build_words(self.h, ptr(self.words))
        self.bits = L.orc_build_bits(self.h)
        self.size_acc = np.zeros(self.nbuckets + 1, dtype=np.uint64)
        self.pos_acc = np.zeros(self.nbuckets + 1, dtype=np.uint64)
        L.orc_build_acc(self.h, ptr(self.size_acc), ptr(self.pos_acc))
        self.ribbon_keys = np.zeros(self.n, dtype=np.uint64)
        self.ribbon_bits = np.zeros(self.n, dtype=np.uint8)
        L.orc_build_ribbon(self.h, ptr(self.ribbon_keys), ptr(self.ribbon_bits))
        nn = L.orc_build_nnodes(self.h)
        self.node_bucket = np.zeros(nn, dtype=np.uint32)
        self.node_m = np.zeros(nn, dtype=np.uint32)
        self.node_level = np.zeros(nn, dtype=np.int32)
        self.node_kind = np.zeros(nn, dtype=np.uint8)
        self.node_code = np.zeros(nn, dtype=np.uint64)
        L.orc_build_nodes(self.h, ptr(self.node_bucket), ptr(self.node_m), ptr(self.node_level),
                          ptr(self.node_kind), ptr(self.node_code))

    def query(self, hashes):
        hashes = u64(hashes).reshape(-1, 2)
        out = np.zeros(hashes.shape[0], dtype=np.uint64)
        lib().orc_build_query(self.h, ptr(hashes), hashes.shape[0], ptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_build_free(self.h)
            self.h = None


class OracleBucketBuild:
    """The oracle's trees of chosen buckets of a partition (the rest is not searched): per node, in the order of
    the chosen buckets and DFS pre-order within each, its bucket, size, level, kind and code."""

    def __init__(self, variant, leaf, hashes, bucket_size, buckets, threads=8):
        hashes = u64(hashes).reshape(-1, 2)
        sel = np.ascontiguousarray(buckets, dtype=np.uint32)
        L = lib()
        h = L.orc_build_buckets(VARIANT[variant], leaf, ptr(hashes), hashes.shape[0], bucket_size, ptr(sel), len(sel),
                                threads)
        if not h:
            raise RuntimeError("oracle bucket build failed")
        try:
            self.nbuckets = L.orc_build_nbuckets(h)
            nn = L.orc_build_nnodes(h)
            self.node_bucket = np.zeros(nn, dtype=np.uint32)
            self.node_m = np.zeros(nn, dtype=np.uint32)
            self.node_level = np.zeros(nn, dtype=np.int32)
            self.node_kind = np.zeros(nn, dtype=np.uint8)
            self.node_code = np.zeros(nn, dtype=np.uint64)
            L.orc_build_nodes(h, ptr(self.node_bucket), ptr(self.node_m), ptr(self.node_level), ptr(self.node_kind),
                              ptr(self.node_code))
        finally:
            L.orc_build_free(h)


class OracleFlat:
    """ShockHash2Flat<k> by the oracle (include/ShockHash2Flat.h:37-302) from the keys' 64-bit hashes."""

    def __init__(self, k, mhcs, threads=8):
        mhcs = u64(mhcs)
        self.k, self.n = k, len(mhcs)
        self.h = lib().orc_flat_build(k, ptr(mhcs), len(mhcs), threads)
        if not self.h:
            raise RuntimeError("oracle flat build failed")
        info = np.zeros(7, dtype=np.uint64)
        lib().orc_flat_info(self.h, ptr(info))
        self.nbuckets, self.layers = int(info[0]), int(info[1])
        self.layer_bases = [int(x) for x in info[2:5]]
        self.bumped, self.seed_fallbacks = int(info[5]), int(info[6])

    def query(self, mhcs):
        mhcs = u64(mhcs)
        out = np.zeros(len(mhcs), dtype=np.uint64)
        lib().orc_flat_query(self.h, ptr(mhcs), len(mhcs), ptr(out))
        return out

    def thresholds(self):
        """Compacted threshold index of every bucket of the layers (the first layer_bases[layers] buckets)."""
        handled = int(self.layer_bases[self.layers])
        out = np.zeros(max(handled, 1), dtype=np.uint32)
        lib().orc_flat_thresholds(self.h, ptr(out))
        return out[:handled]

    def free_positions(self):
        """freePositions: the position list the fallback's values index, one entry per bumped key."""
        out = np.zeros(max(self.bumped, 1), dtype=np.uint64)
        lib().orc_flat_free_positions(self.h, ptr(out))
        return out[: self.bumped]

    def evaluate(self, mhcs):
        mhcs = u64(mhcs)
        b = np.zeros(len(mhcs), dtype=np.uint64)
        s = np.zeros(len(mhcs), dtype=np.uint64)
        lib().orc_flat_evaluate(self.h, ptr(mhcs), len(mhcs), ptr(b), ptr(s))
        return b, s

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_flat_free(self.h)
            self.h = None


def flat_mapping(trange):
    out = np.zeros(trange, dtype=np.uint32)
    lib().orc_flat_mapping(trange, ptr(out))
    return out

