"""ctypes binding of libshockhash_host.so: the drop-in C++ class templates (ShockHash,
SIMDShockHash, ShockHash2) instantiated for a set of leaf sizes. Construction runs the seed searches on
the GPU engine; there is no CPU fallback."""
import ctypes
import os

import numpy as np

from . import HOST_LIB

P = ctypes.c_void_p
_SIGS = {
    "shh_build": (P, [ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_char_p,
                      ctypes.c_size_t]),
    "shh_build_strings": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, P, ctypes.c_size_t, ctypes.c_size_t,
                              ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "shh_free": (None, [P]),
    "shh_query": (None, [P, P, ctypes.c_size_t, P]),
    "shh_query_strings": (None, [P, ctypes.c_char_p, P, ctypes.c_size_t, P]),
    "shh_nwords": (ctypes.c_size_t, [P]),
    "shh_words": (None, [P, P]),
    "shh_ribbon_size": (ctypes.c_size_t, [P]),
    "shh_ribbon": (None, [P, P, P]),
    "shh_nnodes": (ctypes.c_size_t, [P]),
    "shh_codes": (None, [P, P]),
    "shh_bits": (ctypes.c_size_t, [P]),
    "shh_timings": (None, [P, P]),
    "shh_phase_timings": (None, [P, P]),
    "shh_keep_retrieval_input": (None, [ctypes.c_int]),
    "shh_memo": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P]),
    "shh_last_status": (ctypes.c_int, []),
    "shh_save": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    "shh_load": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    "shh_flat_build": (P, [ctypes.c_int, P, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]),
    "shh_flat_free": (None, [P]),
    "shh_flat_query": (ctypes.c_int, [P, P, ctypes.c_size_t, P, ctypes.c_char_p, ctypes.c_size_t]),
    "shh_flat_info": (None, [P, P]),
    "shh_flat_nbuckets": (ctypes.c_size_t, [P]),
    "shh_flat_codes": (None, [P, P]),
    "shh_murmur64a": (ctypes.c_uint64, [P, ctypes.c_size_t, ctypes.c_uint64]),
    "shh_flat_phases": (None, [P, P]),
}
VARIANTS = {"sh1": 0, "simd": 1, "sh2": 2}
ROTATION = 0x100  # ROTATION_FITTING = true instantiations (ShockHash / SIMDShockHash)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing: run __graft_entry__.build()")
        lib = ctypes.CDLL(HOST_LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def memo(variant, leaf):
    out = np.zeros(3000, dtype=np.uint64)
    if load().shh_memo(VARIANTS[variant], leaf, out.ctypes.data_as(P)) != 0:
        raise ValueError(f"leaf size {leaf} not instantiated")
    return out


def _pack_strings(strings):
    data = b"".join(strings)
    offs = np.zeros(len(strings) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in strings])
    return data, offs


class ConstructionError(RuntimeError):
    """A class constructor threw; `status` is the engine's SHGPU_* code (-1: another exception)."""

    def __init__(self, status, msg):
        super().__init__(f"construction failed (status {status}): {msg}")
        self.status = status


class MPHF:
    """A ShockHash / SIMDShockHash / ShockHash2 instance built by the C++ class template."""

    def __init__(self, variant, leaf, keys, bucket_size, threads=1, keep_retrieval_input=False, rotation=False):
        lib = load()
        err = ctypes.create_string_buffer(512)
        vcode = VARIANTS[variant] | (ROTATION if rotation else 0)
        lib.shh_keep_retrieval_input(1 if keep_retrieval_input else 0)
        if isinstance(keys, (list, tuple)):
            data, offs = _pack_strings(keys)
            self._strings = True
            h = lib.shh_build_strings(vcode, leaf, data, offs.ctypes.data_as(P), len(keys), bucket_size,
                                      threads, err, 512)
            self.n = len(keys)
        else:
            arr = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, 2)
            self._strings = False
            h = lib.shh_build(vcode, leaf, arr.ctypes.data_as(P), arr.shape[0], bucket_size, threads, err, 512)
            self.n = arr.shape[0]
        if not h:
            raise ConstructionError(lib.shh_last_status(), err.value.decode(errors="replace"))
        self.h = h

    def __call__(self, keys):
        lib = load()
        if isinstance(keys, (list, tuple)):
            data, offs = _pack_strings(keys)
            out = np.zeros(len(keys), dtype=np.uint64)
            lib.shh_query_strings(self.h, data, offs.ctypes.data_as(P), len(keys), out.ctypes.data_as(P))
            return out
        arr = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, 2)
        out = np.zeros(arr.shape[0], dtype=np.uint64)
        lib.shh_query(self.h, arr.ctypes.data_as(P), arr.shape[0], out.ctypes.data_as(P))
        return out

    @property
    def descriptor_words(self):
        lib = load()
        w = np.zeros(lib.shh_nwords(self.h), dtype=np.uint64)
        lib.shh_words(self.h, w.ctypes.data_as(P))
        return w

    @property
    def ribbon_input(self):
        lib = load()
        n = lib.shh_ribbon_size(self.h)
        k = np.zeros(n, dtype=np.uint64)
        b = np.zeros(n, dtype=np.uint8)
        lib.shh_ribbon(self.h, k.ctypes.data_as(P), b.ctypes.data_as(P))
        return k, b

    @property
    def codes(self):
        lib = load()
        c = np.zeros(lib.shh_nnodes(self.h), dtype=np.uint64)
        lib.shh_codes(self.h, c.ctypes.data_as(P))
        return c

    @property
    def bits(self):
        return load().shh_bits(self.h)

    @property
    def timings_ms(self):
        t = np.zeros(5, dtype=np.float64)
        load().shh_timings(self.h, t.ctypes.data_as(P))
        p = np.zeros(4, dtype=np.float64)
        load().shh_phase_timings(self.h, p.ctypes.data_as(P))
        return {"sort": t[0], "tree": t[1], "search": t[2], "host_encode": t[3], "search_kernel": t[4],
                "rice_ef": p[0], "orientation": p[1], "retrieval": p[2], "constructor": p[3]}

    def save(self, path):
        """Writes the function to `path` (the class template's save())."""
        err = ctypes.create_string_buffer(512)
        if load().shh_save(self.h, str(path).encode(), err, 512) != 0:
            raise RuntimeError(f"save failed: {err.value.decode(errors='replace')}")

    @classmethod
    def load(cls, variant, leaf, path, strings=False, rotation=False):
        """A function read from `path` (the class template's load()); answers queries only."""
        err = ctypes.create_string_buffer(512)
        h = load().shh_load(VARIANTS[variant] | (ROTATION if rotation else 0), leaf, str(path).encode(), err, 512)
        if not h:
            raise RuntimeError(f"load failed: {err.value.decode(errors='replace')}")
        obj = cls.__new__(cls)
        obj.h, obj._strings, obj.n = h, strings, None
        return obj

    def close(self):
        if getattr(self, "h", None):
            load().shh_free(self.h)
            self.h = None

    __del__ = close


class FlatMPHF:
    """A ShockHash2Flat<k> built by the C++ class template from 64-bit key hashes (mhc)."""

    def __init__(self, k, mhcs):
        lib = load()
        err = ctypes.create_string_buffer(512)
        arr = np.ascontiguousarray(mhcs, dtype=np.uint64)
        self.n = len(arr)
        h = lib.shh_flat_build(k, arr.ctypes.data_as(P), len(arr), err, 512)
        if not h:
            raise ConstructionError(lib.shh_last_status(), err.value.decode(errors="replace"))
        self.h = h

    def __call__(self, mhcs):
        arr = np.ascontiguousarray(mhcs, dtype=np.uint64)
        out = np.zeros(len(arr), dtype=np.uint64)
        err = ctypes.create_string_buffer(512)
        if load().shh_flat_query(self.h, arr.ctypes.data_as(P), len(arr), out.ctypes.data_as(P), err, 512) != 0:
            raise RuntimeError(err.value.decode(errors="replace"))
        return out

    @property
    def info(self):
        v = np.zeros(5, dtype=np.float64)
        load().shh_flat_info(self.h, v.ctypes.data_as(P))
        return {"nbuckets": int(v[0]), "layers": int(v[1]), "seed_fallbacks": int(v[2]), "bits": int(v[3]),
                "constructor_ms": float(v[4])}

    @property
    def phases_ms(self):
        """Construction phases: threshold layers and bumping, fallback function, placement of every key, device
        base-case searches with orientation and retrieval, and of that the batched base-case search alone."""
        v = np.zeros(5, dtype=np.float64)
        load().shh_flat_phases(self.h, v.ctypes.data_as(P))
        return {"layers": float(v[0]), "fallback": float(v[1]), "placement": float(v[2]), "device": float(v[3]),
                "search": float(v[4])}

    @property
    def codes(self):
        lib = load()
        c = np.zeros(lib.shh_flat_nbuckets(self.h), dtype=np.uint64)
        lib.shh_flat_codes(self.h, c.ctypes.data_as(P))
        return c

    def close(self):
        if getattr(self, "h", None):
            load().shh_flat_free(self.h)
            self.h = None

    __del__ = close

