"""ctypes binding of libshockhash_amd.so (C ABI in include/shockhash_amd.h).

Device buffers are allocated through the engine's own C ABI helpers (hipMalloc on the engine's
device), so the Python side needs only ctypes and numpy. There is no CPU fallback: every entry point
raises when the library or the GPU is unusable.
"""
import ctypes
import os

import numpy as np

from . import ENGINE_LIB

VARIANTS = {"sh1": 0, "simd": 1, "sh2": 2}
MODES = {"persistent": 0, "plain": 1}
NOT_FOUND = (1 << 64) - 1

# Status codes (include/shockhash_amd.h)
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_SEARCH_LIMIT, ERR_BUCKET_TOO_LARGE, ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_ADMISSION = range(9)


class ShgpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"shgpu status {status}: {msg}")
        self.status = status


class LeafTask(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint32), ("m", ctypes.c_uint32), ("start_seed", ctypes.c_uint64),
                ("max_trials", ctypes.c_uint64)]


class SplitTask(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint32), ("m", ctypes.c_uint32), ("start_seed", ctypes.c_uint64),
                ("max_trials", ctypes.c_uint64), ("unit", ctypes.c_uint32), ("kind", ctypes.c_uint8),
                ("hash", ctypes.c_uint8), ("pad", ctypes.c_uint16)]


class Node(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint32), ("m", ctypes.c_uint16), ("kind", ctypes.c_uint8),
                ("n_children", ctypes.c_uint8), ("unit", ctypes.c_uint32), ("subtree", ctypes.c_uint32),
                ("start_seed", ctypes.c_uint64), ("max_trials", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("sh1_seeds_per_lane", ctypes.c_uint32), ("split_seeds_per_lane", ctypes.c_uint32),
                ("waves_per_cu", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("max_trials", ctypes.c_uint64)]


class ScheduleClass(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("sh1_seeds_per_lane", ctypes.c_uint32),
                ("split_seeds_per_lane", ctypes.c_uint32), ("max_keys", ctypes.c_uint64), ("tuned_ms", ctypes.c_double),
                ("sample_keys", ctypes.c_uint64)]


def _class_dict(c):
    return {f: getattr(c, f) for f, _ in ScheduleClass._fields_}


KEYS_HOST, KEYS_DEVICE, NODES_HOST = 1, 2, 4  # shgpu_build_desc.copy_keys bits


class BuildDesc(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("leaf_size", ctypes.c_uint32), ("bucket_size", ctypes.c_uint64),
                ("mode", ctypes.c_int), ("plain_bucket_limit", ctypes.c_uint32), ("copy_keys", ctypes.c_int),
                ("rotation_fitting", ctypes.c_int)]


LEAF_DTYPE = np.dtype([("offset", "<u4"), ("m", "<u4"), ("start_seed", "<u8"), ("max_trials", "<u8")])
SPLIT_DTYPE = np.dtype([("offset", "<u4"), ("m", "<u4"), ("start_seed", "<u8"), ("max_trials", "<u8"),
                        ("unit", "<u4"), ("kind", "u1"), ("hash", "u1"), ("pad", "<u2")])
NODE_DTYPE = np.dtype([("offset", "<u4"), ("m", "<u2"), ("kind", "u1"), ("n_children", "u1"), ("unit", "<u4"),
                       ("subtree", "<u4"), ("start_seed", "<u8"), ("max_trials", "<u8")])
class ServerConfig(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("leaf_size", ctypes.c_uint32), ("bucket_size", ctypes.c_uint64),
                ("slots", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64), ("copy_keys", ctypes.c_int),
                ("max_batch_keys", ctypes.c_uint64), ("max_batch_requests", ctypes.c_uint32)]


assert LEAF_DTYPE.itemsize == ctypes.sizeof(LeafTask) == 24
assert SPLIT_DTYPE.itemsize == ctypes.sizeof(SplitTask) == 32
assert NODE_DTYPE.itemsize == ctypes.sizeof(Node) == 32

P = ctypes.c_void_p
_SIGS = {
    "shgpu_last_error": (ctypes.c_char_p, []),
    "shgpu_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "shgpu_engine_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
    "shgpu_engine_destroy": (None, [P]),
    "shgpu_engine_info": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.c_char_p, ctypes.c_size_t]),
    "shgpu_get_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
    "shgpu_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
    "shgpu_sync": (ctypes.c_int, [P]),
    "shgpu_get_variant_tuning": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "shgpu_set_variant_tuning": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "shgpu_autotune": (ctypes.c_int, [P, ctypes.POINTER(BuildDesc), P, ctypes.c_uint64, ctypes.POINTER(Tuning),
                                      ctypes.POINTER(ctypes.c_double)]),
    "shgpu_autotune_class": (ctypes.c_int, [P, ctypes.POINTER(BuildDesc), P, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.POINTER(ScheduleClass)]),
    "shgpu_schedule_get": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "shgpu_schedule_set": (ctypes.c_int, [P, P, ctypes.c_uint32]),
    "shgpu_schedule_save": (ctypes.c_int, [P, ctypes.c_char_p]),
    "shgpu_schedule_load": (ctypes.c_int, [P, ctypes.c_char_p]),
    "shgpu_schedule_source": (ctypes.c_char_p, [P]),
    "shgpu_schedule_parse": (ctypes.c_int, [ctypes.c_char_p, P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "shgpu_sh1_leaf_search": (ctypes.c_int, [P, P, P, ctypes.c_uint32, P, P]),
    "shgpu_sh2_leaf_search": (ctypes.c_int, [P, P, P, ctypes.c_uint32, P, P]),
    "shgpu_split_search": (ctypes.c_int, [P, P, P, ctypes.c_uint32, P, P]),
    "shgpu_split_search_noalloc": (ctypes.c_int, [P, P, ctypes.c_uint64, P, ctypes.c_uint32, P, P]),
    "shgpu_build_run": (ctypes.c_int, [P, ctypes.POINTER(BuildDesc), P, ctypes.c_uint64, ctypes.POINTER(P)]),
    "shgpu_hash_strings": (ctypes.c_int, [P, P, P, ctypes.c_uint64, P]),
    "shgpu_leaf_orientation": (ctypes.c_int, [P, ctypes.c_int, P, P, ctypes.c_uint64, P, P]),
    "shgpu_ribbon_shards": (ctypes.c_uint64, [ctypes.c_uint64]),
    "shgpu_ribbon_max_words": (ctypes.c_uint64, [ctypes.c_uint64]),
    "shgpu_flat_layers": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.c_uint32, P, ctypes.c_uint32, P, P, P, P, P, P, P]),
    "shgpu_flat_bucket_keys": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.c_uint32, P, ctypes.c_uint32, P, ctypes.c_uint32,
                                              P, P, P, ctypes.c_uint64, P]),
    "shgpu_ribbon_build": (ctypes.c_int, [P, P, P, ctypes.c_uint64, P, P, P, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "shgpu_build_free": (None, [P]),
    "shgpu_build_nbuckets": (ctypes.c_uint64, [P]),
    "shgpu_build_nnodes": (ctypes.c_uint64, [P]),
    "shgpu_build_bucket_sizes": (ctypes.c_int, [P, P]),
    "shgpu_build_nodes": (ctypes.c_int, [P, P, P]),
    "shgpu_build_keys": (ctypes.c_int, [P, P]),
    "shgpu_build_device_keys": (ctypes.c_int, [P, ctypes.POINTER(P)]),
    "shgpu_build_timings": (ctypes.c_int, [P, P]),
    "shgpu_device_memory": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "shgpu_device_alloc": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
    "shgpu_device_free": (ctypes.c_int, [P, P]),
    "shgpu_host_alloc": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
    "shgpu_host_free": (ctypes.c_int, [P, P]),
    "shgpu_copy_to_device": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "shgpu_copy_to_host": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "shgpu_device_copy": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "shgpu_build_run_host": (ctypes.c_int, [P, ctypes.POINTER(BuildDesc), P, ctypes.c_uint64, ctypes.POINTER(P)]),
    "shgpu_plan_forest": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, P, ctypes.c_uint64, P, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "shgpu_server_create": (ctypes.c_int, [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
    "shgpu_server_destroy": (None, [P]),
    "shgpu_server_submit": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "shgpu_server_submit_device": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "shgpu_server_wait": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "shgpu_server_take": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]),
    "shgpu_server_request_bytes": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "shgpu_server_info": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
}

_lib = None


def load():
    """Loads libshockhash_amd.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ENGINE_LIB):
            raise RuntimeError(f"{ENGINE_LIB} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        lib = ctypes.CDLL(ENGINE_LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != OK:
        lib = load()
        raise ShgpuError(rc, lib.shgpu_last_error().decode(errors="replace"))


class DeviceArray:
    """A device allocation of the engine holding a numpy-typed array."""

    def __init__(self, engine, count, dtype):
        self.engine = engine
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = P()
        _check(load().shgpu_device_alloc(engine.h, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, engine, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(engine, arr.size, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes <= self.nbytes
        _check(load().shgpu_copy_to_device(self.engine.h, self.ptr, arr.ctypes.data_as(P), arr.nbytes))

    def upload_at(self, offset, arr):
        """Copies arr into the allocation starting at element `offset`."""
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert (offset + arr.size) * self.dtype.itemsize <= self.nbytes
        dst = P(self.ptr.value + offset * self.dtype.itemsize)
        _check(load().shgpu_copy_to_device(self.engine.h, dst, arr.ctypes.data_as(P), arr.nbytes))

    def download(self):
        out = np.empty(self.count, dtype=self.dtype)
        _check(load().shgpu_copy_to_host(self.engine.h, out.ctypes.data_as(P), self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None) and self.engine.h:
            load().shgpu_device_free(self.engine.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def plan_forest(variant, leaf_size, size_acc):
    """Host-only tree planning (no GPU): DFS-ordered node records for the given bucket sizes."""
    lib = load()
    size_acc = np.ascontiguousarray(size_acc, dtype=np.uint64)
    nb = len(size_acc) - 1
    cnt = ctypes.c_uint64()
    _check(lib.shgpu_plan_forest(VARIANTS[variant], leaf_size, size_acc.ctypes.data_as(P), nb, None, 0,
                                 ctypes.byref(cnt)))
    out = np.zeros(cnt.value, dtype=NODE_DTYPE)
    _check(lib.shgpu_plan_forest(VARIANTS[variant], leaf_size, size_acc.ctypes.data_as(P), nb,
                                 out.ctypes.data_as(P), cnt.value, ctypes.byref(cnt)))
    return out


class Build:
    """Result of shgpu_build_run: bucket sizes, DFS-ordered nodes and codes, leaf-ordered keys."""

    def __init__(self, handle):
        self.h = handle
        lib = load()
        self.nbuckets = lib.shgpu_build_nbuckets(handle)
        nn = lib.shgpu_build_nnodes(handle)
        self.size_acc = np.zeros(self.nbuckets + 1, dtype=np.uint64)
        _check(lib.shgpu_build_bucket_sizes(handle, self.size_acc.ctypes.data_as(P)))
        self.codes = np.zeros(nn, dtype=np.uint64)
        _check(lib.shgpu_build_nodes(handle, None, self.codes.ctypes.data_as(P)))
        self._nodes = None
        t = np.zeros(4, dtype=np.float64)
        _check(lib.shgpu_build_timings(handle, t.ctypes.data_as(P)))
        self.timings_ms = t

    @property
    def nodes(self):
        """The DFS-ordered nodes (a function of the bucket sizes, made on the host on first use)."""
        if self._nodes is None:
            nodes = np.zeros(len(self.codes), dtype=NODE_DTYPE)
            _check(load().shgpu_build_nodes(self.h, nodes.ctypes.data_as(P), None))
            self._nodes = nodes
        return self._nodes

    def keys(self, n):
        out = np.zeros(n, dtype=np.uint64)
        _check(load().shgpu_build_keys(self.h, out.ctypes.data_as(P)))
        return out

    def device_keys(self, engine, n):
        """The build's device copy of the leaf-ordered keys (copy_keys with KEYS_DEVICE), downloaded."""
        p = P()
        _check(load().shgpu_build_device_keys(self.h, ctypes.byref(p)))
        out = np.zeros(n, dtype=np.uint64)
        if n:
            _check(load().shgpu_copy_to_host(engine.h, out.ctypes.data_as(P), p, n * 8))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            load().shgpu_build_free(self.h)
            self.h = None


class Engine:
    def __init__(self, device=0):
        lib = load()
        h = P()
        _check(lib.shgpu_engine_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device
        cu, clk = ctypes.c_int(), ctypes.c_int()
        arch = ctypes.create_string_buffer(64)
        _check(lib.shgpu_engine_info(h, ctypes.byref(cu), ctypes.byref(clk), arch, 64))
        self.cu_count, self.clock_khz, self.arch = cu.value, clk.value, arch.value.decode()

    def close(self):
        if getattr(self, "h", None):
            load().shgpu_engine_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def tuning(self):
        t = Tuning()
        _check(load().shgpu_get_tuning(self.h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning
        cur.update(kw)
        t = Tuning(**cur)
        _check(load().shgpu_set_tuning(self.h, ctypes.byref(t)))

    def variant_tuning(self, variant):
        t = Tuning()
        _check(load().shgpu_get_variant_tuning(self.h, VARIANTS[variant], ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def set_variant_tuning(self, variant, **kw):
        cur = self.variant_tuning(variant)
        cur.update(kw)
        t = Tuning(**cur)
        _check(load().shgpu_set_variant_tuning(self.h, VARIANTS[variant], ctypes.byref(t)))

    def autotune(self, d_hashes, n, variant, leaf_size, bucket_size):
        """Times every candidate schedule on the sample and keeps the fastest for the variant."""
        desc = BuildDesc(VARIANTS[variant], leaf_size, bucket_size, 0, 0, 0)
        t = Tuning()
        ms = ctypes.c_double()
        _check(load().shgpu_autotune(self.h, ctypes.byref(desc), d_hashes.ptr, n, ctypes.byref(t), ctypes.byref(ms)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}, ms.value

    def autotune_class(self, d_hashes, n, variant, leaf_size, bucket_size, max_keys):
        """Tunes the variant's size class for batches up to max_keys keys on the sample (shgpu_autotune_class)."""
        desc = BuildDesc(VARIANTS[variant], leaf_size, bucket_size, 0, 0, 0)
        c = ScheduleClass()
        _check(load().shgpu_autotune_class(self.h, ctypes.byref(desc), d_hashes.ptr, n, max_keys, ctypes.byref(c)))
        return _class_dict(c)

    @property
    def schedule(self):
        """The size classes of the build schedule, sorted by variant and max_keys."""
        lib, cnt = load(), ctypes.c_uint32()
        _check(lib.shgpu_schedule_get(self.h, None, 0, ctypes.byref(cnt)))
        arr = (ScheduleClass * max(1, cnt.value))()
        _check(lib.shgpu_schedule_get(self.h, arr, len(arr), ctypes.byref(cnt)))
        return [_class_dict(arr[i]) for i in range(cnt.value)]

    def set_schedule(self, classes):
        arr = (ScheduleClass * max(1, len(classes)))()
        for i, c in enumerate(classes):
            for k, v in c.items():
                setattr(arr[i], k, v)
        _check(load().shgpu_schedule_set(self.h, arr, len(classes)))

    def save_schedule(self, path):
        _check(load().shgpu_schedule_save(self.h, os.fsencode(path)))

    def load_schedule(self, path):
        _check(load().shgpu_schedule_load(self.h, os.fsencode(path)))

    @property
    def schedule_source(self):
        return load().shgpu_schedule_source(self.h).decode()

    def sync(self):
        _check(load().shgpu_sync(self.h))

    def memory(self):
        """(free, total) bytes of the engine's device."""
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        _check(load().shgpu_device_memory(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    # ---- batched searches on device arrays (DeviceArray)
    def sh1_leaf_search_dev(self, d_keys, d_tasks, n, d_out):
        _check(load().shgpu_sh1_leaf_search(self.h, d_keys.ptr, d_tasks.ptr, n, d_out.ptr, None))

    def sh2_leaf_search_dev(self, d_keys, d_tasks, n, d_out):
        _check(load().shgpu_sh2_leaf_search(self.h, d_keys.ptr, d_tasks.ptr, n, d_out.ptr, None))

    def split_search_dev(self, d_keys, d_tasks, n, d_out):
        _check(load().shgpu_split_search(self.h, d_keys.ptr, d_tasks.ptr, n, d_out.ptr, None))

    def device_copy(self, d_dst, d_src, nbytes):
        _check(load().shgpu_device_copy(self.h, d_dst.ptr, d_src.ptr, nbytes))

    def array(self, host_array):
        return DeviceArray.from_host(self, host_array)

    # ---- host-array conveniences (copy in, run, copy out)
    def _stage(self, keys, tasks):
        d_keys = DeviceArray.from_host(self, np.asarray(keys, dtype=np.uint64))
        d_tasks = DeviceArray.from_host(self, np.asarray(tasks))
        d_out = DeviceArray(self, len(tasks), np.uint64)
        return d_keys, d_tasks, d_out

    def sh1_leaf_search(self, keys, offsets, sizes, starts, max_trials=1 << 40):
        tasks = np.zeros(len(sizes), dtype=LEAF_DTYPE)
        tasks["offset"], tasks["m"], tasks["start_seed"], tasks["max_trials"] = offsets, sizes, starts, max_trials
        d_keys, d_tasks, d_out = self._stage(keys, tasks)
        self.sh1_leaf_search_dev(d_keys, d_tasks, len(tasks), d_out)
        self.sync()
        return d_out.download()

    def sh2_leaf_search(self, keys, offsets, sizes):
        tasks = np.zeros(len(sizes), dtype=LEAF_DTYPE)
        tasks["offset"], tasks["m"] = offsets, sizes
        d_keys, d_tasks, d_out = self._stage(keys, tasks)
        self.sh2_leaf_search_dev(d_keys, d_tasks, len(tasks), d_out)
        self.sync()
        return d_out.download()

    def split_search(self, keys, tasks):
        """tasks: SPLIT_DTYPE array. Returns (seeds, permuted keys)."""
        d_keys, d_tasks, d_out = self._stage(keys, tasks)
        self.split_search_dev(d_keys, d_tasks, len(tasks), d_out)
        self.sync()
        return d_out.download(), d_keys.download()

    def split_search_noalloc_dev(self, d_keys, n_keys, d_tasks, n, d_out):
        """shgpu_split_search_noalloc on device arrays (no allocation inside the call)."""
        _check(load().shgpu_split_search_noalloc(self.h, d_keys.ptr, n_keys, d_tasks.ptr, n, d_out.ptr, None))

    def split_search_noalloc(self, keys, tasks, n_keys=None):
        """The same as split_search through the entry point that allocates nothing; n_keys defaults to len(keys)."""
        d_keys, d_tasks, d_out = self._stage(keys, tasks)
        self.split_search_noalloc_dev(d_keys, len(keys) if n_keys is None else n_keys, d_tasks, len(tasks), d_out)
        self.sync()
        return d_out.download(), d_keys.download()

    def build_dev(self, d_hashes, n, variant, leaf_size, bucket_size, mode="persistent", plain_bucket_limit=0,
                  copy_keys=True, offset=0, rotation=False):
        """Builds over n fingerprints at `offset` (in fingerprints) of a device array of uint64 pairs; rotation:
        ROTATION_FITTING for ShockHash / SIMDShockHash leaves of exactly leaf_size keys."""
        assert n == 0 or (offset + n) * 16 <= d_hashes.nbytes
        desc = BuildDesc(VARIANTS[variant], leaf_size, bucket_size, MODES[mode], plain_bucket_limit, int(copy_keys),
                         int(rotation))
        h = P()
        ptr = P((d_hashes.ptr.value or 0) + offset * 16)
        _check(load().shgpu_build_run(self.h, ctypes.byref(desc), ptr, n, ctypes.byref(h)))
        return Build(h)

    def hash_strings(self, strings):
        """SpookyHash V2 fingerprints of byte strings computed on the device (shgpu_hash_strings): an (n, 2)
        uint64 array of (first, second)."""
        n = len(strings)
        lens = np.array([len(x) for x in strings], dtype=np.uint64)
        offsets = np.zeros(n + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(lens)
        data = np.frombuffer(b"".join(strings) + b"\0" * 8, dtype=np.uint8)
        d_data = DeviceArray.from_host(self, data)
        d_off = DeviceArray.from_host(self, offsets)
        d_out = DeviceArray(self, 2 * n + 2, np.uint64)
        try:
            _check(load().shgpu_hash_strings(self.h, d_data.ptr, d_off.ptr, n, d_out.ptr))
            out = d_out.download()[: 2 * n].reshape(n, 2)
        finally:
            d_data.free()
            d_off.free()
            d_out.free()
        return out

    def ribbon_build(self, keys, bits):
        """One-bit retrieval structure of distinct keys and their bits, built on the device (shgpu_ribbon_build):
        (slot offsets, seed bytes, packed slot words)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        n = len(keys)
        d_k = DeviceArray.from_host(self, keys if n else np.zeros(1, np.uint64))
        d_b = DeviceArray.from_host(self, bits if n else np.zeros(1, np.uint8))
        lib = load()
        shards = lib.shgpu_ribbon_shards(n)
        offsets = np.zeros(shards + 1, dtype=np.uint32)
        seeds = np.zeros(shards, dtype=np.uint8)
        words = np.zeros(lib.shgpu_ribbon_max_words(n), dtype=np.uint64)
        used = ctypes.c_uint64()
        try:
            _check(lib.shgpu_ribbon_build(self.h, d_k.ptr, d_b.ptr, n, offsets.ctypes.data_as(P), seeds.ctypes.data_as(P),
                                          words.ctypes.data_as(P), len(words), ctypes.byref(used)))
        finally:
            d_k.free()
            d_b.free()
        return offsets, seeds, words[: used.value]

    def build(self, hashes, variant, leaf_size, bucket_size, **kw):
        hashes = np.ascontiguousarray(hashes, dtype=np.uint64).reshape(-1, 2)
        d = DeviceArray.from_host(self, hashes.reshape(-1))
        try:
            return self.build_dev(d, hashes.shape[0], variant, leaf_size, bucket_size, **kw)
        finally:
            d.free()


class Server:
    """Continuous-batching construction server (shgpu_server_*): concurrent builds in `slots` streams,
    admitted in arrival order within an HBM budget. Each result equals Engine.build of the same keys."""

    def __init__(self, engine, variant, leaf_size, bucket_size, slots=4, hbm_budget_bytes=8 << 30, copy_keys=False,
                 max_batch_keys=0, max_batch_requests=0):
        self.engine = engine
        self._pending = {}  # id -> host fingerprints (kept alive until taken)
        cfg = ServerConfig(VARIANTS[variant], leaf_size, bucket_size, slots, hbm_budget_bytes, int(copy_keys),
                           max_batch_keys, max_batch_requests)
        h = P()
        _check(load().shgpu_server_create(engine.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def submit(self, hashes):
        hashes = np.ascontiguousarray(hashes, dtype=np.uint64).reshape(-1, 2)
        rid = ctypes.c_uint64()
        _check(load().shgpu_server_submit(self.h, hashes.ctypes.data_as(P), hashes.shape[0], ctypes.byref(rid)))
        self._pending[rid.value] = hashes
        return rid.value

    def submit_device(self, d_hashes, n, offset=0):
        """Fingerprints n at `offset` (in fingerprints) of a DeviceArray of uint64 pairs."""
        assert (offset + n) * 16 <= d_hashes.nbytes
        rid = ctypes.c_uint64()
        ptr = P((d_hashes.ptr.value or 0) + offset * 16)
        _check(load().shgpu_server_submit_device(self.h, ptr, n, ctypes.byref(rid)))
        self._pending[rid.value] = d_hashes
        return rid.value

    def wait(self, rid, timeout_s=-1.0):
        st = ctypes.c_int()
        _check(load().shgpu_server_wait(self.h, rid, timeout_s, ctypes.byref(st)))
        return st.value

    def take(self, rid):
        """Waits for request `rid` and returns (Build, latency_ms, queue_ms)."""
        self.wait(rid)
        h, lat, q = P(), ctypes.c_double(), ctypes.c_double()
        try:
            _check(load().shgpu_server_take(self.h, rid, ctypes.byref(h), ctypes.byref(lat), ctypes.byref(q)))
        finally:
            self._pending.pop(rid, None)
        return Build(h), lat.value, q.value

    def request_bytes(self, n):
        v = ctypes.c_uint64()
        _check(load().shgpu_server_request_bytes(self.h, n, ctypes.byref(v)))
        return v.value

    def info(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(load().shgpu_server_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("queued", "running", "reserved_bytes", "peak_reserved_bytes"), (x.value for x in v)))

    def close(self):
        if getattr(self, "h", None):
            load().shgpu_server_destroy(self.h)
            self.h = None
        self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def parse_schedule(text):
    """Size classes of a schedule JSON text (shgpu_schedule_parse; no device needed)."""
    lib, cnt = load(), ctypes.c_uint32()
    raw = text.encode() if isinstance(text, str) else text
    _check(lib.shgpu_schedule_parse(raw, None, 0, ctypes.byref(cnt)))
    arr = (ScheduleClass * max(1, cnt.value))()
    _check(lib.shgpu_schedule_parse(raw, arr, len(arr), ctypes.byref(cnt)))
    return [_class_dict(arr[i]) for i in range(cnt.value)]
