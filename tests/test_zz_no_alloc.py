"""shgpu_split_search_noalloc (the split search that allocates nothing) against shgpu_split_search, through the
C ABI: the same seeds and the same partitioned keys, on two calls in a row (the second smaller, so the shared
scratch holds the first call's contents), plus its size limit and device-side task checks."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload
from shockhash_amd.engine import ERR_ARG, NOT_FOUND, SPLIT_DTYPE, ShgpuError

pytestmark = pytest.mark.gpu

NOALLOC_MAX_KEYS = 1 << 20  # SHGPU_NOALLOC_MAX_KEYS


def log(msg):
    print(msg, flush=True)


def split_tasks(variant, leaf, sizes, seed):
    """Split tasks of the given node sizes (the oracle's split geometry), back to back, and their keys."""
    rng = np.random.default_rng(seed)
    tasks, chunks, off = [], [], 0
    for i, m in enumerate(sizes):
        level = int(rng.integers(0, 5))
        keys = workload.fingerprints(m, seed=seed * 1000 + i)[:, 1]
        _, _, kind, fanout, unit, start = O.split(2 if variant == 0 else variant, leaf, level, keys)
        if fanout > 4:
            continue
        t = np.zeros(1, dtype=SPLIT_DTYPE)
        t["offset"], t["m"], t["start_seed"], t["max_trials"] = off, m, start, 1 << 26
        t["unit"], t["kind"], t["hash"] = unit, kind, 1 if variant == 1 else 0
        tasks.append(t)
        chunks.append(keys)
        off += m
    return np.concatenate(tasks), np.concatenate(chunks)


def test_noalloc_split_search_equals_split_search(engine):
    """Two calls in a row, the second with fewer tasks and keys: each equals shgpu_split_search on the same
    input (seeds and the stable partitions)."""
    cases = [(0, 64, [2000, 1536, 768, 700, 464, 256, 256, 200, 130, 128] * 3, 1),
             (1, 16, [512, 256, 130, 64, 48, 33], 2)]
    for variant, leaf, sizes, seed in cases:
        log(f"noalloc split search, variant {variant} leaf {leaf}, {len(sizes)} nodes")
        tasks, keys = split_tasks(variant, leaf, sizes, seed)
        want_seeds, want_keys = engine.split_search(keys, tasks)
        got_seeds, got_keys = engine.split_search_noalloc(keys, tasks)
        assert np.array_equal(got_seeds, want_seeds)
        assert np.array_equal(got_keys, want_keys)
        assert not np.any(got_seeds == np.uint64(NOT_FOUND))


def test_noalloc_size_limit_and_task_checks(engine):
    """More keys than the scratch holds is an argument error; a task reaching past n_keys or with a bad
    fanout is skipped on the device (keys untouched, SHGPU_NOT_FOUND) while the valid ones are searched."""
    log("noalloc limits")
    tasks, keys = split_tasks(0, 64, [768, 256, 256], 3)
    with pytest.raises(ShgpuError) as ex:
        engine.split_search_noalloc(keys, tasks, n_keys=NOALLOC_MAX_KEYS + 1)
    assert ex.value.status == ERR_ARG
    bad = tasks.copy()
    bad["offset"][1] = len(keys) - 100  # reaches past the end of the key array
    bad["unit"][2] = 0                  # no valid fanout
    want_seeds, want_keys = engine.split_search(keys, tasks[:1])
    got_seeds, got_keys = engine.split_search_noalloc(keys, bad)
    assert got_seeds[0] == want_seeds[0]
    assert got_seeds[1] == np.uint64(NOT_FOUND) and got_seeds[2] == np.uint64(NOT_FOUND)
    assert np.array_equal(got_keys[:768], want_keys[:768])
    assert np.array_equal(got_keys[768:], keys[768:])
