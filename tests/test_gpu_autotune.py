"""Autotuned build schedule on the GPU: a size class tuned on a 200k-key sample gives the same codes as other
schedules, the table round-trips through its JSON file, and a new engine loads the file named by SHGPU_SCHEDULE,
the shipped tuned table by default, and rejects a broken file."""
import os

import numpy as np
import pytest

from shockhash_amd import workload
from shockhash_amd.engine import ERR_ARG, DeviceArray, Engine, ShgpuError

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIPPED = os.path.join(ROOT, "shockhash_amd", "schedule_gfx950.json")


def log(msg):
    print(msg, flush=True)


class env_schedule:
    """Sets SHGPU_SCHEDULE while a block runs (engines read it when they are created)."""

    def __init__(self, value):
        self.value, self.old = value, None

    def __enter__(self):
        self.old = os.environ.get("SHGPU_SCHEDULE")
        os.environ["SHGPU_SCHEDULE"] = self.value

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("SHGPU_SCHEDULE", None)
        else:
            os.environ["SHGPU_SCHEDULE"] = self.old


@pytest.fixture(scope="module")
def sample():
    n = 200_000
    with env_schedule("none"):
        eng = Engine()
    fp = workload.fingerprints(n, seed=21)
    d = DeviceArray.from_host(eng, fp.reshape(-1))
    yield eng, d, n
    d.free()
    eng.close()


def test_tuned_class_keeps_the_codes(sample, tmp_path):
    eng, d, n = sample
    assert eng.schedule_source == "built-in"
    before = eng.build_dev(d, n, "sh2", 64, 2000, copy_keys=False).codes
    log(f"  tuning ShockHash2 leaf 64 on {n} keys")
    c = eng.autotune_class(d, n, "sh2", 64, 2000, 300_000)
    log(f"  chosen: {c}")
    assert c["variant"] == 2 and c["max_keys"] == 300_000 and c["sample_keys"] == n and c["tuned_ms"] > 0
    assert c["waves_per_cu"] in (4, 8, 12, 16, 24) and c["split_seeds_per_lane"] in (1, 2, 4)
    mine = [x for x in eng.schedule if x["variant"] == 2 and x["max_keys"] == 300_000]
    assert mine == [c]
    # a build of this size takes the tuned class; the codes do not depend on the schedule
    tuned = eng.build_dev(d, n, "sh2", 64, 2000, copy_keys=False).codes
    assert np.array_equal(tuned, before)
    eng.set_variant_tuning("sh2", waves_per_cu=24, split_seeds_per_lane=1)
    try:
        other = eng.build_dev(d, n, "sh2", 64, 2000, copy_keys=False).codes
    finally:
        eng.set_variant_tuning("sh2", waves_per_cu=0)
    assert np.array_equal(other, before)
    # the table round-trips through its file, and a new engine loads the file SHGPU_SCHEDULE names
    path = str(tmp_path / "schedule.json")
    eng.save_schedule(path)
    with env_schedule(path):
        e2 = Engine()
    try:
        assert e2.schedule_source == path
        assert e2.schedule == eng.schedule
    finally:
        e2.close()


def test_sh1_class_keeps_the_codes(sample):
    eng, d, n = sample
    ref = eng.build_dev(d, n, "sh1", 30, 2000, copy_keys=False).codes
    c = eng.autotune_class(d, n, "sh1", 30, 2000, 300_000)
    log(f"  chosen: {c}")
    assert c["sh1_seeds_per_lane"] in (1, 2, 4)
    assert np.array_equal(eng.build_dev(d, n, "sh1", 30, 2000, copy_keys=False).codes, ref)


def test_engine_loads_the_shipped_schedule():
    assert os.path.exists(SHIPPED)
    old = os.environ.pop("SHGPU_SCHEDULE", None)
    try:
        eng = Engine()
        try:
            assert os.path.realpath(eng.schedule_source) == os.path.realpath(SHIPPED)
            assert {c["variant"] for c in eng.schedule} == {0, 1, 2}
            assert all(c["sample_keys"] > 0 for c in eng.schedule)
        finally:
            eng.close()
    finally:
        if old is not None:
            os.environ["SHGPU_SCHEDULE"] = old


def test_broken_schedule_file_is_an_error(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('{"arch": "gfx950", "classes": [{"variant": 2, "max_keys": 10, "waves_per_cu": 99, '
                   '"sh1_seeds_per_lane": 4, "split_seeds_per_lane": 2}]}')
    with env_schedule(str(bad)):
        with pytest.raises(ShgpuError) as ei:
            Engine()
    assert ei.value.status == ERR_ARG
    with env_schedule(str(tmp_path / "missing.json")):
        with pytest.raises(ShgpuError):
            Engine()
