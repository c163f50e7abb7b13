"""Build schedule size classes (no GPU): the JSON form parses through the library's C ABI, invalid classes are
rejected, and the shipped tuned schedule covers every variant with valid classes."""
import json
import os

import pytest

from shockhash_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIPPED = os.path.join(ROOT, "shockhash_amd", "schedule_gfx950.json")

CLASS = {"variant": 2, "max_keys": 350000, "waves_per_cu": 4, "sh1_seeds_per_lane": 4, "split_seeds_per_lane": 2,
         "tuned_ms": 2.5, "sample_keys": 200000}


def doc(*classes, **extra):
    return json.dumps({"arch": "gfx950", **extra, "classes": list(classes)})


def test_parse_round_trip():
    last = dict(CLASS, max_keys=2**64 - 1, waves_per_cu=24, split_seeds_per_lane=4, tuned_ms=0.0, sample_keys=0)
    got = E.parse_schedule(doc(CLASS, last, note={"skipped": [1, 2, "x"]}))
    assert got == [CLASS, last]


def test_missing_optional_fields_default_to_untuned():
    c = {k: v for k, v in CLASS.items() if k not in ("tuned_ms", "sample_keys")}
    (got,) = E.parse_schedule(doc(c))
    assert got["tuned_ms"] == 0.0 and got["sample_keys"] == 0 and got["waves_per_cu"] == 4


@pytest.mark.parametrize("bad", [
    dict(CLASS, waves_per_cu=0), dict(CLASS, waves_per_cu=33), dict(CLASS, sh1_seeds_per_lane=3),
    dict(CLASS, split_seeds_per_lane=8), dict(CLASS, variant=3), dict(CLASS, max_keys=0),
    dict(CLASS, waves_per_cu=-4), dict(CLASS, max_keys=1.5),
    {k: v for k, v in CLASS.items() if k != "variant"},
])
def test_invalid_classes_are_rejected(bad):
    with pytest.raises(E.ShgpuError) as ei:
        E.parse_schedule(doc(bad))
    assert ei.value.status == E.ERR_ARG


@pytest.mark.parametrize("text", ["", "[]", "{", '{"arch": "gfx950"}', '{"classes": [}', '{"classes": []} x',
                                  '{"classes": [{"variant": 2,}]}'])
def test_malformed_json_is_rejected(text):
    with pytest.raises(E.ShgpuError):
        E.parse_schedule(text)


def test_shipped_schedule_covers_every_variant():
    text = open(SHIPPED).read()
    classes = E.parse_schedule(text)
    assert json.loads(text)["arch"] == "gfx950"
    for v in (0, 1, 2):
        mine = [c for c in classes if c["variant"] == v]
        assert mine, f"variant {v} has no class"
        sizes = [c["max_keys"] for c in mine]
        assert sizes == sorted(sizes) and len(set(sizes)) == len(sizes)
        # tuned on the GPU: every class records its sample and the kernel time it was chosen on
        assert all(c["sample_keys"] > 0 and c["tuned_ms"] > 0 for c in mine)
