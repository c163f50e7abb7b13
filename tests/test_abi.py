"""C ABI surface (no GPU): both libraries load, export every symbol their headers declare, and report
errors through status codes when no device can be used."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import ENGINE_LIB, HOST_LIB
from shockhash_amd import engine as E
from shockhash_amd import host as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shgpu_\w+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    names = declared(os.path.join(ROOT, "include", "shockhash_amd.h"))
    assert len(names) >= 20
    lib = ctypes.CDLL(ENGINE_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers them all
    assert set(names) <= set(E.exported_symbols())


def test_host_library_exports():
    lib = ctypes.CDLL(HOST_LIB)
    for n in H.exported_symbols():
        assert hasattr(lib, n), n


def test_status_strings():
    lib = E.load()
    for st in range(9):
        assert lib.shgpu_status_string(st)


def test_null_arguments_are_rejected_without_a_device():
    lib = E.load()
    assert lib.shgpu_engine_create(0, None) == E.ERR_ARG
    assert lib.shgpu_sh1_leaf_search(None, None, None, 1, None, None) == E.ERR_ARG
    assert lib.shgpu_build_run(None, None, None, 0, None) == E.ERR_ARG


@pytest.mark.parametrize("variant,leaf", [("sh1", 30), ("sh1", 8), ("simd", 58), ("simd", 24), ("sh2", 64),
                                          ("sh2", 16), ("sh2", 4)])
def test_compile_time_memo_equals_runtime_formula(variant, leaf):
    """The class templates' constexpr Golomb-Rice tables (g++ folding) equal the oracle's run-time
    evaluation of the same formula (libm), entry for entry."""
    assert np.array_equal(H.memo(variant, leaf), O.golomb_memo(variant, leaf))


def test_memo_leaf_entries_match_reference_tables():
    m1 = H.memo("sh1", 30)
    assert [int(m1[s] >> 27) for s in (0, 6, 10, 20, 30)] == [0, 1, 2, 7, 11]  # bij_memo, ShockHash.h:53-61
    m2 = H.memo("sh2", 64)
    assert [int(m2[s] >> 27) for s in (7, 9, 21, 40, 64)] == [3, 4, 10, 16, 26]  # bij_memo2, ShockHash2.h:33-47
    # subtree node counts: a leaf is one node, sizes 0/1 are none
    assert int((m2[1] >> 16) & 0x7FF) == 0 and int((m2[64] >> 16) & 0x7FF) == 1
