"""The compiled ShockHash2 seam (libShockHash2Precompiled.so) without a GPU: it exports the reference's C++
signatures (include/shockhash/ShockHash2-precompiled.h:8-10), a program written against them compiles and links
(tests/native/seam_driver.cpp), and shockhash2query, which is pure host code, equals the oracle's query."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "shockhash_amd", "_build", "libShockHash2Precompiled.so")
DRIVER = os.path.join(ROOT, "tests", "native", "_build", "seam_driver")


def test_exports_the_reference_signatures():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    assert ("shockhash::shockhash2construct(unsigned long, std::vector<unsigned long, std::allocator<unsigned long> >&, "
            "std::vector<std::pair<unsigned long, unsigned char>, std::allocator<std::pair<unsigned long, unsigned char> > >&)"
            in out)
    assert "shockhash::shockhash2query(unsigned long, unsigned long, unsigned long, unsigned long)" in out


def test_program_against_the_signatures_links(tmp_path):
    """A program that declares the two functions exactly as the reference's header does links against the
    library (the C++ ABI of the seam, not just its names)."""
    src = tmp_path / "decl.cpp"
    src.write_text('#include <cstddef>\n#include <cstdint>\n#include <vector>\n#include <utility>\n'
                   "namespace shockhash {\n"
                   "    size_t shockhash2query(size_t param, size_t seed, uint64_t key, size_t retrieved);\n"
                   "    size_t shockhash2construct(size_t param, std::vector<uint64_t> &leafKeys,\n"
                   "                               std::vector<std::pair<uint64_t, uint8_t>> &ribbonInput);\n"
                   "}\n"
                   "int main() { std::vector<uint64_t> k; std::vector<std::pair<uint64_t, uint8_t>> r;\n"
                   "  return (int)(shockhash::shockhash2construct(1, k, r) + shockhash::shockhash2query(1, 5, 7, 0)); }\n")
    exe = tmp_path / "decl"
    lib_dir = os.path.dirname(LIB)
    subprocess.run(["g++", "-std=c++17", "-o", str(exe), str(src), f"-L{lib_dir}", "-lShockHash2Precompiled",
                    f"-Wl,-rpath,{lib_dir}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_trivial_leaf_sizes_need_no_search():
    """param 0 and 1 (dispatchLeafSize<I <= 1>) and past 128 return 0 and append nothing
    (src/ShockHash2-precompiled.cpp:24-47)."""
    r = subprocess.run([DRIVER, "trivial"], capture_output=True, text=True, check=True)
    assert r.stdout.split("\n")[:3] == ["0 0 0", "1 0 0", "129 0 0"]


def test_query_equals_oracle():
    rng = np.random.default_rng(3)
    n = 4000
    params = rng.integers(0, 130, n).astype(np.uint64)
    seeds = rng.integers(0, 1 << 40, n, dtype=np.uint64)
    keys = workload.fingerprints(n, seed=9)[:, 1]
    retrieved = rng.integers(0, 2, n).astype(np.uint64)
    inp = np.stack([params, seeds, keys, retrieved], axis=1).astype("<u8").tobytes()
    out = np.frombuffer(subprocess.run([DRIVER, "query"], input=inp, capture_output=True, check=True).stdout, dtype="<u8")
    L = O.lib()
    want = np.array([0 if p > 128 else L.orc_sh2_query_cell(int(p), int(s), int(k), int(b))
                     for p, s, k, b in zip(params, seeds, keys, retrieved)], dtype=np.uint64)
    assert np.array_equal(out, want)
