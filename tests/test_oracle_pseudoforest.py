"""The cuckoo/union-find predicates of the reference agree with the textbook pseudoforest test, so any
exact pseudoforest predicate (the GPU's) reproduces the reference's seed choice."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("native") / "pf"
    lib_dir = os.path.dirname(O.LIB_PATH)
    subprocess.run(["g++", "-O2", "-o", str(out), os.path.join(HERE, "native", "pseudoforest_exhaustive.cpp"),
                    f"-L{lib_dir}", "-loracle", f"-Wl,-rpath,{lib_dir}"], check=True)
    return str(out)


def test_exhaustive_small_multigraphs(checker):
    # every sequence of <= 5 edges on <= 4 vertices and <= 4 edges on 5 vertices
    for n, k in [(4, 5), (5, 4)]:
        r = subprocess.run([checker, str(n), str(k)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr


def textbook_prefix(a, b, nc):
    parent = list(range(nc))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x
    edges = {}
    for i, (x, y) in enumerate(zip(a, b)):
        rx, ry = find(x), find(y)
        if rx != ry:
            parent[rx] = ry
        # recount the component of y
        r = find(y)
        comp = [v for v in range(nc) if find(v) == r]
        e = sum(1 for j in range(i + 1) if find(a[j]) == r)
        if e > len(comp):
            return i
    return len(a)


@pytest.mark.parametrize("shape", ["sh1", "sh2_even", "sh2_odd"])
def test_random_leaf_shaped_graphs(shape):
    """Random graphs of the leaf shapes (m <= 64 cells and edges): bipartite ShockHash1 cells, and
    ShockHash2's shifted halves (which overlap in one cell for odd m)."""
    rng = np.random.default_rng({"sh1": 1, "sh2_even": 2, "sh2_odd": 3}[shape])
    L = O.lib()
    for trial in range(300):
        m = int(rng.integers(2, 65))
        if shape == "sh2_odd" and m % 2 == 0:
            m += 1 if m < 64 else -1
        if shape == "sh2_even" and m % 2 == 1:
            m -= 1
        m = max(m, 2)
        if shape == "sh1":
            a = rng.integers(0, m // 2, size=m)
            b = rng.integers(m // 2, m, size=m)
        else:
            H = (m + 1) // 2
            a = rng.integers(0, H, size=m) + m // 2
            b = rng.integers(0, H, size=m)
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        want = textbook_prefix(list(a), list(b), m)
        for p in range(3):
            got = L.orc_pseudoforest_prefix(p, O.ptr(a), O.ptr(b), m, m)
            assert got == want, (shape, m, p)
