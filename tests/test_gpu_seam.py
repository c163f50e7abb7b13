"""The compiled ShockHash2 seam on the GPU: shockhash2construct (libShockHash2Precompiled.so, the reference's C++
signatures) called through a C++ driver for leaves of every size, against the oracle's findSeed and
constructRetrieval (include/shockhash/ShockHash2-internal.h:501-572) for m = 2..128: the code, the appended retrieval
input and
the cells shockhash2query gives back, which form a bijection onto the leaf's cells."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "_build", "seam_driver")


def run_construct(leaves):
    inp = b"".join(np.uint32(len(k)).tobytes() + np.asarray(k, dtype="<u8").tobytes() for k in leaves)
    r = subprocess.run([DRIVER, "construct"], input=inp, capture_output=True, timeout=600)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-2000:]
    out, pos, res = r.stdout, 0, []
    for k in leaves:
        m = len(k)
        code = int(np.frombuffer(out, "<u8", 1, pos)[0])
        pos += 8
        rec = np.frombuffer(out, np.dtype([("key", "<u8"), ("bit", "u1")]), m, pos)
        pos += 9 * m
        cells = np.frombuffer(out, "<u8", m, pos)
        pos += 8 * m
        res.append((code, rec["key"].copy(), rec["bit"].copy(), cells.copy()))
    assert pos == len(out)
    return res


@pytest.mark.parametrize("sizes", [list(range(2, 129)), [64] * 40, [128] * 8])
def test_seam_construct_matches_oracle(sizes):
    print(f"seam: {len(sizes)} leaves, sizes {sizes[0]}..{sizes[-1]}", flush=True)
    keys = workload.fingerprints(int(sum(sizes)), seed=71 + len(sizes))[:, 1]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    leaves = [keys[o:o + m] for o, m in zip(offs, sizes)]
    got = run_construct(leaves)
    for k, (code, rk, rb, cells) in zip(leaves, got):
        m = len(k)
        assert code == O.sh2_find_seed(k), f"m={m}"
        assert np.array_equal(rk, k)
        assert np.array_equal(rb, O.sh2_retrieval(k, code)), f"m={m}"
        assert sorted(cells.tolist()) == list(range(m)), f"m={m}"
