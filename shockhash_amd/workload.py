"""Synthetic inputs of the ShockHash workloads, generated from a seed (no datasets).

Fingerprints are SplitMix64 outputs: fingerprint i = (sm(2i), sm(2i+1)) of the stream started at
`seed` (Vigna's splitmix64: state += 0x9e3779b97f4a7c15, output = Mix13(state)). The reference's own
benchmark draws keys from XorShift64 (benchmark/BenchmarkData.h:7-33, un-vendored); any uniform
128-bit source exercises the same code path.
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix13(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, count, offset=0):
    """Outputs offset .. offset+count-1 of the SplitMix64 stream seeded with `seed`."""
    idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + idx * GAMMA
    return mix13(state)


def fingerprints(n, seed=1):
    """n uniform 128-bit fingerprints as an (n, 2) uint64 array of (first, second); distinct .second."""
    out = splitmix64(seed, 2 * n).reshape(n, 2)
    # 64-bit collisions among 1e7 values are ~3e-6 likely; redraw any duplicate .second so that no
    # leaf can hold two equal keys (the reference would search forever on them).
    sec = out[:, 1]
    if len(np.unique(sec)) != n:
        _, first_idx = np.unique(sec, return_index=True)
        dup = np.setdiff1d(np.arange(n), first_idx)
        extra = splitmix64(seed ^ 0x5DEECE66D, 2 * len(dup) + 64)
        out[dup, 1] = extra[: len(dup)]
        assert len(np.unique(out[:, 1])) == n
    return out


def isolated_leaves(n_leaves, m, seed=2):
    """Keys of n_leaves independent leaves of m keys (the .second of uniform fingerprints)."""
    return fingerprints(n_leaves * m, seed)[:, 1].reshape(n_leaves, m).copy()


def random_strings(n, seed=3, min_len=10, max_len=50):
    """n distinct printable-ASCII strings of length uniform in [min_len, max_len]."""
    rng = np.random.default_rng(seed)
    out, seen = [], set()
    while len(out) < n:
        k = n - len(out)
        lens = rng.integers(min_len, max_len + 1, size=k)
        chars = rng.integers(33, 127, size=(k, max_len), dtype=np.uint8)
        for i in range(k):
            s = chars[i, : lens[i]].tobytes()
            if s not in seen:
                seen.add(s)
                out.append(s)
    return out
