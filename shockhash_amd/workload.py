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
    """n uniform 128-bit fingerprints as an (n, 2) uint64 array of (first, second).

    The .second values are distinct by construction: they are outputs 2, 4, ..., 2n of one SplitMix64
    stream, whose states seed + k * GAMMA are distinct for k < 2^64 (GAMMA is odd) and Mix13 is a
    bijection (xor-shifts and odd multipliers are invertible mod 2^64). No leaf can hold two equal keys,
    on which the seed searches would run to their bounds."""
    return splitmix64(seed, 2 * n).reshape(n, 2)


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


def bucket_sizes(hashes, bucket_size):
    """Cumulative bucket sizes (nbuckets + 1 entries) of fingerprints partitioned as the constructors do:
    nbuckets = max(1, ceil(n / bucket_size)), bucket = floor(first * nbuckets / 2^64) (remap128). Bucket b holds
    the fingerprints with first >= ceil(b * 2^64 / nbuckets), so the bounds are found by searching the sorted
    .first values for those thresholds (exact integer arithmetic, no 128-bit products over the keys)."""
    first = np.sort(np.asarray(hashes, dtype=np.uint64).reshape(-1, 2)[:, 0])
    n = len(first)
    nb = max(1, -(-n // bucket_size))
    acc = np.zeros(nb + 1, dtype=np.uint64)
    thr = np.array([-(-(b << 64) // nb) for b in range(1, nb)], dtype=np.uint64)
    acc[1:nb] = np.searchsorted(first, thr, side="left")
    acc[nb] = n
    return acc
