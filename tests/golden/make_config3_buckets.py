"""Writes config3_simd58_buckets.npz: the oracle's node codes of 32 buckets of the config-3 workload,
SIMDShockHash<58> with bucket size 2000 over 10,000,000 uniform fingerprints (workload.fingerprints(10**7,
seed=1)), i.e. every split and leaf seed of those buckets' trees in DFS pre-order (the tree the reference's
recSplit builds, include/SIMDShockHash.hpp:563-617, with leafLevel's search, :507-542). The whole workload is
about 10^14 hash evaluations, core-hours for the CPU; these 32 buckets (every 156th, ~64,000 keys) take a few
minutes on 8 threads. The GPU test builds all 10,000,000 keys and compares these buckets node for node.

Run from the repository root after oracle/ is built:  python tests/golden/make_config3_buckets.py [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from shockhash_amd import workload  # noqa: E402

N, SEED, BUCKET, LEAF = 10_000_000, 1, 2000, 58
BUCKETS = np.arange(32, dtype=np.uint32) * 156 + 7


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 8)
    h = workload.fingerprints(N, seed=SEED)
    t0 = time.time()
    ob = O.OracleBucketBuild("simd", LEAF, h, BUCKET, BUCKETS, threads=threads)
    print(f"{len(BUCKETS)} buckets, {len(ob.node_code)} nodes in {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "config3_simd58_buckets.npz"), buckets=BUCKETS, node_bucket=ob.node_bucket,
                        node_m=ob.node_m, node_kind=ob.node_kind, node_code=ob.node_code, n=N, seed=SEED,
                        bucket_size=BUCKET, leaf=LEAF)


if __name__ == "__main__":
    main()
