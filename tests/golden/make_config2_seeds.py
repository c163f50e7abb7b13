"""Writes config2_seeds.npz: the oracle's minimal seed (from 0) of each of the 200,000 isolated ShockHash leaves
of 40 keys of the config-2 workload (workload.isolated_leaves(200000, 40, seed=11), the .second of uniform
fingerprints), as benchmark/bijections/ShockHash1.h:13-35 computes it. Takes a few minutes on 8 threads;
the GPU test compares every seed with this file.

Run from the repository root after oracle/ is built:  python tests/golden/make_config2_seeds.py [threads]
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

N_LEAVES, M, SEED = 200_000, 40, 11


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 8)
    keys = workload.isolated_leaves(N_LEAVES, M, seed=SEED)
    t0 = time.time()
    seeds = O.sh1_leaf_batch(keys, start=0, threads=threads)
    print(f"{N_LEAVES} leaves in {time.time() - t0:.0f} s, max seed {int(seeds.max())}", flush=True)
    np.savez_compressed(os.path.join(HERE, "config2_seeds.npz"), seeds=seeds, n_leaves=N_LEAVES, m=M, seed=SEED)


if __name__ == "__main__":
    main()
