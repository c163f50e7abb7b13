"""Autotunes the build schedule's size classes on this GPU and writes them as JSON.

    python -m shockhash_amd.tune --out shockhash_amd/schedule_gfx950.json

For each variant (ShockHash leaf 30, SIMDShockHash leaf 58, ShockHash2 leaf 64; bucket 2000) and each sample size,
shgpu_autotune_class builds a sample of uniform fingerprints under the candidate schedules (waves per CU, then seeds
per lane at the best wave count), checks that every candidate gives the same codes, and keeps the one with the
shortest search kernel. SIMDShockHash leaves carry ~1000x the work per key of the others (a 1M-key build takes ~6
s), so its samples are smaller.
A class covers batches up to the geometric mean of its sample size and the next one; the largest class covers
every larger batch. The engine loads the written file when it is created (schedule_<arch>.json beside the library
or in its parent directory, or the file named by SHGPU_SCHEDULE).
"""
import argparse
import math
import sys
import time

from . import workload
from .engine import DeviceArray, Engine

U64_MAX = (1 << 64) - 1

# variant, leaf size, sample sizes (keys)
PLAN = [
    ("sh2", 64, [200_000, 500_000, 1_000_000, 2_000_000, 4_000_000, 10_000_000]),
    ("sh1", 30, [200_000, 1_000_000, 2_000_000, 4_000_000, 10_000_000]),
    ("simd", 58, [200_000, 500_000]),
]


def class_bounds(sizes):
    """max_keys of each sample size's class: the geometric mean with the next size, the last unbounded."""
    return [int(math.sqrt(a * b)) for a, b in zip(sizes, sizes[1:])] + [U64_MAX]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True, help="JSON file to write")
    ap.add_argument("--bucket", type=int, default=2000)
    ap.add_argument("--variants", default="sh2,sh1,simd", help="comma-separated subset of sh2, sh1, simd")
    ap.add_argument("--max-keys", type=int, default=0, help="skip samples larger than this (0: all)")
    args = ap.parse_args(argv)
    eng = Engine()
    wanted = set(args.variants.split(","))
    tuned = []
    for variant, leaf, sizes in PLAN:
        if variant not in wanted:
            continue
        if args.max_keys:
            sizes = [n for n in sizes if n <= args.max_keys] or sizes[:1]
        for n, max_keys in zip(sizes, class_bounds(sizes)):
            fp = workload.fingerprints(n, seed=11)
            d = DeviceArray.from_host(eng, fp.reshape(-1))
            t0 = time.perf_counter()
            try:
                c = eng.autotune_class(d, n, variant, leaf, args.bucket, max_keys)
            finally:
                d.free()
            print(f"{variant} leaf {leaf}, {n} keys (class up to {max_keys if max_keys < U64_MAX else 'any'}): "
                  f"{c['waves_per_cu']} waves/CU, seeds per lane leaf {c['sh1_seeds_per_lane']} split "
                  f"{c['split_seeds_per_lane']}, kernel {c['tuned_ms']:.3f} ms (tuning took {time.perf_counter() - t0:.1f} s)",
                  flush=True)
            tuned.append(c)
    eng.set_schedule(tuned)
    eng.save_schedule(args.out)
    print(f"wrote {args.out}: {len(tuned)} classes", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
