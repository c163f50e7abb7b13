"""MPHF construction throughput of the MI355X ShockHash engine.

Workload (one step = one whole construction): bipartite ShockHash2, leaf size 64, bucket size 2000,
10,000,000 uniform random 128-bit fingerprints resident in HBM before the timed region. A step sorts the
fingerprints, buckets them, plans every RecSplit tree and searches every split and leaf seed on the
GPU (one persistent kernel with a device work queue); the per-node codes come back to the host. The
same step in the plain form (one kernel launch per tree node) is timed on the first buckets, the CPU
oracle is timed on a bounded sample, and the GPU codes are checked against the oracle on that sample.
Two more legs follow: the ShockHash2<64> class constructor over the same keys from host memory (upload,
device search, orientation and retrieval, host encoding; every key queried, descriptor words compared
with the oracle's), and a serving mix of 2048 Poisson-arriving requests on the continuous-batching
server (p50/p99 latency).

Prints one JSON line (see the keys below). One GPU only: --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HBM bytes per k_build launch on this workload: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of
# this script, scaled by the same counters on the 256 MiB copy probe (profiles/k_build_shockhash2_L64_10M.md).
MEASURED_TRAFFIC_BYTES = 3.154e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


VARIANT_NAMES = {"sh2": "ShockHash2", "sh1": "ShockHash", "simd": "SIMDShockHash"}


def O_build(h, threads, variant="sh2", leaf=64):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O.OracleBuild(variant, leaf, h, 2000, threads=threads)


def cpu_baseline(n_sample, threads, variant="sh2", leaf=64):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from shockhash_amd import workload
    h = workload.fingerprints(n_sample, seed=7)
    t0 = time.perf_counter()
    ob = O.OracleBuild(variant, leaf, h, 2000, threads=threads)
    dt = time.perf_counter() - t0
    return h, ob, dt


def serving_mix(eng, args):
    """Serving mix: args.serve_requests build requests, Poisson arrivals at args.serve_rate requests/s,
    sizes log-uniform in [1e3, 1e6], fingerprints resident in HBM, each built as ShockHash2 leaf 64 /
    bucket 2000 by the continuous-batching server. Latency = submission to completion."""
    from shockhash_amd import workload
    from shockhash_amd.engine import DeviceArray, Server
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e6), args.serve_requests)).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, args.serve_requests))
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    total = int(offs[-1])
    # one SplitMix64 stream (seed 2), generated in chunks straight into the device array
    d = DeviceArray(eng, 2 * total, np.uint64)
    chunk = 1 << 24
    for a in range(0, total, chunk):
        b = min(total, a + chunk)
        d.upload_at(2 * a, workload.splitmix64(2, 2 * (b - a), offset=2 * a))
    log(f"serving mix: {args.serve_requests} requests, {total} keys resident, rate {args.serve_rate}/s, "
        f"{args.serve_slots} slots, budget {args.serve_budget_gib} GiB")
    srv = Server(eng, "sh2", 64, 2000, slots=args.serve_slots, hbm_budget_bytes=int(args.serve_budget_gib * (1 << 30)),
                 max_batch_keys=args.serve_batch_keys)
    # warm the slots (first-use buffer allocation)
    for rid in [srv.submit_device(d, int(sizes[i]), int(offs[i])) for i in range(args.serve_slots)]:
        srv.take(rid)
    ids = []
    t0 = time.perf_counter()
    for i in range(args.serve_requests):
        wait = t0 + arrivals[i] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit_device(d, int(sizes[i]), int(offs[i])))
    results = [srv.take(rid) for rid in ids]
    wall = time.perf_counter() - t0
    info = srv.info()
    lat = np.array([r[1] for r in results])
    queue = np.array([r[2] for r in results])
    # identical alone and in the mix: rebuild a sample of the requests on their own
    check = sorted(set([int(np.argmin(sizes)), int(np.argmax(sizes))] + list(rng.choice(len(sizes), 6, replace=False))))
    same = True
    for i in check:
        alone = eng.build_dev(d, int(sizes[i]), "sh2", 64, 2000, copy_keys=False, offset=int(offs[i]))
        same = same and bool(np.array_equal(alone.codes, results[i][0].codes)) and \
            bool(np.array_equal(alone.size_acc, results[i][0].size_acc))
    srv.close()
    d.free()
    out = {"requests": args.serve_requests, "rate_per_s": args.serve_rate, "keys": total, "slots": args.serve_slots,
           "batch_keys": args.serve_batch_keys,
           "hbm_budget_bytes": int(args.serve_budget_gib * (1 << 30)), "peak_reserved_bytes": info["peak_reserved_bytes"],
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max()), "queue_p99_ms": float(np.percentile(queue, 99)),
           "keys_per_s": total / wall, "wall_s": wall, "identical_alone_sample": len(check), "identical": same}
    log(f"serving: p50 {out['p50_ms']:.2f} ms, p99 {out['p99_ms']:.2f} ms, {out['keys_per_s'] / 1e6:.1f} M keys/s, "
        f"identical alone and in the mix: {same}")
    return out


def class_build(h, args, ob):
    """The drop-in class constructor ShockHash2<64>(fingerprints, 2000) on the same 10M host fingerprints:
    upload, device sort/plan/search, orientation and retrieval, host Golomb-Rice + Elias-Fano. The second of
    two builds is reported; every key is queried, and the descriptor words are compared with the oracle's
    when the CPU baseline built the same keys."""
    from shockhash_amd import host
    for rep in range(2):
        f = host.MPHF("sh2", 64, h, 2000, threads=1)
        t = f.timings_ms
        if rep == 0:
            f.close()
    t0 = time.perf_counter()
    q = f(h)
    query_s = time.perf_counter() - t0
    bij = bool(len(q) == len(h) and int(q.max()) < len(h) and len(np.unique(q)) == len(h))
    desc_eq = None if ob is None else bool(np.array_equal(f.descriptor_words, ob.words))
    out = {"constructor_ms": t["constructor"], "keys_per_s": len(h) / (t["constructor"] / 1e3),
           "phases_ms": {k: float(t[k]) for k in ("sort", "tree", "search", "rice_ef", "orientation", "retrieval")},
           "bits_per_key": f.bits / len(h), "queries_bijective": bij, "query_ns_per_key": 1e9 * query_s / len(h),
           "descriptors_equal_oracle": desc_eq}
    f.close()
    log(f"class build: ShockHash2<64> constructor {t['constructor']:.1f} ms ({out['keys_per_s'] / 1e6:.0f} M keys/s), "
        f"phases {out['phases_ms']}, {out['bits_per_key']:.3f} bits/key, queries bijective {bij}, "
        f"descriptors equal the oracle's: {desc_eq}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--variant", choices=sorted(VARIANT_NAMES), default="sh2",
                    help="class variant of the build (the headline workload is sh2 with leaf 64)")
    ap.add_argument("--leaf", type=int, default=64, help="leaf size")
    ap.add_argument("--plain-buckets", type=int, default=100, help="plain-form buckets (0: no plain-form run)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="CPU oracle sample size (0: the whole workload)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--split-seeds-per-lane", type=int, default=None)
    ap.add_argument("--leaf-seeds-per-lane", type=int, default=None)
    ap.add_argument("--waves-per-cu", type=int, default=None)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2048)
    ap.add_argument("--serve-rate", type=float, default=800.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=1,
                    help="concurrent builds; two or more overlap kernels that then stretch each other's tails")
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-batch-keys", type=int, default=1 << 22, help="keys a slot builds together at most")
    ap.add_argument("--no-class-build", action="store_true",
                    help="skip the ShockHash2<64> class-constructor timing (host fingerprints in, queries out)")
    ap.add_argument("--copy-probe-mib", type=int, default=256,
                    help="size of the plain copy kernel run once as the HBM-counter calibration reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the construction engine does not shard across GPUs "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from shockhash_amd import workload
    from shockhash_amd.engine import Engine
    from shockhash_amd.roofline import roofline

    eng = Engine(0)
    tune = {k: v for k, v in (("split_seeds_per_lane", args.split_seeds_per_lane),
                                ("sh1_seeds_per_lane", args.leaf_seeds_per_lane),
                                ("waves_per_cu", args.waves_per_cu)) if v is not None}
    V, L = args.variant, args.leaf
    headline = V == "sh2" and L == 64
    if tune:
        eng.set_variant_tuning(V, **tune)
    log(f"device {eng.arch} CUs {eng.cu_count} clock {eng.clock_khz / 1e6:.2f} GHz; "
        f"{VARIANT_NAMES[V]} schedule {eng.variant_tuning(V)}")
    h = workload.fingerprints(args.n, seed=1)
    d_h = eng.array(h.reshape(-1))  # fingerprints resident in HBM
    log(f"{args.n} fingerprints resident on the device")
    if args.copy_probe_mib:
        # known-size reference for FETCH_SIZE / WRITE_SIZE calibration under rocprofv3 --pmc
        from shockhash_amd.engine import DeviceArray
        nb = args.copy_probe_mib << 20
        src, dst = DeviceArray(eng, nb // 8, np.uint64), DeviceArray(eng, nb // 8, np.uint64)
        t0 = time.perf_counter()
        eng.device_copy(dst, src, nb)
        log(f"copy probe: {nb >> 20} MiB in {1e3 * (time.perf_counter() - t0):.2f} ms (k_copy_u64)")
        src.free()
        dst.free()

    for i in range(args.warmup):
        b = eng.build_dev(d_h, args.n, V, L, 2000, copy_keys=False)
        log(f"warmup {i}: timings {np.round(b.timings_ms, 2)}")
    step_s, kernel_ms = [], []
    for i in range(args.steps):
        eng.sync()
        t0 = time.perf_counter()
        b = eng.build_dev(d_h, args.n, V, L, 2000, copy_keys=False)
        eng.sync()
        step_s.append(time.perf_counter() - t0)
        kernel_ms.append(b.timings_ms[3])
        log(f"step {i}: {step_s[-1] * 1e3:.1f} ms (sort {b.timings_ms[0]:.1f}, tree {b.timings_ms[1]:.1f}, "
            f"search {b.timings_ms[2]:.1f}, kernel {b.timings_ms[3]:.1f})")
    ms_per_step = 1e3 * float(np.mean(step_s))
    value = args.n / (ms_per_step / 1e3)
    kms = float(np.mean(kernel_ms))
    roof = roofline(b.nodes, b.codes, V, kms, MEASURED_TRAFFIC_BYTES if headline and args.n == 10_000_000 else None)

    # plain form: one launch per node over the first buckets, same device input
    plain_value, n_plain_keys, plain_match = None, 0, True
    if args.plain_buckets:
        t0 = time.perf_counter()
        bp = eng.build_dev(d_h, args.n, V, L, 2000, mode="plain", plain_bucket_limit=args.plain_buckets,
                           copy_keys=False)
        plain_s = time.perf_counter() - t0
        n_plain_keys = int(b.size_acc[args.plain_buckets])
        nn = len(bp.codes)
        plain_match = bool(np.array_equal(bp.codes, b.codes[:nn]))
        # the plain run also sorts all keys; charge it only for the searched buckets' share of sort time
        plain_search_s = (bp.timings_ms[1] + bp.timings_ms[2]) / 1e3 + bp.timings_ms[0] / 1e3 * n_plain_keys / args.n
        plain_value = n_plain_keys / plain_search_s
        log(f"plain form: {args.plain_buckets} buckets ({n_plain_keys} keys) in {plain_s:.2f} s, codes equal: {plain_match}")

    cpu = None
    bit_exact = plain_match
    if not args.no_cpu_baseline:
        n_cpu = args.cpu_sample or args.n
        if n_cpu == args.n:
            hs = h
            t0 = time.perf_counter()
            ob = O_build(hs, args.cpu_threads, V, L)
            dt = time.perf_counter() - t0
            bs = b  # the timed GPU build of the same keys
            sample = f"the whole workload ({args.n} keys, seed 1)"
        else:
            hs, ob, dt = cpu_baseline(n_cpu, args.cpu_threads, V, L)
            bs = eng.build(hs, V, L, 2000)
            sample = f"{n_cpu} fingerprints (same generator, seed 7)"
        bit_exact = bit_exact and bool(np.array_equal(bs.codes, ob.node_code)) and \
            bool(np.array_equal(bs.size_acc, ob.size_acc))
        cpu = {"value": n_cpu / dt, "unit": "keys/s", "cores": args.cpu_threads, "kind": "port",
               "sample": f"{sample}, {VARIANT_NAMES[V]} leaf {L}, bucket 2000, oracle C++ restatement, "
                         f"{args.cpu_threads} threads, {dt:.1f} s"}
        log(f"cpu oracle: {cpu['value']:.0f} keys/s; GPU codes equal the oracle's: {bit_exact}")

    cls = None
    if args.no_class_build or not headline:
        log("class-constructor leg: skipped" + ("" if headline else " (ShockHash2 leaf 64 only)"))
    else:
        cls = class_build(h, args, ob if (not args.no_cpu_baseline and (args.cpu_sample or args.n) == args.n) else None)
        bit_exact = bit_exact and cls["queries_bijective"] and cls["descriptors_equal_oracle"] is not False

    serving = None if args.no_serving or not headline else serving_mix(eng, args)
    if serving is not None:
        bit_exact = bit_exact and serving["identical"]

    out = {
        "metric": "MPHF build keys/s at 10M keys (bit-exact) vs plain GPU form; p99 build latency",
        "value": value,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"{VARIANT_NAMES[V]} leaf {L}, bucket 2000, "
                               f"{f'{args.n // 1_000_000}M' if args.n % 1_000_000 == 0 else args.n} uniform 128-bit fingerprints "
                               "(SplitMix64, seed 1), persistent device-queue build", "keys": args.n, "leaf_size": L,
                   "bucket_size": 2000, "variant": VARIANT_NAMES[V]},
        "plain_form": {"keys_per_s": plain_value, "buckets": args.plain_buckets, "keys": n_plain_keys,
                       "codes_equal_persistent": plain_match},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "bit_exact": bit_exact,
        "search_kernel_ms": kms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "class_build": cls,
        "p99_build_latency_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
