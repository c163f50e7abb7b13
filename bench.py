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

Three more workloads print lines of their own (not the headline): --workload strings times the ShockHash<30>
class constructor over 1,000,000 random strings (device fingerprinting, build, orientation, retrieval, host
encoding), --workload leaves times 200,000 isolated ShockHash leaves of 40 keys searched from seed 0 (hash
evaluations per second), --workload flat times the ShockHash2Flat<k> constructor over 10,000,000 key hashes. --leaf / --variant select other leaf sizes and classes for the build workload.

Prints one JSON line (see the keys below). One GPU only: --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HBM bytes per k_build launch on this workload: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of
# this script, scaled by the same counters on the 256 MiB copy probe (profiles/k_build_shockhash2_L64_10M.md).
MEASURED_TRAFFIC_BYTES = 2.375e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


VARIANT_NAMES = {"sh2": "ShockHash2", "sh1": "ShockHash", "simd": "SIMDShockHash"}


def host_cpu():
    """(CPU model, threads the baseline uses): the host's model name, and the threads this process may run on
    (its CPU share on a shared host: OMP_NUM_THREADS when set, else the affinity mask)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    return model, threads


def native_oracle():
    """Compiles the CPU oracle with -march=native for this host (the prebuilt library targets x86-64-v3 so that it
    runs anywhere) into a temporary directory; returns its path and a description, or the prebuilt library when
    the compile fails."""
    src = os.path.join(ROOT, "oracle", "shockhash_oracle.cpp")
    out = os.path.join(tempfile.mkdtemp(prefix="shh_oracle_"), "liboracle_native.so")
    t0 = time.perf_counter()
    r = subprocess.run(["g++", "-O3", "-march=native", "-std=c++20", "-fPIC", "-pthread", "-shared", "-o", out, src],
                       capture_output=True, text=True)
    if r.returncode != 0:
        log(f"native oracle build failed, using the x86-64-v3 build: {r.stderr[-300:]}")
        return os.path.join(ROOT, "oracle", "_build", "liboracle.so"), "g++ -O3 -march=x86-64-v3"
    log(f"oracle built with -march=native in {time.perf_counter() - t0:.1f} s")
    return out, "g++ -O3 -march=native"


def cpu_baseline(h, threads, variant, leaf, reps=3, rotation=False):
    """The oracle's construction of the fingerprints h (the CPU baseline), reps times: (last build, median s,
    all times)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    times, ob = [], None
    for _ in range(reps):
        ob = None
        t0 = time.perf_counter()
        ob = O.OracleBuild(variant, leaf, h, 2000, threads=threads, rotation=rotation)
        times.append(time.perf_counter() - t0)
    return ob, float(np.median(times)), times


def serving_requests(eng, args):
    """The serving mix's requests: sizes log-uniform in [1e3, 1e6] (seed 5), their offsets, and one device array
    holding every request's fingerprints (one SplitMix64 stream, seed 2, generated in chunks straight into it)."""
    from shockhash_amd import workload
    from shockhash_amd.engine import DeviceArray
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e6), args.serve_requests)).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    total = int(offs[-1])
    d = DeviceArray(eng, 2 * total, np.uint64)
    chunk = 1 << 24
    for a in range(0, total, chunk):
        b = min(total, a + chunk)
        d.upload_at(2 * a, workload.splitmix64(2, 2 * (b - a), offset=2 * a))
    return sizes, offs, d


def serving_mix(eng, args, mix, rate, alone_digests=None):
    """Serving mix: args.serve_requests build requests, Poisson arrivals at `rate` requests/s,
    sizes log-uniform in [1e3, 1e6], fingerprints resident in HBM, each built as ShockHash2 leaf 64 /
    bucket 2000 by the continuous-batching server. Latency = submission to completion. alone_digests: digests of
    every request built alone (filled on the first call), against which each result in the mix is compared."""
    from shockhash_amd.engine import Server
    sizes, offs, d = mix
    total = int(offs[-1])
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, args.serve_requests))
    log(f"serving mix: {args.serve_requests} requests, {total} keys resident, rate {rate}/s, "
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
    # identical alone and in the mix: every request is built on its own once (digests kept across rates)
    if alone_digests is not None and not alone_digests:
        for i in range(len(sizes)):
            alone = eng.build_dev(d, int(sizes[i]), "sh2", 64, 2000, copy_keys=False, offset=int(offs[i]))
            alone_digests.append(hash((alone.codes.tobytes(), alone.size_acc.tobytes())))
    check = len(sizes) if alone_digests else 0
    same = all(hash((r[0].codes.tobytes(), r[0].size_acc.tobytes())) == a for r, a in zip(results, alone_digests or []))
    srv.close()
    out = {"requests": args.serve_requests, "rate_per_s": rate, "keys": total, "slots": args.serve_slots,
           "batch_keys": args.serve_batch_keys,
           "hbm_budget_bytes": int(args.serve_budget_gib * (1 << 30)), "peak_reserved_bytes": info["peak_reserved_bytes"],
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max()), "queue_p99_ms": float(np.percentile(queue, 99)),
           "keys_per_s": total / wall, "offered_keys_per_s": float(sizes.mean() * rate), "wall_s": wall,
           "identical_alone_checked": check, "identical": same}
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


def strings_workload(args, eng):
    """ShockHash<30,false> class constructor over 1,000,000 distinct random ASCII strings of length 10..50 (seed
    3), bucket 2000: strings packed and uploaded, SpookyV2 fingerprints on the device, sort, plan, search,
    orientation, retrieval, host encoding. One step = one constructor call (its own wall clock)."""
    from shockhash_amd import host, workload
    from shockhash_amd.engine import plan_forest
    from shockhash_amd.roofline import roofline
    strings = workload.random_strings(args.n, seed=3)
    ms, kms, f = [], [], None
    for i in range(args.warmup + args.steps):
        if f is not None:
            f.close()
        f = host.MPHF("sh1", 30, strings, 2000)
        t = f.timings_ms
        if i >= args.warmup:
            ms.append(t["constructor"])
            kms.append(t["search_kernel"])
        log(f"{'warmup' if i < args.warmup else 'step'} {i}: constructor {t['constructor']:.1f} ms, "
            f"search kernel {t['search_kernel']:.2f} ms")
    q = f(strings)
    bij = bool(int(q.max()) < len(strings) and len(np.unique(q)) == len(strings))
    fp = eng.hash_strings(strings)
    nodes = plan_forest("sh1", 30, workload.bucket_sizes(fp, 2000))
    roof = roofline(nodes, f.codes, "sh1", float(np.mean(kms)))
    ms_step = float(np.mean(ms))
    return {"metric": "ShockHash<30> constructor keys/s over 1M strings", "value": args.n / (ms_step / 1e3),
            "unit": "keys/s", "ms_per_step": ms_step, "search_kernel_ms": float(np.mean(kms)), "roofline": roof,
            "queries_bijective": bij, "bits_per_key": f.bits / args.n,
            "config": {"workload": f"ShockHash<30,false> class constructor, bucket 2000, {args.n} random strings of "
                                   "length 10..50 (seed 3), SpookyV2 fingerprints on the device", "keys": args.n,
                       "leaf_size": 30, "bucket_size": 2000, "variant": "ShockHash"}}


def flat_workload(args, eng):
    """ShockHash2Flat<k> (k = --leaf, 64 by default) over n (--n, 10,000,000 by default) uniform 64-bit key hashes (the
    .second of SplitMix64 fingerprints, seed 1), built by the class template from host memory: the threshold layers,
    every key's placement and every bucket's base case (ShockHash2's leaf search, one batched launch), orientation and
    retrieval on the device; thresholds, the fallback function over the bumped keys and their placement on the host.
    One step = one construction (its own wall clock). All n queries must form a bijection onto [0, n) (n a multiple of
    k); on a sample (--cpu-sample, 1,000,000 by default) the function equals the oracle's restatement key for key, and
    the oracle's construction of that sample is the CPU baseline. The roofline is the batched base-case search's."""
    from shockhash_amd import host, workload
    from shockhash_amd.roofline import INT32_PEAK_TOPS, KIND_LEAF, SH2_HALF_SEED_OPS, evaluations
    k, n = args.leaf, args.n
    mh = workload.fingerprints(n, seed=1)[:, 1].copy()
    ms, search_ms, f = [], [], None
    for i in range(args.warmup + args.steps):
        if f is not None:
            f.close()
        f = host.FlatMPHF(k, mh)
        ph = f.phases_ms
        if i >= args.warmup:
            ms.append(f.info["constructor_ms"])
            search_ms.append(ph["search"])
        log(f"{'warmup' if i < args.warmup else 'step'} {i}: constructor {f.info['constructor_ms']:.1f} ms, phases "
            + ", ".join(f"{a} {b:.1f}" for a, b in ph.items()))
    q = f(mh)
    bij = bool(len(np.unique(q)) == n and int(q.max()) == n - 1) if n % k == 0 else bool(len(np.unique(q)) == n)
    codes = f.codes
    nodes = np.zeros(len(codes), dtype=[("m", "<u2"), ("kind", "u1")])
    nodes["m"], nodes["kind"] = k, KIND_LEAF
    evals = float(evaluations(nodes, codes, "sh2").sum())
    s_search = float(np.median(search_ms)) / 1e3
    achieved = evals * SH2_HALF_SEED_OPS / s_search / 1e12
    info = f.info
    f.close()
    ms_step = float(np.median(ms))
    out = {"metric": f"ShockHash2Flat<{k}> constructor keys/s", "value": n / (ms_step / 1e3), "unit": "keys/s",
           "ms_per_step": ms_step, "phases_ms": ph, "base_case_search_ms": 1e3 * s_search,
           "roofline": {"bound": "valu", "achieved": achieved, "peak": INT32_PEAK_TOPS, "unit": "TOP/s",
                        "frac": achieved / INT32_PEAK_TOPS, "traffic": None, "ops_per_eval": SH2_HALF_SEED_OPS,
                        "hash_evals": evals, "evals_per_s": evals / s_search,
                        "kernel": "k_sh2_batch (every bucket's base case, one launch)"},
           "queries_bijective": bij, "bits_per_key": info["bits"] / n, "buckets": info["nbuckets"],
           "config": {"workload": f"ShockHash2Flat<{k}> class constructor over {n} uniform 64-bit key hashes (SplitMix64 "
                                  "fingerprints' .second, seed 1) in host memory", "keys": n, "k": k}}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        lib_path, build_desc = native_oracle()
        O.use_library(lib_path)
        model, share = host_cpu()
        threads = args.cpu_threads or share
        ns = min(n, args.cpu_sample or 1_000_000)
        ms_ = workload.fingerprints(ns, seed=7)[:, 1].copy()
        times, of = [], None
        for _ in range(max(1, args.cpu_reps)):
            of = None
            t0 = time.perf_counter()
            of = O.OracleFlat(k, ms_, threads=threads)
            times.append(time.perf_counter() - t0)
        fs = host.FlatMPHF(k, ms_)
        equal = bool(np.array_equal(fs(ms_), of.query(ms_)))
        fs.close()
        dt = float(np.median(times))
        out["bit_exact_sample"] = equal
        out["cpu_baseline"] = {"value": ns / dt, "unit": "keys/s", "cores": threads, "kind": "port",
                               "sample": f"{ns} key hashes (same generator, seed 7), ShockHash2Flat<{k}>; the oracle's C++ "
                                         f"restatement of the reference construction ({build_desc}), {threads} threads "
                                         f"(this process's CPU share) on {model}; median of {len(times)} runs: "
                                         f"{', '.join(f'{t:.2f}' for t in times)} s; its values equal the GPU "
                                         f"construction's on every key: {equal}",
                               "cpu_model": model}
        log(f"flat CPU baseline: {ns / dt / 1e6:.2f} M keys/s on {ns} keys, values equal {equal}")
    log(f"flat: {ms_step:.1f} ms per construction ({n / ms_step / 1e3:.0f} M keys/s), base-case search {1e3 * s_search:.2f} ms, "
        f"roofline {out['roofline']['frac']:.3f}, bijective {bij}")
    return out


def leaves_workload(args, eng):
    """200,000 isolated ShockHash leaves of 40 keys (the .second of uniform fingerprints, seed 11), each searched
    from seed 0 for its minimal seed (benchmark/bijections/ShockHash1.h:13-35) by the batched leaf search with the
    keys resident in HBM. One step = the whole batch; value = hash evaluations per second."""
    from shockhash_amd import workload
    from shockhash_amd.engine import LEAF_DTYPE, DeviceArray
    from shockhash_amd.roofline import INT32_PEAK_TOPS, SH1_LEAF_OPS
    nl, m = args.n, 40
    keys = workload.isolated_leaves(nl, m, seed=11)
    tasks = np.zeros(nl, dtype=LEAF_DTYPE)
    tasks["offset"], tasks["m"], tasks["start_seed"], tasks["max_trials"] = np.arange(nl) * m, m, 0, 1 << 40
    d_keys = DeviceArray.from_host(eng, keys.reshape(-1))
    d_tasks = DeviceArray.from_host(eng, tasks)
    d_out = DeviceArray(eng, nl, np.uint64)
    step_s = []
    for i in range(args.warmup + args.steps):
        eng.sync()
        t0 = time.perf_counter()
        eng.sh1_leaf_search_dev(d_keys, d_tasks, nl, d_out)
        eng.sync()
        if i >= args.warmup:
            step_s.append(time.perf_counter() - t0)
        log(f"{'warmup' if i < args.warmup else 'step'} {i}: {1e3 * (time.perf_counter() - t0):.1f} ms")
    seeds = d_out.download()
    evals = float(m * (seeds.astype(np.float64) + 1).sum())
    s_step = float(np.mean(step_s))
    gold = os.path.join(ROOT, "tests", "golden", "config2_seeds.npz")
    equal = None
    if nl == 200_000 and os.path.exists(gold):
        equal = bool(np.array_equal(seeds, np.load(gold)["seeds"]))
    achieved = evals * SH1_LEAF_OPS / s_step / 1e12
    return {"metric": "isolated ShockHash leaf hash evaluations/s (200k leaves of 40 keys)", "value": evals / s_step,
            "unit": "evals/s", "ms_per_step": 1e3 * s_step,
            "roofline": {"bound": "valu", "achieved": achieved, "peak": INT32_PEAK_TOPS, "unit": "TOP/s",
                         "frac": achieved / INT32_PEAK_TOPS, "traffic": None, "ops_per_eval": SH1_LEAF_OPS,
                         "hash_evals": evals, "evals_per_s": evals / s_step},
            "seeds_equal_oracle_fixture": equal,
            "config": {"workload": f"{nl} isolated ShockHash leaves of {m} keys (seed 11), seeds from 0, batched leaf "
                                   "search", "leaves": nl, "leaf_size": m}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["build", "strings", "leaves", "flat"], default="build",
                    help="build: the headline persistent build (default); strings: the ShockHash<30> constructor over "
                         "1M strings; leaves: 200k isolated leaves of 40 keys; flat: the ShockHash2Flat<k> constructor "
                         "(k = --leaf) over 10M key hashes")
    ap.add_argument("--n", type=int, default=None, help="keys (build, strings, flat) or leaves (leaves)")
    ap.add_argument("--variant", choices=sorted(VARIANT_NAMES), default="sh2",
                    help="class variant of the build (the headline workload is sh2 with leaf 64)")
    ap.add_argument("--leaf", type=int, default=64, help="leaf size")
    ap.add_argument("--rotation", action="store_true",
                    help="ROTATION_FITTING build (ShockHash / SIMDShockHash: full leaves rotation-fitted)")
    ap.add_argument("--plain-buckets", type=int, default=100, help="plain-form buckets (0: no plain-form run)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="CPU oracle sample size (0: the whole workload)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: this process's CPU share)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="CPU baseline runs (the median is reported)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--split-seeds-per-lane", type=int, default=None)
    ap.add_argument("--leaf-seeds-per-lane", type=int, default=None)
    ap.add_argument("--waves-per-cu", type=int, default=None)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2048)
    ap.add_argument("--serve-rates", type=str, default="800,1300,1800",
                    help="Poisson arrival rates, requests/s (the first one is the headline p99)")
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
    if args.workload != "build":
        args.n = args.n or {"strings": 1_000_000, "leaves": 200_000, "flat": 10_000_000}[args.workload]
        run = {"strings": strings_workload, "leaves": leaves_workload, "flat": flat_workload}[args.workload]
        out = run(args, eng)
        out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic"})
        print(json.dumps(out))
        eng.close()
        return
    args.n = args.n or 10_000_000
    tune = {k: v for k, v in (("split_seeds_per_lane", args.split_seeds_per_lane),
                                ("sh1_seeds_per_lane", args.leaf_seeds_per_lane),
                                ("waves_per_cu", args.waves_per_cu)) if v is not None}
    V, L = args.variant, args.leaf
    if args.rotation and V == "sh2":
        ap.error("--rotation applies to the sh1 and simd variants")
    headline = V == "sh2" and L == 64
    if tune:
        eng.set_variant_tuning(V, **tune)
    vidx = {"sh1": 0, "simd": 1, "sh2": 2}[V]
    sched_class = next((c for c in eng.schedule if c["variant"] == vidx and args.n <= c["max_keys"]),
                       [c for c in eng.schedule if c["variant"] == vidx][-1])
    log(f"device {eng.arch} CUs {eng.cu_count} clock {eng.clock_khz / 1e6:.2f} GHz; "
        f"{VARIANT_NAMES[V]} tuning {eng.variant_tuning(V)}; schedule from {eng.schedule_source}, "
        f"class for {args.n} keys {sched_class}")
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
        b = eng.build_dev(d_h, args.n, V, L, 2000, copy_keys=False, rotation=args.rotation)
        log(f"warmup {i}: timings {np.round(b.timings_ms, 2)}")
    step_s, kernel_ms = [], []
    for i in range(args.steps):
        eng.sync()
        t0 = time.perf_counter()
        b = eng.build_dev(d_h, args.n, V, L, 2000, copy_keys=False, rotation=args.rotation)
        eng.sync()
        step_s.append(time.perf_counter() - t0)
        kernel_ms.append(b.timings_ms[3])
        log(f"step {i}: {step_s[-1] * 1e3:.1f} ms (sort {b.timings_ms[0]:.1f}, tree {b.timings_ms[1]:.1f}, "
            f"search {b.timings_ms[2]:.1f}, kernel {b.timings_ms[3]:.1f})")
    ms_per_step = 1e3 * float(np.mean(step_s))
    value = args.n / (ms_per_step / 1e3)
    kms = float(np.mean(kernel_ms))
    roof = roofline(b.nodes, b.codes, V, kms, MEASURED_TRAFFIC_BYTES if headline and args.n == 10_000_000 else None,
                    rotation_leaf=L if args.rotation else 0)

    # plain form: one launch per node over the first buckets, same device input
    plain_value, n_plain_keys, plain_match = None, 0, True
    if args.plain_buckets:
        t0 = time.perf_counter()
        bp = eng.build_dev(d_h, args.n, V, L, 2000, mode="plain", plain_bucket_limit=args.plain_buckets,
                           copy_keys=False, rotation=args.rotation)
        plain_s = time.perf_counter() - t0
        n_plain_keys = int(b.size_acc[args.plain_buckets])
        nn = len(bp.codes)
        plain_match = bool(np.array_equal(bp.codes, b.codes[:nn]))
        # the plain run also sorts all keys; charge it only for the searched buckets' share of sort time
        plain_search_s = (bp.timings_ms[1] + bp.timings_ms[2]) / 1e3 + bp.timings_ms[0] / 1e3 * n_plain_keys / args.n
        plain_value = n_plain_keys / plain_search_s
        log(f"plain form: {args.plain_buckets} buckets ({n_plain_keys} keys) in {plain_s:.2f} s, codes equal: {plain_match}")

    cpu, ob = None, None
    bit_exact = plain_match
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        lib_path, build_desc = native_oracle()
        O.use_library(lib_path)
        model, share = host_cpu()
        threads = args.cpu_threads or share
        n_cpu = args.cpu_sample or args.n
        if n_cpu == args.n:
            hs, bs = h, b  # the timed GPU build of the same keys
            sample = f"the whole workload ({args.n} keys, seed 1)"
        else:
            hs = workload.fingerprints(n_cpu, seed=7)
            bs = eng.build(hs, V, L, 2000, rotation=args.rotation)
            sample = f"{n_cpu} fingerprints (same generator, seed 7)"
        ob, dt, times = cpu_baseline(hs, threads, V, L, reps=max(1, args.cpu_reps), rotation=args.rotation)
        bit_exact = bit_exact and bool(np.array_equal(bs.codes, ob.node_code)) and \
            bool(np.array_equal(bs.size_acc, ob.size_acc))
        cpu = {"value": n_cpu / dt, "unit": "keys/s", "cores": threads, "kind": "port",
               "sample": f"{sample}, {VARIANT_NAMES[V]} leaf {L}, bucket 2000; the oracle's C++ restatement of the "
                         f"reference construction ({build_desc}), {threads} threads (this process's CPU share) on "
                         f"{model}; median of {len(times)} runs: {', '.join(f'{t:.2f}' for t in times)} s",
               "cpu_model": model}
        if n_cpu != args.n:
            ob = None
        log(f"cpu oracle: {cpu['value']:.0f} keys/s (median of {times}); GPU codes equal the oracle's: {bit_exact}")

    cls = None
    if args.no_class_build or not headline:
        log("class-constructor leg: skipped" + ("" if headline else " (ShockHash2 leaf 64 only)"))
    else:
        cls = class_build(h, args, ob)
        bit_exact = bit_exact and cls["queries_bijective"] and cls["descriptors_equal_oracle"] is not False

    serving = None
    if not args.no_serving and headline:
        digests = []
        mix = serving_requests(eng, args)
        serving = [serving_mix(eng, args, mix, float(r), digests) for r in args.serve_rates.split(",")]
        mix[2].free()
        bit_exact = bit_exact and all(x["identical"] for x in serving)

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
        "config": {"workload": f"{VARIANT_NAMES[V]}{'<ROTATION_FITTING>' if args.rotation else ''} leaf {L}, bucket 2000, "
                               f"{f'{args.n // 1_000_000}M' if args.n % 1_000_000 == 0 else args.n} uniform 128-bit fingerprints "
                               "(SplitMix64, seed 1), persistent device-queue build", "keys": args.n, "leaf_size": L,
                   "bucket_size": 2000, "variant": VARIANT_NAMES[V],
                   "schedule": {"source": os.path.basename(eng.schedule_source),
                                "waves_per_cu": tune.get("waves_per_cu") or sched_class["waves_per_cu"],
                                "tuned_class": None if tune.get("waves_per_cu") else sched_class}},
        "plain_form": {"keys_per_s": plain_value, "buckets": args.plain_buckets, "keys": n_plain_keys,
                       "codes_equal_persistent": plain_match},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "bit_exact": bit_exact,
        "search_kernel_ms": kms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "mphf_build_keys_per_s": cls["keys_per_s"] if cls else None,
        "class_build": cls,
        "p99_build_latency_ms": serving[0]["p99_ms"] if serving else None,
        "serving": serving,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
