"""Serving loop (shgpu_server_*): continuous batching of construction requests on one GPU. Every
request's result must be identical to the same keys built alone, and to the CPU oracle."""
import numpy as np
import pytest

import oracle_lib as O
from shockhash_amd import workload
from shockhash_amd.engine import ERR_ADMISSION, DeviceArray, Server, ShgpuError

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def request_mix(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    return [workload.fingerprints(int(n), seed=1000 + seed * 100 + i) for i, n in enumerate(sizes)]


def same(a, b):
    return np.array_equal(a.size_acc, b.size_acc) and np.array_equal(a.codes, b.codes) and \
        np.array_equal(a.nodes, b.nodes)


def describe(a, b, oracle=None):
    """The first differences between two builds (and which of them the oracle agrees with)."""
    out = []
    if not np.array_equal(a.size_acc, b.size_acc):
        out.append("bucket sizes differ")
    if len(a.nodes) != len(b.nodes) or not np.array_equal(a.nodes, b.nodes):
        out.append(f"forests differ ({len(a.nodes)} vs {len(b.nodes)} nodes)")
    elif not np.array_equal(a.codes, b.codes):
        for i in np.nonzero(a.codes != b.codes)[0][:6]:
            nd = a.nodes[i]
            o = f", oracle {oracle.node_code[i]}" if oracle is not None else ""
            out.append(f"node {i} (kind {nd['kind']}, m {nd['m']}, unit {nd['unit']}): {a.codes[i]} vs {b.codes[i]}{o}")
    return "; ".join(out)


@pytest.mark.parametrize("variant,leaf,bucket,slots", [("sh2", 64, 2000, 4), ("sh2", 64, 2000, 1), ("sh2", 16, 500, 3),
                                                       ("sh1", 30, 1000, 2), ("simd", 24, 1000, 4)])
def test_results_identical_alone_and_in_the_mix(engine, variant, leaf, bucket, slots):
    log(f"serving {variant} L{leaf} b{bucket} slots {slots}")
    mix = request_mix(12, 50, 120_000, seed=leaf)
    mix.append(workload.fingerprints(0, seed=1))  # empty request
    mix.append(workload.fingerprints(1, seed=2))
    srv = Server(engine, variant, leaf, bucket, slots=slots, hbm_budget_bytes=4 << 30, copy_keys=True)
    # every other request from fingerprints already resident on the device
    resident = [DeviceArray.from_host(engine, h.reshape(-1)) if i % 2 and len(h) else None for i, h in enumerate(mix)]
    try:
        ids = [srv.submit_device(r, len(h)) if r is not None else srv.submit(h) for h, r in zip(mix, resident)]
        got = [srv.take(i) for i in ids]
        info = srv.info()
    finally:
        srv.close()
    log(f"  served {len(ids)} requests, peak reserved {info['peak_reserved_bytes'] >> 20} MiB")
    assert info["queued"] == 0 and info["running"] == 0
    assert info["peak_reserved_bytes"] <= 4 << 30
    for h, (b, lat, q) in zip(mix, got):
        alone = engine.build(h, variant, leaf, bucket)
        if not same(b, alone):
            ob = O.OracleBuild(variant, leaf, h, bucket, threads=8)
            pytest.fail(f"{len(h)} keys, in the mix vs alone: {describe(b, alone, ob)}")
        assert np.array_equal(b.keys(len(h)), alone.keys(len(h)))
        assert lat >= q >= 0
    # the smallest requests also against the oracle
    for h, (b, _, _) in zip(mix, got):
        if 0 < len(h) <= 20_000:
            ob = O.OracleBuild(variant, leaf, h, bucket, threads=8)
            assert np.array_equal(b.codes, ob.node_code) and np.array_equal(b.size_acc, ob.size_acc)
    log("  identical alone, in the mix and to the oracle")


def test_requests_batched_behind_a_long_one(engine):
    """One slot: while a large request runs, the small ones queue up and are then built together in one
    kernel; each equals its build alone."""
    log("serving batch")
    big = workload.fingerprints(1_000_000, seed=21)
    small = request_mix(16, 100, 30_000, seed=22)
    srv = Server(engine, "sh2", 64, 2000, slots=1, max_batch_keys=1 << 20, max_batch_requests=8)
    try:
        ids = [srv.submit(big)] + [srv.submit(h) for h in small]
        got = [srv.take(i)[0] for i in ids]
    finally:
        srv.close()
    for h, b in zip([big] + small, got):
        assert same(b, engine.build(h, "sh2", 64, 2000))


def test_budget_serialises_and_rejects(engine):
    """A budget that holds one request at a time still completes every request, bit-exact; a request
    larger than the budget is refused at submission."""
    log("serving budget")
    mix = request_mix(6, 20_000, 60_000, seed=3)
    with pytest.raises(ShgpuError) as ei:
        Server(engine, "sh2", 64, 2000, slots=4, hbm_budget_bytes=16 << 20)  # below the fixed leaf scratch
    assert ei.value.status == ERR_ADMISSION
    probe = Server(engine, "sh2", 64, 2000, slots=4, hbm_budget_bytes=64 << 30)
    fixed = probe.info()["reserved_bytes"]
    one = probe.request_bytes(max(len(h) for h in mix))
    probe.close()
    budget = fixed + one + one // 2  # room for one running request at a time
    srv = Server(engine, "sh2", 64, 2000, slots=4, hbm_budget_bytes=budget)
    try:
        with pytest.raises(ShgpuError) as ei:
            srv.submit(workload.fingerprints(2_000_000, seed=4))
        assert ei.value.status == ERR_ADMISSION
        ids = [srv.submit(h) for h in mix]
        got = [srv.take(i)[0] for i in ids]
        info = srv.info()
    finally:
        srv.close()
    assert info["peak_reserved_bytes"] <= budget and info["reserved_bytes"] == fixed
    for h, b in zip(mix, got):
        assert same(b, engine.build(h, "sh2", 64, 2000))


def test_untaken_requests_are_released(engine):
    srv = Server(engine, "sh2", 32, 1000, slots=2)
    for i in range(4):
        srv.submit(workload.fingerprints(5000, seed=50 + i))
    srv.close()  # waits for running requests, drops queued ones, frees results
