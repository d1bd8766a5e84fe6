"""The continuous-batching serving loop (fh_serve_run): every request completes, the fused
mixed-level batches produce exactly the results of one-call-per-operation execution (checksum of
every request's result), and latency statistics are consistent."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sctx(fh):
    # N=2^12, 12 Q primes (start level 11: a request consumes 8 levels), dnum 3
    c = fh.Context(12, 12, 2, 3)
    yield c
    c.close()


def _keys(ctx, seed):
    full = list(range(ctx.L1 + ctx.K))
    buf = ctx.array((ctx.dnum * 2 * (ctx.L1 + ctx.K), ctx.N))
    keys = []
    for i in range(9):
        ctx.fill_uniform(buf, full, num_polys=ctx.dnum * 2, seed=seed + i)
        keys.append(ctx.key_from_device(buf))
    buf.free()
    return keys[:8], [ctx.galois_index(1 << k) for k in range(8)], keys[8]


@pytest.mark.parametrize("rate,requests", [(2000.0, 24), (200.0, 16)])
def test_serving_batched_equals_per_op(sctx, rate, requests):
    rk, rg, relin = _keys(sctx, 77)
    print(f"[serve] rate={rate} requests={requests}", flush=True)
    b = sctx.serve(rk, rg, relin, rate, requests, mode=0, seed=5, checksum=True)
    p = sctx.serve(rk, rg, relin, rate, requests, mode=1, seed=5, checksum=True)
    for r in (b, p):
        assert r["completed"] == requests
        assert r["inference"] + r["polynomial"] == requests
        assert 0 < r["p50_ms"] <= r["p99_ms"] <= r["max_ms"]
        assert r["ticks"] >= 17
    assert b["checksum"] == p["checksum"], "fused batches differ from per-operation execution"
    for k in rk + [relin]:
        k.free()


def test_each_result_same_alone_and_in_the_mix(sctx):
    """Every request's result is bit for bit the same in the batched mix, in per-operation
    execution and when it runs alone (mode 2: one request admitted at a time)."""
    rk, rg, relin = _keys(sctx, 83)
    runs = {mode: sctx.serve(rk, rg, relin, 3000.0, 12, mode=mode, seed=9, request_checksums=True, request_outputs=True)
            for mode in (0, 1, 2)}
    # the host waiting for every tick, and running far ahead of the GPU
    runs["sync"] = sctx.serve(rk, rg, relin, 3000.0, 12, mode=0, seed=9, request_checksums=True, request_outputs=True, ticks_in_flight=1)
    runs["ahead"] = sctx.serve(rk, rg, relin, 3000.0, 12, mode=0, seed=9, request_checksums=True, request_outputs=True, ticks_in_flight=16)
    assert runs[2]["max_active"] == 1 and runs[0]["max_active"] >= 3
    per = {mode: r["request_checksums"] for mode, r in runs.items()}
    assert all(v != 0 for v in per[2])
    outs = {mode: r["request_outputs"] for mode, r in runs.items()}
    for i in range(12):
        assert outs[2][i].size > 0
        for mode in (0, 1, "sync", "ahead"):
            assert np.array_equal(outs[mode][i], outs[2][i]), f"request {i}: result in mode {mode} differs from the request alone"
    assert all(per[mode] == per[2] for mode in per)
    for r in runs.values():
        assert r["p99_other_ms"] == r["p99_ms"]  # no bootstrap requests in this mix
    for k in rk + [relin]:
        k.free()


def test_serving_rejects_bad_config(sctx, fh):
    rk, rg, relin = _keys(sctx, 91)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, start_level=5)      # too few levels for a request
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, mode=3)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, mode=-1)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, bootstrap_share=0.2)  # no bootstrap function
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, inference_share=0.9, bootstrap_share=0.2)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, ticks_in_flight=65)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, bootstrap_concurrent=3)
    for k in rk + [relin]:
        k.free()


# ---- bootstrap requests through the C++ facade (tests/cpp/serve_check.cpp) ----
import json
import os
import subprocess

from conftest import PKG

SCHECK = os.path.join(PKG, "_build", "serve_check")


@pytest.mark.parametrize("share,concurrent", [(0.3, 0), (1.0, 0), (0.3, 1), (1.0, 1), (0.3, 2)])
def test_serving_with_bootstrap_requests(share, concurrent):
    """A session whose arrivals include bootstrap requests, served by the facade's EvalBootstrap
    (N=2^12, 8 slots, level budget {1,1}, 16 towers): batched and per-operation execution give the
    same checksum, every bootstrap request went through the callback once per mode. concurrent:
    the bootstraps run on the loop's worker thread through a second context with the same keys
    (2: yielding, one bootstrap operation at a time between the loop's ticks)."""
    if not os.path.exists(SCHECK):
        pytest.fail("serve_check is not built (run __graft_entry__.build())")
    args = ["12", "8", "1", "1", "15", "50", "2000", "16", str(share), "5", str(0.5 * (1 - share)), "16", "0", str(concurrent)]
    print(f"[serve_check] {SCHECK} {' '.join(args)}", flush=True)
    r = subprocess.run([SCHECK] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    b, p = res["batched"], res["per_op"]
    for x in (b, p):
        assert x["completed"] == 16
        assert x["inference"] + x["polynomial"] + x["bootstraps"] == 16
        assert 0 < x["p50_ms"] <= x["p99_ms"] <= x["max_ms"]
    assert b["bootstraps"] == p["bootstraps"] > 0
    if share == 1.0:
        assert b["bootstraps"] == 16
    assert res["callbacks"] == 4 * b["bootstraps"]  # warm-up, batched, per-operation and alone sessions
    assert b["checksum"] == p["checksum"], "fused batches differ from per-operation execution"
    assert res["each_result_same_alone_and_in_mix"]
    assert res["bitwise_same_alone_and_in_mix"], "a request's result differs between the mix and the request alone"
    assert res["concurrent"] == concurrent


@pytest.mark.parametrize("technique,concurrent", [(0, 0), (1, 0), (0, 1), (0, 2)],
                         ids=["fixedmanual", "flexibleauto", "fixedmanual-concurrent", "fixedmanual-yielding"])
def test_serving_workload_n65536_with_bootstraps(technique, concurrent):
    """The serving workload's own shape: N=2^16, 25 towers, real keys (sparse ternary h=32), all
    32768 slots, bootstrap level budget {3,3} with 59-bit scaling; arrivals at 50/s, 70% inference,
    25% polynomial and 5% bootstrap requests (the bootstrap requests under the context's scaling
    technique). Every request's result is bit for bit the same in the batched mix, in
    per-operation execution and alone; concurrent: with the bootstraps on the loop's worker thread
    (a second context, same keys) while the other requests keep ticking (yielding: one bootstrap
    operation at a time, between ticks)."""
    if not os.path.exists(SCHECK):
        pytest.fail("serve_check is not built (run __graft_entry__.build())")
    args = ["16", "32768", "3", "3", "24", "59", "50", "40", "0.05", "2026", "0.70", "32", str(technique), str(concurrent)]
    print(f"[serve_check] {SCHECK} {' '.join(args)}", flush=True)
    r = subprocess.run([SCHECK] + args, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print({k: res[k] for k in ("technique", "concurrent", "callbacks", "each_result_same_alone_and_in_mix", "bitwise_same_alone_and_in_mix")})
    print({m: {k: res[m][k] for k in ("p50_ms", "p99_ms", "p99_other_ms")} for m in ("batched", "per_op", "alone")})
    b, p, a = res["batched"], res["per_op"], res["alone"]
    for x in (b, p, a):
        assert x["completed"] == 40
        assert x["inference"] + x["polynomial"] + x["bootstraps"] == 40
    assert b["bootstraps"] == p["bootstraps"] == a["bootstraps"] > 0, "the seeded mix holds bootstrap requests"
    assert a["max_active"] == 1 and b["max_active"] >= 2
    assert res["bitwise_same_alone_and_in_mix"], "a request's result differs between the mix and the request alone"
