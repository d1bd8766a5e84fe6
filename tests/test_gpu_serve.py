"""The continuous-batching serving loop (fh_serve_run): every request completes, the fused
mixed-level batches produce exactly the results of one-call-per-operation execution (checksum of
every request's result), and latency statistics are consistent."""
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


def test_serving_rejects_bad_config(sctx, fh):
    rk, rg, relin = _keys(sctx, 91)
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, start_level=5)      # too few levels for a request
    with pytest.raises(fh.FHError):
        sctx.serve(rk, rg, relin, 100.0, 4, mode=3)
    for k in rk + [relin]:
        k.free()
