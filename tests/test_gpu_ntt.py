"""NTT / INTT parity: the HIP kernels against the CPU oracle, bit-exact, through the C ABI."""
import numpy as np
import pytest

from conftest import CFG1, CFG2, make_oracle_ctx

pytestmark = pytest.mark.gpu


def test_context_moduli_match_oracle(gctx1, gctx2, octx1, octx2):
    for g, o in ((gctx1, octx1), (gctx2, octx2)):
        assert [int(x) for x in g.moduli] == [int(x) for x in o.primes]
        assert [int(x) for x in g.psi] == o.psi


@pytest.mark.parametrize("cfgname", ["cfg1", "cfg2"])
def test_ntt_forward_inverse_all_primes(cfgname, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    pids = list(range(o.L1 + o.K))
    x = o.random_poly(7, pids)
    d = g.to_device(x)
    g.ntt_forward(d, pids)
    got = d.download()
    exp = o.ntt(x, pids)
    assert np.array_equal(got, exp), "forward NTT differs from the oracle"
    g.ntt_inverse(d, pids)
    back = d.download()
    assert np.array_equal(back, x), "INTT(NTT(x)) != x"
    # inverse on independent input
    y = o.random_poly(8, pids)
    d.upload(y)
    g.ntt_inverse(d, pids)
    assert np.array_equal(d.download(), o.intt(y, pids))
    d.free()


def test_ntt_batched_strided(gctx2, octx2):
    """Three polynomials of 5 limbs each, 7 limbs apart (stride > limbs), primes out of order."""
    g, o = gctx2, octx2
    N = o.N
    pids = [24, 0, 30, 3, 25]
    npoly, stride_limbs = 3, 7
    host = np.zeros((npoly * stride_limbs, N), dtype=np.uint64)
    for p in range(npoly):
        host[p * stride_limbs:p * stride_limbs + 5] = o.random_poly(100 + p, pids)
    d = g.to_device(host)
    g.ntt_forward(d, pids, num_polys=npoly, poly_stride=stride_limbs * N)
    got = d.download()
    for p in range(npoly):
        blk = slice(p * stride_limbs, p * stride_limbs + 5)
        assert np.array_equal(got[blk], o.ntt(host[blk], pids))
        gap = slice(p * stride_limbs + 5, (p + 1) * stride_limbs)
        assert not got[gap].any(), "gap limbs between strided polynomials were written"
    d.free()


def test_ntt_edge_values(gctx1, octx1):
    g, o = gctx1, octx1
    pids = list(range(o.L1 + o.K))
    x = np.zeros((len(pids), o.N), dtype=np.uint64)
    for i, p in enumerate(pids):
        x[i, ::2] = o.primes[p] - 1  # maximal residues
    d = g.to_device(x)
    g.ntt_forward(d, pids)
    assert np.array_equal(d.download(), o.ntt(x, pids))
    d.upload(np.zeros_like(x))
    g.ntt_forward(d, pids)
    assert not d.download().any()
    d.free()


@pytest.mark.parametrize("logN", [10, 11, 12, 14, 15])
def test_ntt_all_ring_sizes(fh, oracle_mod, logN):
    cfg = dict(logN=logN, L1=3, K=1, dnum=1)
    o = make_oracle_ctx(oracle_mod, cfg)
    g = fh.Context(logN, 3, 1, 1)
    try:
        pids = [0, 1, 2, 3]
        x = o.random_poly(logN, pids)
        d = g.to_device(x)
        g.ntt_forward(d, pids)
        assert np.array_equal(d.download(), o.ntt(x, pids))
        g.ntt_inverse(d, pids)
        assert np.array_equal(d.download(), x)
        d.free()
    finally:
        g.close()


# radix splits (log2 of the column-pass length) the kernels are built for, by logN
SPLITS = {10: [4, 5, 6], 11: [5, 6], 12: [5, 6, 7], 13: [6, 7], 14: [6, 7, 8], 15: [7, 8], 16: [8]}
THREAD_KNOBS = ["ntt_fwd_cols_threads", "ntt_fwd_rows_threads", "ntt_inv_rows_threads", "ntt_inv_cols_threads"]


@pytest.mark.parametrize("logN,l1", [(n, s) for n, ss in SPLITS.items() for s in ss])
def test_ntt_radix_splits_and_tiles(fh, oracle_mod, logN, l1):
    """Every radix split gives the oracle's transform; at the default split so do the 512-thread
    tiles of each pass."""
    cfg = dict(logN=logN, L1=3, K=1, dnum=1)
    o = make_oracle_ctx(oracle_mod, cfg)
    g = fh.Context(logN, 3, 1, 1)
    try:
        pids = [0, 1, 2, 3]
        x = o.random_poly(100 + logN + l1, pids)
        want = o.ntt(x, pids)
        g.tune_set("ntt_split", l1)
        shapes = [[256] * 4]
        if l1 == logN // 2:
            shapes += [[512] * 4, [512, 256, 256, 512], [256, 512, 512, 256]]
        for sh in shapes:
            for k, v in zip(THREAD_KNOBS, sh):
                g.tune_set(k, v)
            print(f"[ntt split] logN={logN} l1={l1} threads={sh}", flush=True)
            d = g.to_device(x)
            g.ntt_forward(d, pids)
            assert np.array_equal(d.download(), want), (l1, sh)
            g.ntt_inverse(d, pids)
            assert np.array_equal(d.download(), x), (l1, sh)
            d.free()
    finally:
        g.close()


@pytest.mark.parametrize("streams", [1, 2])
def test_ntt_wide_tiles_chunked_strided(streams, gctx2, octx2):
    """512-thread tiles with 2-limb launch chunks over strided polynomials at N=2^16, the chunks on
    one stream or alternating between two (ntt_streams)."""
    g, o = gctx2, octx2
    N = o.N
    pids = [24, 0, 30, 3, 25]
    npoly, stride_limbs = 3, 7
    host = np.zeros((npoly * stride_limbs, N), dtype=np.uint64)
    for p in range(npoly):
        host[p * stride_limbs:p * stride_limbs + 5] = o.random_poly(200 + p, pids)
    saved = {k: g.tune_get(k) for k in THREAD_KNOBS + ["ntt_chunk_mb", "ntt_streams"]}
    try:
        for k in THREAD_KNOBS:
            g.tune_set(k, 512)
        g.tune_set("ntt_chunk_mb", 1)
        g.tune_set("ntt_streams", streams)
        d = g.to_device(host)
        g.ntt_forward(d, pids, num_polys=npoly, poly_stride=stride_limbs * N)
        got = d.download()
        for p in range(npoly):
            blk = slice(p * stride_limbs, p * stride_limbs + 5)
            assert np.array_equal(got[blk], o.ntt(host[blk], pids))
        g.ntt_inverse(d, pids, num_polys=npoly, poly_stride=stride_limbs * N)
        assert np.array_equal(d.download(), host)
        d.free()
    finally:
        for k, v in saved.items():
            g.tune_set(k, v)


def test_tunable_validation(gctx1, fh):
    g = gctx1
    with pytest.raises(fh.FHError):
        g.tune_set("ntt_split", 3)            # below the smallest built split
    with pytest.raises(fh.FHError):
        g.tune_set("ntt_split", 9)
    with pytest.raises(fh.FHError):
        g.tune_set("ntt_fwd_cols_threads", 384)
    with pytest.raises(fh.FHError):
        g.tune_set("ntt_mix", 2)
    with pytest.raises(fh.FHError):
        g.tune_set("ntt_streams", 3)
    with pytest.raises(fh.FHError):
        g.tune_set("rot_streams", 0)
    with pytest.raises(fh.FHError):
        g.tune_set("no_such_knob", 1)
    assert g.tune_get("ntt_split") == 0


@pytest.mark.parametrize("logN,polys", [(12, 16), (16, 4)])
def test_autotune_choice_is_valid_and_exact(fh, oracle_mod, logN, polys):
    """The autotuner keeps a built shape no slower than the default, and transforms stay exact."""
    cfg = dict(logN=logN, L1=3, K=1, dnum=1)
    o = make_oracle_ctx(oracle_mod, cfg)
    g = fh.Context(logN, 3, 1, 1)
    try:
        r = g.autotune(polys)
        print(f"[autotune] logN={logN} {r}", flush=True)
        assert r["split"] in SPLITS[logN]
        assert set(r["threads"].values()) <= {256, 512}
        assert r["candidates"] >= 3 and r["limbs"] == polys * 4
        assert 0 < r["ms_best"] <= r["ms_default"]
        assert g.tune_get("ntt_split") in (0, r["split"])
        pids = [0, 1, 2, 3]
        x = o.random_poly(300 + logN, pids)
        d = g.to_device(x)
        g.ntt_forward(d, pids)
        assert np.array_equal(d.download(), o.ntt(x, pids))
        g.ntt_inverse(d, pids)
        assert np.array_equal(d.download(), x)
        d.free()
    finally:
        g.close()


@pytest.mark.skip(reason="CU-masked context streams are switched off: the process crashed (SIGSEGV) inside "
                         "the CU-masked stream's creation on the GPU host, also after the mask-size argument "
                         "was corrected; the stream_cu_share tunable was withdrawn")
def test_ntt_on_cu_limited_stream(fh, octx1):
    """A context whose stream is restricted to a share of the CUs (stream_cu_share, a CU-masked
    stream) computes the same transforms; the share is validated and can be set back to all CUs."""
    from conftest import CFG1
    o = octx1
    g = fh.Context(CFG1["logN"], CFG1["L1"], CFG1["K"], CFG1["dnum"])
    try:
        pids = list(range(o.L1 + o.K))
        x = o.random_poly(17, pids)
        for share in (50, 13, 100):
            g.tune_set("stream_cu_share", share)
            assert g.tune_get("stream_cu_share") == share
            d = g.to_device(x)
            g.ntt_forward(d, pids)
            assert np.array_equal(d.download(), o.ntt(x, pids)), f"forward NTT differs at a {share}% CU share"
            d.free()
        for bad in (0, 101):
            with pytest.raises(fh.FHError):
                g.tune_set("stream_cu_share", bad)
    finally:
        g.close()
