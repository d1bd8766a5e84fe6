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
