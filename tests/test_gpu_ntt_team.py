"""The single-launch team NTT (tunable ntt_team = 1): forward and inverse transforms, batches with
many limbs per team, and the key-switching / Rescale compositions that use its fused epilogues,
bit-exact against the oracle at the production shape (N=2^16) and at N=2^12 / 2^14 (team sizes 1
and 4). After each case fh_synchronize confirms that no team barrier wait ran past its bound."""
import numpy as np
import pytest

from conftest import make_oracle_ctx

pytestmark = pytest.mark.gpu


class team_on:
    """ntt_team = 1 (and teams per XCD) on a context for the duration of a block."""

    def __init__(self, g, per_xcd=None):
        self.g, self.per_xcd = g, per_xcd

    def __enter__(self):
        self.saved = (self.g.tune_get("ntt_team"), self.g.tune_get("ntt_team_per_xcd"))
        self.g.tune_set("ntt_team", 1)
        if self.per_xcd:
            self.g.tune_set("ntt_team_per_xcd", self.per_xcd)
        return self.g

    def __exit__(self, *exc):
        self.g.tune_set("ntt_team", self.saved[0])
        self.g.tune_set("ntt_team_per_xcd", self.saved[1])


def test_team_ntt_all_primes_n65536(gctx2, octx2):
    g, o = gctx2, octx2
    pids = list(range(o.L1 + o.K))          # 25 fp64-path primes and the 60-bit integer-path ones
    x = o.random_poly(11, pids)
    with team_on(g):
        print("[team ntt N=2^16] forward + inverse over all 34 primes", flush=True)
        d = g.to_device(x)
        g.ntt_forward(d, pids)
        g.sync()
        assert np.array_equal(d.download(), o.ntt(x, pids)), "team forward NTT differs from the oracle"
        g.ntt_inverse(d, pids)
        g.sync()
        assert np.array_equal(d.download(), x), "team INTT(NTT(x)) != x"
        y = o.random_poly(12, pids)
        d.upload(y)
        g.ntt_inverse(d, pids)
        g.sync()
        assert np.array_equal(d.download(), o.intt(y, pids))
        d.free()


@pytest.mark.parametrize("per_xcd", [1, 4, 8])
def test_team_ntt_many_limbs_per_team(gctx2, octx2, per_xcd):
    """32 polynomials of 25 limbs with 2 gap limbs between them (800 limbs): every team transforms many
    limbs in turn; the gap limbs between the polynomials stay untouched."""
    g, o = gctx2, octx2
    N, npoly, nl, stride = o.N, 32, 25, 27
    pids = list(range(nl))
    host = np.zeros((npoly * stride, N), dtype=np.uint64)
    for p in range(npoly):
        host[p * stride:p * stride + nl] = o.random_poly(300 + p, pids)
    exp = [o.ntt(host[p * stride:p * stride + nl], pids) for p in range(npoly)]
    with team_on(g, per_xcd):
        print(f"[team ntt N=2^16] {npoly * nl} limbs, {per_xcd} teams per XCD", flush=True)
        d = g.to_device(host)
        g.ntt_forward(d, pids, num_polys=npoly, poly_stride=stride * N)
        g.sync()
        got = d.download()
        for p in range(npoly):
            assert np.array_equal(got[p * stride:p * stride + nl], exp[p]), p
            assert not got[p * stride + nl:(p + 1) * stride].any(), "gap limbs were written"
        g.ntt_inverse(d, pids, num_polys=npoly, poly_stride=stride * N)
        g.sync()
        assert np.array_equal(d.download(), host)
        d.free()


def test_team_keyswitch_and_rescale_n65536(gctx2, octx2):
    """ModUp (INTT + target NTTs), ModDown (INTT + NTT with the ModDown epilogue), the rotation built
    on them and Rescale (NTT with the Rescale epilogue) through the team launches."""
    g, o = gctx2, octx2
    l = 24
    ids = list(range(l + 1))
    ct = np.concatenate([o.random_poly(400, ids), o.random_poly(401, ids)])
    key = o.random_poly(402, list(range(o.L1 + o.K)) * (2 * o.dnum))
    k = o.rot_index(5)
    with team_on(g):
        print("[team ntt N=2^16] modup / rotation / rescale at level 24", flush=True)
        d = g.to_device(ct)
        dig = g.array((o.num_digits(l) * (l + 1 + o.K), o.N))
        c1 = g.to_device(ct[l + 1:])
        g.modup(c1, l, dig)
        g.sync()
        assert np.array_equal(dig.download(), o.modup(ct[l + 1:], l)), "ModUp"
        dk = g.key(key)
        out = g.array((2 * (l + 1), o.N))
        g.eval_rotate(d, l, dk, k, out)
        g.sync()
        assert np.array_equal(out.download(), o.eval_rotate(ct, l, key, k)), "EvalRotate"
        rs = g.array((2 * l, o.N))
        g.rescale(d, l, 2, rs)
        g.sync()
        assert np.array_equal(rs.download(), o.rescale_ct(ct, l, 2)), "Rescale"
        for x in (d, c1, dig, dk, out, rs):
            x.free()


@pytest.mark.parametrize("logN", [12, 14])
def test_team_ntt_other_ring_sizes(fh, oracle_mod, logN):
    """N = 2^12 (teams of 1 workgroup) and 2^14 (teams of 4)."""
    o = make_oracle_ctx(oracle_mod, dict(logN=logN, L1=3, K=1, dnum=1))
    g = fh.Context(logN, 3, 1, 1)
    try:
        g.tune_set("ntt_team", 1)
        pids = [0, 1, 2, 3] * 8
        x = o.random_poly(logN + 40, pids)
        d = g.to_device(x)
        g.ntt_forward(d, pids)
        g.sync()
        assert np.array_equal(d.download(), o.ntt(x, pids))
        g.ntt_inverse(d, pids)
        g.sync()
        assert np.array_equal(d.download(), x)
        d.free()
    finally:
        g.close()
