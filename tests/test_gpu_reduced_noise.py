"""ReducedNoise basis conversions (fh_context_set_reduced_noise) against the oracle's ReducedNoise
restatement: ModUp, ModDown, EvalRotate, EvalMult and a mixed-level batch, bit-exact; switching
the option off restores the plain conversions."""
import numpy as np
import pytest

from conftest import CFG1, CFG2, make_oracle_ctx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["cfg1", "cfg2"])
def pair(request, oracle_mod, fh):
    cfg = CFG1 if request.param == "cfg1" else CFG2
    o = make_oracle_ctx(oracle_mod, cfg)
    o.set_reduced_noise(True)
    g = fh.Context(cfg["logN"], cfg["L1"], cfg["K"], cfg["dnum"])
    g.set_reduced_noise(True)
    yield o, g
    g.close()


def test_rn_modup_moddown(pair):
    o, g = pair
    for l in sorted({o.L1 - 1, o.L1 // 2, 1}):
        print(f"[rn] N={o.N} level={l}", flush=True)
        x = o.random_poly(300 + l, list(range(l + 1)))
        d = g.to_device(x)
        beta = o.num_digits(l)
        dig = g.array((beta * (l + 1 + o.K), o.N))
        g.modup(d, l, dig)
        assert np.array_equal(dig.download(), o.modup(x, l))
        xe = np.concatenate([o.random_poly(400 + l + k, o.ext_pids(l)) for k in range(2)])
        de = g.to_device(xe)
        out = g.array((2 * (l + 1), o.N))
        g.moddown(de, l, 2, out)
        exp = np.concatenate([o.moddown(xe[k * (l + 1 + o.K):(k + 1) * (l + 1 + o.K)], l) for k in range(2)])
        assert np.array_equal(out.download(), exp)
        for a in (d, dig, de, out):
            a.free()


def test_rn_rotate_mult_and_batch(pair, fh):
    o, g = pair
    l = o.L1 - 1
    ct = np.concatenate([o.random_poly(500 + c, list(range(l + 1))) for c in range(2)])
    ct2 = np.concatenate([o.random_poly(510 + c, list(range(l + 1))) for c in range(2)])
    key = o.random_poly(520, list(range(o.L1 + o.K)) * (2 * o.dnum))
    k = o.rot_index(1)
    dk = g.key(key)
    d1, d2 = g.to_device(ct), g.to_device(ct2)
    out = g.array(ct.shape)
    g.eval_rotate(d1, l, dk, k, out)
    assert np.array_equal(out.download(), o.eval_rotate(ct, l, key, k))
    g.eval_mult(d1, d2, l, dk, out)
    assert np.array_equal(out.download(), o.eval_mult(ct, ct2, l, key))
    # mixed-level batch with the option on
    b = fh.Batch(g)
    lv = max(1, l // 2)
    c3 = np.concatenate([o.random_poly(530 + c, list(range(lv + 1))) for c in range(2)])
    d3 = g.to_device(c3)
    o1, o2 = g.array(ct.shape), g.array((2 * lv, o.N))
    b.rotate(d1.ptr, l, dk, k, o1.ptr)
    b.mult(d3.ptr, d3.ptr, lv, dk, True, o2.ptr)
    b.finalize()
    b.run()
    g.sync()
    assert np.array_equal(o1.download(), o.eval_rotate(ct, l, key, k))
    assert np.array_equal(o2.download(), o.rescale_ct(o.eval_mult(c3, c3, lv, key), lv))
    b.free()
    # switched off: the plain conversions again (and they differ from ReducedNoise)
    g.set_reduced_noise(False)
    o.set_reduced_noise(False)
    g.eval_rotate(d1, l, dk, k, out)
    plain = o.eval_rotate(ct, l, key, k)
    assert np.array_equal(out.download(), plain)
    o.set_reduced_noise(True)
    assert not np.array_equal(plain, o.eval_rotate(ct, l, key, k))
    g.set_reduced_noise(True)
    for a in (dk, d1, d2, d3, out, o1, o2):
        a.free()
