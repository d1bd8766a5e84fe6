"""Element-wise RNS kernels and the tensor product against the oracle, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op", [0, 1, 2])
def test_binary(op, gctx2, octx2):
    g, o = gctx2, octx2
    nl = 25
    pids = list(range(nl))
    a, b = o.random_poly(1, pids), o.random_poly(2, pids)
    da, db, dc = g.to_device(a), g.to_device(b), g.array(a.shape)
    g.eltwise(op, da, db, dc)
    exp = [o.add, o.sub, o.mul][op](a, b)
    assert np.array_equal(dc.download(), exp)
    for x in (da, db, dc):
        x.free()


@pytest.mark.parametrize("op", [0, 1, 2])
def test_scalar(op, gctx1, octx1):
    g, o = gctx1, octx1
    nl = 4
    a = o.random_poly(3, list(range(nl)))
    s = np.array([int(o.primes[i]) - 5 - i for i in range(nl)], dtype=np.uint64)
    da, dc = g.to_device(a), g.array(a.shape)
    g.scalar(op, da, s if op != 2 else None, dc)
    exp = [lambda: o.add_scalar(a, s), lambda: o.mul_scalar(a, s), lambda: o.neg(a)][op]()
    assert np.array_equal(dc.download(), exp)
    da.free(); dc.free()


@pytest.mark.parametrize("square", [False, True])
def test_tensor(square, gctx2, octx2):
    g, o = gctx2, octx2
    l = 24
    pids = list(range(l + 1))
    c1 = np.concatenate([o.random_poly(10, pids), o.random_poly(11, pids)])
    c2 = c1 if square else np.concatenate([o.random_poly(12, pids), o.random_poly(13, pids)])
    d1 = g.to_device(c1)
    d2 = d1 if square else g.to_device(c2)
    out = g.array((3 * (l + 1), o.N))
    g.tensor(d1, d2, l, out)
    assert np.array_equal(out.download(), o.tensor(c1, c2, l))
