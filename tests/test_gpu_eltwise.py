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


@pytest.mark.parametrize("n,towers,with_add", [(1, 5, False), (7, 12, True), (16, 10, True), (3, 25, False), (14, 23, True)])
def test_linear_wsum(n, towers, with_add, gctx2, octx2):
    """fh_linear_wsum against the oracle's scalar products and additions: inputs at mixed tower
    counts (only the first `towers` limbs of each component count), 2 components; up to 320
    weights travel in the kernel arguments, more (14 x 23) through an uploaded table."""
    g, o = gctx2, octx2
    rng = np.random.default_rng(n)
    cts, devs, tin = [], [], []
    for k in range(n):
        t = towers + int(rng.integers(0, 3)) if towers + 2 <= o.L1 else towers
        pids = list(range(t))
        c = np.concatenate([o.random_poly(100 + 2 * k, pids), o.random_poly(101 + 2 * k, pids)])
        cts.append(c); tin.append(t); devs.append(g.to_device(c))
    w = np.array([[int(rng.integers(0, int(o.primes[i]))) for i in range(towers)] for _ in range(n)], dtype=np.uint64)
    add = np.array([int(rng.integers(0, int(o.primes[i]))) for i in range(towers)], dtype=np.uint64) if with_add else None
    out = g.array((2 * towers, o.N))
    g.linear_wsum(devs, tin, w, towers, 2, out, add)
    exp = np.zeros((2 * towers, o.N), dtype=np.uint64)
    for k in range(n):
        for comp in range(2):
            part = cts[k][comp * tin[k]: comp * tin[k] + towers]
            exp[comp * towers:(comp + 1) * towers] = o.add(exp[comp * towers:(comp + 1) * towers], o.mul_scalar(part, w[k]))
    if with_add:
        exp[:towers] = o.add_scalar(exp[:towers], add)
    assert np.array_equal(out.download(), exp)
    for x in devs + [out]:
        x.free()


def test_linear_wsum_pair_constant(gctx2, octx2):
    """Four components (two ciphertexts back to back), the constant on components 0 and 2."""
    g, o = gctx2, octx2
    towers = 6
    pids = list(range(towers))
    x = np.concatenate([o.random_poly(140 + j, pids) for j in range(4)])
    w = np.array([[3 + i for i in range(towers)]], dtype=np.uint64)
    add = np.array([int(o.primes[i]) - 7 for i in range(towers)], dtype=np.uint64)
    d, out = g.to_device(x), g.array((4 * towers, o.N))
    g.linear_wsum([d], [towers], w, towers, 4, out, add, add_components=5)
    exp = np.concatenate([o.mul_scalar(x[j * towers:(j + 1) * towers], w[0]) for j in range(4)])
    for j in (0, 2):
        exp[j * towers:(j + 1) * towers] = o.add_scalar(exp[j * towers:(j + 1) * towers], add)
    assert np.array_equal(out.download(), exp)
    d.free(); out.free()


def test_linear_wsum_rejects_bad_shapes(gctx2, fh):
    g = gctx2
    a = g.array((2 * 4, g.N))
    out = g.array((2 * 5, g.N))
    with pytest.raises(fh.FHError):
        g.linear_wsum([a], [4], np.ones((1, 5), dtype=np.uint64), 5, 2, out)    # input below the output's towers
    with pytest.raises(fh.FHError):
        g.linear_wsum([a] * 17, [4] * 17, np.ones((17, 4), dtype=np.uint64), 4, 2, out)
    a.free(); out.free()


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
