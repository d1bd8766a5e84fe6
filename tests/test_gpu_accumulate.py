"""Rotation-fold sums against the oracle, bit for bit: the lazy fold (fh_accumulate_sum, the reference's
Accumulate(ct, bStep, stride, size)) and the cascade from a start factor (fh_accumulate_sum_from,
Accumulate(..., startFactor)) at power-of-radix and ragged sizes, strides, radices 2/4/8, the no-op
cases and the rejected calls; at N=2^16 the 256-summand radix-4 fold (four levels)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ct(o, seed, l):
    ids = list(range(l + 1))
    return np.concatenate([o.random_poly(seed, ids), o.random_poly(seed + 1, ids)])


@pytest.fixture(scope="module")
def secret1(octx1):
    return octx1.gen_secret(71)


@pytest.mark.parametrize("radix,stride,size,start,l", [(4, 1, 16, 1, 3), (2, 3, 8, 1, 2), (4, 2, 9, 1, 3), (4, 1, 16, 2, 1),
                                                        (8, 1, 64, 1, 3), (4, 5, 3, 1, 0), (2, 1, 32, 1, 3), (4, 1, 64, 4, 2)])
def test_accumulate_matches_oracle(radix, stride, size, start, l, gctx1, octx1, fh, secret1):
    g, o = gctx1, octx1
    print(f"[accumulate] radix={radix} stride={stride} size={size} start={start} l={l}", flush=True)
    idx = g.accumulate_indices(radix, stride, size, start)
    assert idx == o.accumulate_indices(radix, stride, size, start)
    ks = [o.rot_index(r) for r in idx]
    hkeys = [o.gen_rot_key(secret1, k, 500 + i) for i, k in enumerate(ks)]
    x = _ct(o, 700 + size, l)
    keys = [g.key(k) for k in hkeys]
    d = g.to_device(x)
    out = g.array(x.shape)
    # the cascade from the start factor
    exp = o.accumulate_sum_from(x, l, radix, stride, size, hkeys, ks, start)
    g.accumulate_sum_from(d, l, radix, stride, size, start, keys, ks, out)
    assert np.array_equal(out.download(), exp)
    if start == 1:
        # the lazy fold, out of place and in place
        exp = o.accumulate_sum(x, l, radix, stride, size, hkeys, ks)
        g.accumulate_sum(d, l, radix, stride, size, keys, ks, out)
        assert np.array_equal(out.download(), exp)
        g.accumulate_sum(d, l, radix, stride, size, keys, ks, d)
        assert np.array_equal(d.download(), exp)
    else:
        g.accumulate_sum_from(d, l, radix, stride, size, start, keys, ks, d)      # in place
        assert np.array_equal(d.download(), exp)
    for k in keys + [d, out]:
        k.free()


def test_accumulate_full_size(gctx2, octx2):
    """N = 2^16 at level 24: the 256-summand lazy fold at radix 4 (four levels, 12 rotations)."""
    g, o = gctx2, octx2
    l, radix, size = 24, 4, 256
    s = o.gen_secret(81)
    idx = o.accumulate_indices(radix, 1, size)
    assert idx == [1, 2, 3, 4, 8, 12, 16, 32, 48, 64, 128, 192]
    ks = [o.rot_index(r) for r in idx]
    print("[accumulate] N=2^16 l=24 256 summands: keys", flush=True)
    hkeys = [o.gen_rot_key(s, k, 800 + i) for i, k in enumerate(ks)]
    x = _ct(o, 900, l)
    keys = [g.key(k) for k in hkeys]
    d = g.to_device(x)
    out = g.array(x.shape)
    print("[accumulate] N=2^16 l=24 256 summands: GPU and oracle", flush=True)
    g.accumulate_sum(d, l, radix, 1, size, keys, ks, out)
    assert np.array_equal(out.download(), o.accumulate_sum(x, l, radix, 1, size, hkeys, ks))
    for k in keys + [d, out]:
        k.free()


def test_accumulate_no_op_and_errors(gctx1, octx1, fh, secret1):
    g, o = gctx1, octx1
    l = 2
    x = _ct(o, 950, l)
    d = g.to_device(x)
    out = g.array(x.shape)
    for size, start in ((1, 1), (0, 1), (8, 8), (8, 0)):
        assert g.accumulate_indices(4, 1, size, start) == []
        g.accumulate_sum_from(d, l, 4, 1, size, start, [], [], out)
        assert np.array_equal(out.download(), x)
        if start == 1:
            g.accumulate_sum(d, l, 4, 1, size, [], [], out)
            assert np.array_equal(out.download(), x)
    k1 = o.rot_index(1)
    key = g.key(o.gen_rot_key(secret1, k1, 960))
    with pytest.raises(fh.FHError):
        g.accumulate_sum(d, l, 3, 1, 4, [key], [k1], out)                  # radix not a power of two
    with pytest.raises(fh.FHError):
        g.accumulate_sum(d, l, 4, 1, 4, [key], [k1], out)                  # three rotations, one key
    with pytest.raises(fh.FHError):
        g.accumulate_sum(d, l, 2, 1, 2, [key], [2], out)                   # even Galois element
    with pytest.raises(fh.FHError):
        g.accumulate_indices(4, 0, 8)                                      # stride below 1
    for z in (d, out, key):
        z.free()
