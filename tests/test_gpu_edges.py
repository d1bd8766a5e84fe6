"""Edge cases of the C ABI: the lowest level (one limb), empty batches and transforms, and the
contract violations each entry point rejects with an error code (never a launch)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def level0(octx1):
    o = octx1
    s = o.gen_secret(41)
    return dict(s=s, relin=o.gen_relin_key(s, 42), rot=o.gen_rot_key(s, o.rot_index(1), 43))


def _ct(o, seed, l):
    ids = list(range(l + 1))
    return np.concatenate([o.random_poly(seed, ids), o.random_poly(seed + 1, ids)])


def test_level0_mult_and_rotate(gctx1, octx1, level0):
    """Key switching at level 0 (one Q limb, one digit): products and rotations equal the oracle."""
    g, o, m = gctx1, octx1, level0
    a, b = _ct(o, 600, 0), _ct(o, 602, 0)
    da, db = g.to_device(a), g.to_device(b)
    out = g.array(a.shape)
    k = g.key(m["relin"])
    g.eval_mult(da, db, 0, k, out)
    assert np.array_equal(out.download(), o.eval_mult(a, b, 0, m["relin"]))
    rk = g.key(m["rot"])
    g.eval_rotate(da, 0, rk, o.rot_index(1), out)
    assert np.array_equal(out.download(), o.eval_rotate(a, 0, m["rot"], o.rot_index(1)))
    for x in (da, db, out, k, rk):
        x.free()


def test_empty_work_is_a_no_op(gctx1, octx1, level0):
    g, o, m = gctx1, octx1, level0
    d = g.to_device(_ct(o, 610, 2))
    before = d.download()
    g.ntt_forward(d, [], num_limbs=0)
    rk = g.key(m["rot"])
    g.rotate_batch(d, 0, 2, [], [], d)
    assert np.array_equal(d.download(), before)
    d.free(); rk.free()


def test_contract_violations(gctx1, octx1, level0, fh):
    g, o, m = gctx1, octx1, level0
    l = 2
    d = g.to_device(_ct(o, 620, l))
    out = g.array((2 * (l + 1), o.N))
    k = g.key(m["relin"])
    rk = g.key(m["rot"])
    with pytest.raises(fh.FHError):
        g.rescale(d, 0, 2, out)                                   # no limb to drop at level 0
    with pytest.raises(fh.FHError):
        g.rescale(d, l, 9, out)                                   # more than 8 components
    with pytest.raises(fh.FHError):
        g.rescale(d, o.L1, 2, out)                                # level beyond the chain
    with pytest.raises(fh.FHError):
        g.eval_rotate(d, l, rk, 2, out)                           # even Galois element
    with pytest.raises(fh.FHError):
        g.eval_mult_many(0, d, d, l, k, out)                      # empty batch of products
    with pytest.raises(fh.FHError):
        g.eval_mult_many(1, d, d, l, k, out, towers1=l)           # operand below the product's level
    with pytest.raises(fh.FHError):
        g.mod_raise(d, l, 2, o.L1, out)                           # raise beyond the chain
    with pytest.raises(fh.FHError):
        g.linear_wsum([d], [l + 1], np.ones((1, o.L1 + 1), dtype=np.uint64), o.L1 + 1, 2, out)
    with pytest.raises(fh.FHError):
        g.encode_coeffs(np.zeros((0, o.N), dtype=np.int64), [0], out)
    for x in (d, out, k, rk):
        x.free()


def test_even_galois_rejected(gctx1, octx1, level0, fh):
    """Every entry point taking Galois elements rejects an even one (not a bijection of the slots,
    so the key inner product's scatter would leave outputs unwritten) before any launch, and so does
    one at or above 2N."""
    g, o, m = gctx1, octx1, level0
    l = 2
    src = _ct(o, 630, l)
    d = g.to_device(src)
    out = g.array((2 * (l + 1), o.N))
    rk = g.key(m["rot"])
    good = o.rot_index(1)
    for bad in (2, 2 * o.N + 1):
        with pytest.raises(fh.FHError, match="galois"):
            g.rotate_batch(d, 1, l, [rk], [bad], out)
        with pytest.raises(fh.FHError, match="galois"):
            g.rotate_hoisted(d, l, [rk, rk], [good, bad], out)
        with pytest.raises(fh.FHError, match="galois"):
            g.linear_transform(d, l, o.N // 2, 2, 1, d, [None, rk], [1, bad], out)
        with pytest.raises(fh.FHError, match="galois"):
            g.linear_transform(d, l, o.N // 2, 2, 1, d, [None, rk], [1, good], out, giant_key=rk, giant_galois=bad)
        with pytest.raises(fh.FHError, match="galois"):
            g.linear_transform_batch(d, 1, l, o.N // 2, 2, 1, d, [None, rk], [1, bad], out)
        with pytest.raises(fh.FHError, match="galois"):
            g.linear_transform_flat(d, l, 4, 2, 1, d, [None, rk], [1, good], out, giant_keys=[rk], giant_galois=[bad])
    # the flat form needs a key for every giant multiple that rotates
    with pytest.raises(fh.FHError, match="null key"):
        g.linear_transform_flat(d, l, 4, 2, 1, d, [None, rk], [1, good], out, giant_keys=[None], giant_galois=[1])
    with pytest.raises(fh.FHError, match="giant"):
        g.linear_transform_flat(d, l, 4, 2, 1, d, [None, rk], [1, good], out)
    # nothing was launched: the output buffer is untouched and the engine still works
    g.rotate_batch(d, 1, l, [rk], [good], out)
    assert np.array_equal(out.download(), o.eval_rotate(src, l, m["rot"], good))
    for x in (d, out, rk):
        x.free()


def test_stream_copy(gctx1, fh):
    """The STREAM copy kernel (the bandwidth reference of the counter calibration) copies exactly:
    a size that is no multiple of its 16 KiB chunk, one chunk, a single 16-byte word; and it rejects
    sizes and pointers that are not 16-byte aligned."""
    g = gctx1
    lib = fh.load()
    rng = np.random.default_rng(5)
    for words in (2 * (3 * 2048 + 17), 2048, 2):
        src = g.array((words,))
        dst = g.array((words + 2,))
        host = rng.integers(0, 2**63, size=words, dtype=np.uint64)
        src.upload(host)
        dst.upload(np.full(words + 2, 7, dtype=np.uint64))
        fh._check(lib.fh_stream_copy(g.h, dst.ptr, src.ptr, words * 8))
        got = dst.download()
        assert np.array_equal(got[:words], host), words
        assert np.all(got[words:] == 7), "the copy wrote past its end"
        src.free()
        dst.free()
    a = g.array((4,))
    with pytest.raises(fh.FHError, match="16-byte"):
        fh._check(lib.fh_stream_copy(g.h, a.ptr, a.ptr + 8, 16))
    with pytest.raises(fh.FHError, match="16-byte"):
        fh._check(lib.fh_stream_copy(g.h, a.ptr, a.ptr, 8))
    a.free()
