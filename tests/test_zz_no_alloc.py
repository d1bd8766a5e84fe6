"""fh_rescale_noalloc (Rescale through the library's fixed per-device array, no allocation) against
fh_rescale and the oracle, bit for bit, on two calls in a row (the second smaller, so it runs over
the first call's leftovers in the shared array), plus its size and overlap contract."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ct(o, seed, l, ncomp):
    return np.concatenate([o.random_poly(seed + k, list(range(l + 1))) for k in range(ncomp)])


def test_rescale_noalloc_matches_rescale(gctx2, octx2, fh):
    g, o = gctx2, octx2
    # N=2^16: 2 components at level 24 use 2 * 25 * 2^16 of the 2^22 scratch words; then a smaller call
    for l, ncomp, seed in [(24, 2, 900), (10, 1, 950)]:
        print(f"[rescale_noalloc N=2^16] level {l}, {ncomp} component(s)", flush=True)
        x = _ct(o, seed, l, ncomp)
        d = g.to_device(x)
        ref, got = g.array((ncomp * l, o.N)), g.array((ncomp * l, o.N))
        g.rescale(d, l, ncomp, ref)
        g.rescale_noalloc(d, l, ncomp, got)
        g.sync()
        r, v = ref.download(), got.download()
        assert np.array_equal(v, r), f"level {l}: fh_rescale_noalloc differs from fh_rescale"
        assert np.array_equal(v, o.rescale_ct(x, l, ncomp)), f"level {l}: differs from the oracle"
        d.free(); ref.free(); got.free()


def test_rescale_noalloc_contract(gctx1, gctx2, octx1, fh):
    # 3 components at level 24, N=2^16: 3 * 25 * 2^16 words > 2^22 -> rejected before any launch
    d = gctx2.array((3 * 25, 1 << 16))
    out = gctx2.array((3 * 24, 1 << 16))
    with pytest.raises(fh.FHError, match="scratch"):
        gctx2.rescale_noalloc(d, 24, 3, out)
    d.free(); out.free()
    # in place is rejected (fh_rescale allows it through a pool buffer)
    o = octx1
    x = _ct(o, 990, 3, 2)
    d1 = gctx1.to_device(x)
    with pytest.raises(fh.FHError, match="overlap"):
        gctx1.rescale_noalloc(d1, 3, 2, d1)
    # the small context works too, after the rejected calls
    out1 = gctx1.array((2 * 3, o.N))
    gctx1.rescale_noalloc(d1, 3, 2, out1)
    assert np.array_equal(out1.download(), o.rescale_ct(x, 3, 2))
    d1.free(); out1.free()
