"""Device frees (fh_device_free) do not wait for the stream: a freed block is handed out again only
after the work queued before the free has completed. Sources freed right after the rotations that
read them are queued, then overwritten through new allocations of the same size, must leave every
rotation's result equal to the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_freed_source_reused_only_after_its_readers(gctx1, octx1):
    g, o = gctx1, octx1
    s = o.gen_secret(51)
    l = 3
    gal = o.rot_index(1)
    rk_h = o.gen_rot_key(s, gal, 52)
    rk = g.key(rk_h)
    ids = list(range(l + 1))
    srcs = [np.concatenate([o.random_poly(700 + 2 * i, ids), o.random_poly(701 + 2 * i, ids)]) for i in range(8)]
    outs, junk = [], []
    for i, src in enumerate(srcs):
        d = g.to_device(src)
        mine = [g.array(src.shape) for _ in range(4)]
        for out in mine:
            g.eval_rotate(d, l, rk, gal, out)  # queued; each reads d
        d.free()  # returns at once
        # same size class: d's block only once the rotations above have completed
        y = g.array(src.shape)
        g.fill_uniform(y, ids, num_polys=2, seed=900 + i)
        outs.append(mine)
        junk.append(y)
    for i, src in enumerate(srcs):
        want = o.eval_rotate(src, l, rk_h, gal)
        for out in outs[i]:
            assert np.array_equal(out.download(), want), f"rotation of source {i} saw its freed source overwritten"
    for x in [a for m in outs for a in m] + junk + [rk]:
        x.free()
