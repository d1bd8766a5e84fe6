"""The headline paths at their own shapes, against the oracle bit for bit.

* fh_rotate_batch at N=2^16, 25 limbs (level 24), dnum 3, K 9: 20 ciphertexts rotated by steps
  drawn (seeded) from the 16 keyed steps of the batched rotation workload; every limb of every
  output compared with the oracle's EvalRotate (reference: CudaEngine::evalRotateMany /
  Ciphertext::rotate_many, api/engine/cuda/CudaEngine.cpp:656-702, src/CKKS/Ciphertext.cpp:1519-1599),
  with the key inner product separate and fused into the targets' row stages (NTT_KSK_DOT).
* the mixed-level batch (fh_batch) at N=2^16 with levels drawn uniformly from [1, 24]: 40%
  EvalRotate, 40% EvalMult + relinearization + Rescale, 20% EvalAdd, run eagerly and replayed from
  the captured HIP graph with fresh inputs (plain conversions, not ReducedNoise).

Keys are uniform residues of the key shape: the hot-path arithmetic does not depend on how a key was
sampled, and the bit-exact comparison only needs the same words on both sides.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS16 = [1 << i for i in range(15)] + [3]


def _key(o, seed):
    return o.random_poly(seed, list(range(o.L1 + o.K)) * (2 * o.dnum))


def _ct(o, seed, l, ncomp=2):
    return np.concatenate([o.random_poly(seed * 7 + c, list(range(l + 1))) for c in range(ncomp)])


@pytest.mark.parametrize("fused", [0, 1])
def test_rotate_batch_headline_shape(fused, gctx2, octx2):
    """fused 1: the ModUp targets' row stages run fused with the key inner product (tunable
    rot_fused), the same words expected."""
    g, o = gctx2, octx2
    g.tune_set("rot_fused", fused)
    l, B = 24, 20
    rng = np.random.default_rng(2026)
    pick = [STEPS16[i] for i in rng.integers(0, len(STEPS16), size=B)]
    used = sorted(set(pick))
    print(f"[rotate_batch N=2^16 l=24 B={B}] steps {pick}", flush=True)
    hkeys = {r: _key(o, 3000 + r) for r in used}
    dk = {r: g.key(hkeys[r]) for r in used}
    cts = [_ct(o, 4000 + b, l) for b in range(B)]
    d = g.to_device(np.concatenate(cts))
    rows = 2 * (l + 1)
    out = g.array((B * rows, o.N))
    try:
        g.rotate_batch(d, B, l, [dk[r] for r in pick], [o.rot_index(r) for r in pick], out)
        g.sync()
    finally:
        g.tune_set("rot_fused", 0)
    got = out.download().reshape(B, rows, o.N)
    for b in range(B):
        exp = o.eval_rotate(cts[b], l, hkeys[pick[b]], o.rot_index(pick[b]))
        bad = int(np.count_nonzero(got[b] != exp))
        assert bad == 0, f"ciphertext {b} (step {pick[b]}): {bad} words differ from the oracle"
    d.free()
    out.free()
    for k in dk.values():
        k.free()


def test_mixed_level_batch_headline_shape(gctx2, octx2, fh):
    g, o = gctx2, octx2
    n = 16
    rng = np.random.default_rng(77)
    kinds = rng.choice(3, size=n, p=[0.4, 0.4, 0.2])          # 0 rotate, 1 mult + relin + rescale, 2 add
    levels = [int(v) for v in rng.integers(1, o.L1, size=n)]
    steps = [STEPS16[i] for i in rng.integers(0, len(STEPS16), size=n)]
    print(f"[batch N=2^16] kinds {list(kinds)} levels {levels}", flush=True)
    relin = _key(o, 5000)
    drl = g.key(relin)
    hrot = {r: _key(o, 5100 + r) for r in set(steps[i] for i in range(n) if kinds[i] == 0)}
    drot = {r: g.key(k) for r, k in hrot.items()}
    b = fh.Batch(g)
    host, dev = [], []
    for i in range(n):
        l = levels[i]
        a, c = _ct(o, 6000 + 2 * i, l), _ct(o, 6001 + 2 * i, l)
        da, dc = g.to_device(a), g.to_device(c)
        dout = g.array((2 * l if kinds[i] == 1 else 2 * (l + 1), o.N))
        if kinds[i] == 0:
            b.rotate(da.ptr, l, drot[steps[i]], o.rot_index(steps[i]), dout.ptr)
        elif kinds[i] == 1:
            b.mult(da.ptr, dc.ptr, l, drl, True, dout.ptr)
        else:
            b.add(da.ptr, dc.ptr, l, 2, dout.ptr)
        host.append([a, c])
        dev.append((da, dc, dout))
    b.finalize()

    def expected(i):
        l, (a, c) = levels[i], host[i]
        if kinds[i] == 0:
            return o.eval_rotate(a, l, hrot[steps[i]], o.rot_index(steps[i]))
        if kinds[i] == 1:
            return o.rescale_ct(o.eval_mult(a, c, l, relin), l)
        return np.concatenate([o.add(a[:l + 1], c[:l + 1]), o.add(a[l + 1:], c[l + 1:])])

    def check(tag):
        for i in range(n):
            got = dev[i][2].download()
            assert np.array_equal(got, expected(i)), f"{tag}: item {i} (kind {kinds[i]}, level {levels[i]}) differs"

    b.run(graph=False)
    g.sync()
    check("eager")
    for rep in range(2):
        for i in range(n):
            host[i] = [_ct(o, 70000 + 1000 * rep + 2 * i, levels[i]), _ct(o, 70001 + 1000 * rep + 2 * i, levels[i])]
            dev[i][0].upload(host[i][0])
            dev[i][1].upload(host[i][1])
        g.sync()
        print(f"[batch N=2^16] graph replay {rep}", flush=True)
        b.run(graph=True)
        g.sync()
        check(f"graph replay {rep}")
    b.free()
    for t in dev:
        for x in t:
            x.free()
    for k in list(drot.values()) + [drl]:
        k.free()
