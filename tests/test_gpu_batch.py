"""Mixed-level operation batches (fh_batch_*) against the oracle.

A batch holds rotations, multiplications (with and without Rescale) and additions of
ciphertexts at different levels; every item must equal the oracle's single operation bit for bit,
eagerly and on every replay of the captured HIP graph (with the inputs changed between replays,
and eager work and synchronizations in between).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# N=2^12, 8 Q primes, dnum 3 (alpha 3): levels 1..7 use 1, 2 or 3 digits
CFG = dict(logN=12, L1=8, dnum=3)


# ntt_mix 0: prime-major NTT job tables; 1: jobs of 60-bit primes spread among the 50-bit ones
@pytest.fixture(scope="module", params=[0, 1], ids=["prime_major", "mixed_classes"])
def env(request, oracle_mod, fh):
    O = oracle_mod
    K = 2
    primes = O.gen_primes(CFG["logN"], CFG["L1"], 60, 50, K)
    o = O.Ctx(CFG["logN"], primes, CFG["L1"], K, CFG["dnum"])
    g = fh.Context(CFG["logN"], CFG["L1"], K, CFG["dnum"])
    assert [int(x) for x in g.moduli] == [int(x) for x in primes]
    g.tune_set("ntt_mix", request.param)
    yield o, g
    g.close()


def _rand_ct(o, seed, l, ncomp=2):
    return np.concatenate([o.random_poly(seed * 7 + c, list(range(l + 1))) for c in range(ncomp)])


def _key(o, seed):
    return o.random_poly(seed, list(range(o.L1 + o.K)) * (2 * o.dnum))


KINDS = ["rot", "mult", "mult_rs", "add2", "add3", "mulpt", "mulpt_rs", "addpt", "drop"]


def _items(o, rng, n):
    """(kind, level, extra): extra picks the rotation key (rot) or the target level offset (drop)"""
    out = []
    for i in range(n):
        kind = KINDS[rng.integers(0, len(KINDS))]
        l = int(rng.integers(1 if kind in ("mult_rs", "mulpt_rs") else 0, o.L1))
        out.append((kind, l, int(rng.integers(0, 3))))
    return out


def _ncomp(kind):
    return 3 if kind == "add3" else 2


def _drop_to(l, extra):
    return max(0, l - extra)


def _expected(o, kind, l, a, b, keys, gal, extra):
    t = l + 1
    comp = lambda x, c: x[c * t:(c + 1) * t]
    if kind == "rot":
        return o.eval_rotate(a, l, keys[0], gal)
    if kind == "mult":
        return o.eval_mult(a, b, l, keys[1])
    if kind == "mult_rs":
        return o.rescale_ct(o.eval_mult(a, b, l, keys[1]), l)
    if kind in ("mulpt", "mulpt_rs"):
        pt = comp(b, 0)
        prod = np.concatenate([o.mul(comp(a, 0), pt), o.mul(comp(a, 1), pt)])
        return o.rescale_ct(prod, l) if kind == "mulpt_rs" else prod
    if kind == "addpt":
        return np.concatenate([o.add(comp(a, 0), comp(b, 0)), comp(a, 1)])
    if kind == "drop":
        lo = _drop_to(l, extra) + 1
        return np.concatenate([comp(a, 0)[:lo], comp(a, 1)[:lo]])
    return np.concatenate([o.add(comp(a, c), comp(b, c)) for c in range(_ncomp(kind))])


def _out_limbs(kind, l, extra):
    t = l + 1
    return {"rot": 2 * t, "mult": 2 * t, "mult_rs": 2 * l, "add2": 2 * t, "add3": 3 * t, "mulpt": 2 * t, "mulpt_rs": 2 * l,
            "addpt": 2 * t, "drop": 2 * (_drop_to(l, extra) + 1)}[kind]


@pytest.mark.parametrize("n_items,seed", [(12, 1), (40, 2), (90, 3)])
def test_batch_matches_oracle(env, fh, n_items, seed):
    o, g = env
    rng = np.random.default_rng(seed)
    rot_keys = [_key(o, 100 + i) for i in range(3)]
    relin = _key(o, 200)
    drk = [g.key(k) for k in rot_keys]
    drl = g.key(relin)
    steps = [1, -2, 5]
    gals = [o.rot_index(r) for r in steps]
    items = _items(o, rng, n_items)
    print(f"[batch] n_items={n_items} seed={seed} kinds={[k for k, _, _ in items]}", flush=True)
    b = fh.Batch(g)
    host, dev = [], []
    for i, (kind, l, ki) in enumerate(items):
        nc = _ncomp(kind)
        a = _rand_ct(o, 1000 * seed + 2 * i, l, nc)
        bb = _rand_ct(o, 1000 * seed + 2 * i + 1, l, nc)
        da, db = g.to_device(a), g.to_device(bb)
        dout = g.array((_out_limbs(kind, l, ki), o.N))
        if kind == "rot":
            b.rotate(da.ptr, l, drk[ki], gals[ki], dout.ptr)
        elif kind in ("mult", "mult_rs"):
            b.mult(da.ptr, db.ptr, l, drl, kind == "mult_rs", dout.ptr)
        elif kind in ("mulpt", "mulpt_rs"):
            b.mult_plain(da.ptr, db.ptr, l, kind == "mulpt_rs", dout.ptr)   # first component of b as the plaintext
        elif kind == "addpt":
            b.add_plain(da.ptr, db.ptr, l, dout.ptr)
        elif kind == "drop":
            b.level_reduce(da.ptr, l, _drop_to(l, ki), 2, dout.ptr)
        else:
            b.add(da.ptr, db.ptr, l, nc, dout.ptr)
        host.append((kind, l, ki, a, bb))
        dev.append((da, db, dout))
    b.finalize()
    n, scratch = b.info()
    assert n == n_items and scratch > 0

    def check(tag):
        for i, (kind, l, ki, a, bb) in enumerate(host):
            want = _expected(o, kind, l, a, bb, [rot_keys[ki], relin], gals[ki], ki)
            got = dev[i][2].download()
            assert np.array_equal(got, want), f"{tag}: item {i} ({kind}, level {l}) differs from the oracle"

    b.run(graph=False)
    g.sync()
    check("eager")
    # graph: capture on the first run, replay with fresh inputs each time
    for rep in range(3):
        for i, (kind, l, ki, a, bb) in enumerate(host):
            nc = _ncomp(kind)
            a2 = _rand_ct(o, 50000 + 1000 * rep + 2 * i, l, nc)
            b2 = _rand_ct(o, 50000 + 1000 * rep + 2 * i + 1, l, nc)
            dev[i][0].upload(a2)
            dev[i][1].upload(b2)
            host[i] = (kind, l, ki, a2, b2)
        g.sync()
        b.run(graph=True)
        g.sync()
        check(f"graph replay {rep}")
        # eager single-operation work between replays
        k0, l0, ki0, a0, b0 = host[0]
        if k0 == "rot":
            tmp = g.array((2 * (l0 + 1), o.N))
            g.eval_rotate(dev[0][0], l0, drk[ki0], gals[ki0], tmp)
            g.sync()
            assert np.array_equal(tmp.download(), dev[0][2].download())
            tmp.free()
    b.free()
    for da, db, dout in dev:
        da.free()
        db.free()
        dout.free()
    for k in drk + [drl]:
        k.free()


def test_batch_reset_reuse(env, fh):
    """A batch reset and refilled (the serving loop's pattern) keeps producing oracle results."""
    o, g = env
    relin = _key(o, 300)
    drl = g.key(relin)
    b = fh.Batch(g)
    for rnd in range(3):
        n = 5 + 7 * rnd
        rng = np.random.default_rng(400 + rnd)
        bufs = []
        for i in range(n):
            l = int(rng.integers(1, o.L1))
            x, y = _rand_ct(o, 9000 + 100 * rnd + 2 * i, l), _rand_ct(o, 9001 + 100 * rnd + 2 * i, l)
            dx, dy, do = g.to_device(x), g.to_device(y), g.array((2 * l, o.N))
            b.mult(dx.ptr, dy.ptr, l, drl, True, do.ptr)
            bufs.append((l, x, y, dx, dy, do))
        b.finalize()
        b.run(graph=(rnd == 1))
        g.sync()
        for l, x, y, dx, dy, do in bufs:
            assert np.array_equal(do.download(), o.rescale_ct(o.eval_mult(x, y, l, relin), l)), f"round {rnd}"
            for d in (dx, dy, do):
                d.free()
        b.reset()
    b.free()
    drl.free()


def test_batch_mult_operands_above_level(env, fh):
    """fh_batch_mult_strided: products (with and without Rescale) whose operands sit above the
    product's level, read in place, equal the oracle's product of the operands' first level + 1
    limbs; operand towers outside [level + 1, L1] are rejected."""
    o, g = env
    relin = _key(o, 310)
    drl = g.key(relin)
    b = fh.Batch(g)
    rng = np.random.default_rng(77)
    bufs = []
    for i in range(9):
        l = int(rng.integers(1, o.L1 - 1))
        ta, tb = int(rng.integers(l + 1, o.L1 + 1)), int(rng.integers(l + 1, o.L1 + 1))
        x, y = _rand_ct(o, 9500 + 2 * i, ta - 1), _rand_ct(o, 9501 + 2 * i, tb - 1)
        rs = bool(i % 2)
        dx, dy, do = g.to_device(x), g.to_device(y), g.array((2 * (l if rs else l + 1), o.N))
        b.mult_strided(dx.ptr, ta, dy.ptr, tb, l, drl, rs, do.ptr)
        bufs.append((l, ta, tb, rs, x, y, dx, dy, do))
    with pytest.raises(fh.FHError):
        b.mult_strided(bufs[0][6].ptr, bufs[0][0], bufs[0][7].ptr, bufs[0][2], bufs[0][0], drl, False, bufs[0][8].ptr)  # towers_a = level
    with pytest.raises(fh.FHError):
        b.mult_strided(bufs[0][6].ptr, o.L1 + 1, bufs[0][7].ptr, bufs[0][2], bufs[0][0], drl, False, bufs[0][8].ptr)   # above L1
    b.finalize()
    b.run(graph=False)
    g.sync()
    for l, ta, tb, rs, x, y, dx, dy, do in bufs:
        lo = l + 1
        xa = np.concatenate([x[:lo], x[ta:ta + lo]])
        yb = np.concatenate([y[:lo], y[tb:tb + lo]])
        exp = o.eval_mult(xa, yb, l, relin)
        if rs:
            exp = o.rescale_ct(exp, l)
        assert np.array_equal(do.download(), exp), (l, ta, tb, rs)
        for d in (dx, dy, do):
            d.free()
    b.free()
    drl.free()


def test_batch_errors(env, fh):
    o, g = env
    b = fh.Batch(g)
    key = g.key(_key(o, 7))
    x = g.array((4, o.N))
    with pytest.raises(fh.FHError):
        b.mult(x.ptr, x.ptr, 0, key, True, x.ptr)   # rescale at level 0
    with pytest.raises(fh.FHError):
        b.rotate(x.ptr, 1, key, 4, x.ptr)           # even galois index
    with pytest.raises(fh.FHError):
        b.add(x.ptr, x.ptr, o.L1, 2, x.ptr)         # level out of range
    b.finalize()
    with pytest.raises(fh.FHError):
        b.add(x.ptr, x.ptr, 0, 2, x.ptr)            # finalized
    b.run()                                         # an empty batch is a no-op
    g.sync()
    b.free()
    x.free()
    key.free()
