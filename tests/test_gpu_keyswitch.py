"""Rescale, ModUp, key inner product, ModDown, automorphism, EvalMult and EvalRotate on the GPU
against the CPU oracle, bit-exact on every limb (plus decrypted-slot checks for the full ops)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ids(o, l):
    return list(range(l + 1))


@pytest.mark.parametrize("cfgname,level,ncomp", [("cfg1", 3, 2), ("cfg1", 1, 3), ("cfg2", 24, 2), ("cfg2", 9, 2)])
def test_rescale(cfgname, level, ncomp, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    x = np.concatenate([o.random_poly(40 + k, _ids(o, level)) for k in range(ncomp)])
    d = g.to_device(x)
    out = g.array((ncomp * level, o.N))
    g.rescale(d, level, ncomp, out)
    assert np.array_equal(out.download(), o.rescale_ct(x, level, ncomp))
    d.free(); out.free()


@pytest.mark.parametrize("cfgname,level", [("cfg1", 3), ("cfg1", 2), ("cfg1", 0), ("cfg2", 24), ("cfg2", 17), ("cfg2", 8), ("cfg2", 3)])
def test_modup(cfgname, level, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    x = o.random_poly(50 + level, _ids(o, level))
    beta = o.num_digits(level)
    assert g.num_digits(level) == beta
    d = g.to_device(x)
    dig = g.array((beta * (level + 1 + o.K), o.N))
    g.modup(d, level, dig)
    assert np.array_equal(dig.download(), o.modup(x, level))
    d.free(); dig.free()


@pytest.mark.parametrize("cfgname,level", [("cfg1", 3), ("cfg2", 24), ("cfg2", 17), ("cfg2", 3)])
def test_modup_mixed_job_order(cfgname, level, gctx1, gctx2, octx1, octx2):
    """ModUp with the target NTT jobs of 60-bit primes spread among the 50-bit ones (ntt_mix=1)."""
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    x = o.random_poly(60 + level, _ids(o, level))
    beta = o.num_digits(level)
    d = g.to_device(x)
    dig = g.array((beta * (level + 1 + o.K), o.N))
    g.tune_set("ntt_mix", 1)
    try:
        g.modup(d, level, dig)
        g.sync()
    finally:
        g.tune_set("ntt_mix", 0)
    assert np.array_equal(dig.download(), o.modup(x, level))
    d.free(); dig.free()


@pytest.mark.parametrize("cfgname,level,galois,with_c0", [("cfg1", 3, 1, False), ("cfg1", 2, 5, True), ("cfg2", 24, 25, True), ("cfg2", 13, 1, False)])
def test_key_inner_product(cfgname, level, galois, with_c0, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    ext = level + 1 + o.K
    beta = o.num_digits(level)
    pids = o.ext_pids(level)
    digits = np.concatenate([o.random_poly(60 + d, pids) for d in range(beta)])
    full = list(range(o.L1 + o.K))
    key = np.concatenate([o.random_poly(70 + i, full) for i in range(2 * o.dnum)])
    c0 = o.random_poly(80, _ids(o, level))
    dk = g.key(key)
    dd = g.to_device(digits)
    dc0 = g.to_device(c0)
    o0, o1 = g.array((ext, o.N)), g.array((ext, o.N))
    g.key_inner_product(dd, level, dk, o0, o1, c0=dc0 if with_c0 else None, galois=galois)
    e0, e1 = o.ksk_inner(digits, level, key)
    if with_c0:
        Pm = np.array([int(o.primes[i]) for i in range(level + 1)], dtype=object)
        from oracle.oracle import lib
        pmod = [int(lib().orc_P_mod(o.h, i)) for i in range(level + 1)]
        e0 = e0.copy()
        e0[: level + 1] = o.add(e0[: level + 1], o.mul_scalar(c0, np.array(pmod, dtype=np.uint64)))
    if galois != 1:
        e0, e1 = o.automorph(e0, galois), o.automorph(e1, galois)
    assert np.array_equal(o0.download(), e0)
    assert np.array_equal(o1.download(), e1)
    for x in (dd, dc0, o0, o1):
        x.free()
    dk.free()


@pytest.mark.parametrize("cfgname,level,npolys", [("cfg1", 3, 2), ("cfg1", 0, 1), ("cfg2", 24, 2), ("cfg2", 11, 1)])
def test_moddown(cfgname, level, npolys, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    ext = level + 1 + o.K
    x = np.concatenate([o.random_poly(90 + k, o.ext_pids(level)) for k in range(npolys)])
    d = g.to_device(x)
    out = g.array((npolys * (level + 1), o.N))
    g.moddown(d, level, npolys, out)
    exp = np.concatenate([o.moddown(x[k * ext:(k + 1) * ext], level) for k in range(npolys)])
    assert np.array_equal(out.download(), exp)
    d.free(); out.free()


@pytest.mark.parametrize("rot", [1, -1, 7, 1000, "conj"])
def test_automorph(rot, gctx1, octx1):
    g, o = gctx1, octx1
    k = 2 * o.N - 1 if rot == "conj" else o.rot_index(rot)
    if rot != "conj":
        assert g.galois_index(rot) == k
    x = o.random_poly(5, list(range(o.L1)))
    d, out = g.to_device(x), g.array(x.shape)
    g.automorph(d, o.L1, k, out)
    assert np.array_equal(out.download(), o.automorph(x, k))
    d.free(); out.free()


def _keys(o, seed=1):
    s = o.gen_secret(seed)
    return s


def _slots(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


@pytest.fixture(scope="module")
def cfg1_material(octx1):
    o = octx1
    s = o.gen_secret(11)
    relin = o.gen_relin_key(s, 12)
    k1 = o.rot_index(1)
    rot1 = o.gen_rot_key(s, k1, 13)
    n = o.N // 2
    z1, z2 = _slots(n, 1), _slots(n, 2)
    l = o.L1 - 1
    ct1 = o.encrypt(o.encode(z1, 2.0 ** 50, l), l, s, 21)
    ct2 = o.encrypt(o.encode(z2, 2.0 ** 50, l), l, s, 22)
    return dict(s=s, relin=relin, rot1=rot1, k1=k1, z1=z1, z2=z2, ct1=ct1, ct2=ct2, l=l)


def test_cfg1_mult_rescale_rotate_chain(gctx1, octx1, cfg1_material):
    """cfg1 pipeline: EvalMult + relinearize -> Rescale -> EvalRotate(+1); every limb bit-exact,
    decrypted slots equal z1*z2 rotated by one within 2^-25."""
    g, o, m = gctx1, octx1, cfg1_material
    l = m["l"]
    d1, d2 = g.to_device(m["ct1"]), g.to_device(m["ct2"])
    krel, krot = g.key(m["relin"]), g.key(m["rot1"])
    prod = g.array((2 * (l + 1), o.N))
    g.eval_mult(d1, d2, l, krel, prod)
    exp_prod = o.eval_mult(m["ct1"], m["ct2"], l, m["relin"])
    assert np.array_equal(prod.download(), exp_prod)
    rs = g.array((2 * l, o.N))
    g.rescale(prod, l, 2, rs)
    exp_rs = o.rescale_ct(exp_prod, l, 2)
    assert np.array_equal(rs.download(), exp_rs)
    rot = g.array((2 * l, o.N))
    g.eval_rotate(rs, l - 1, krot, m["k1"], rot)
    exp_rot = o.eval_rotate(exp_rs, l - 1, m["rot1"], m["k1"])
    got = rot.download()
    assert np.array_equal(got, exp_rot)
    # decrypt: scale after rescale = 2^100 / q_l
    scale = 2.0 ** 100 / float(o.primes[l])
    dec = o.decode(o.decrypt(got, l - 1, m["s"]), l - 1, o.N // 2, scale)
    want = np.roll(m["z1"] * m["z2"], -1)
    err = np.max(np.abs(dec - want))
    assert err < 2.0 ** -25, err
    for x in (d1, d2, prod, rs, rot):
        x.free()
    krel.free(); krot.free()


def test_relinearize_degree2(gctx1, octx1, cfg1_material):
    g, o, m = gctx1, octx1, cfg1_material
    l = m["l"]
    t = o.tensor(m["ct1"], m["ct2"], l)
    dt = g.to_device(t)
    k = g.key(m["relin"])
    out = g.array((2 * (l + 1), o.N))
    g.relinearize(dt, l, k, out)
    assert np.array_equal(out.download(), o.relinearize(t, l, m["relin"]))
    dt.free(); out.free(); k.free()


def test_rotate_ext_matches(gctx1, octx1, cfg1_material):
    g, o, m = gctx1, octx1, cfg1_material
    l = m["l"]
    d = g.to_device(m["ct1"])
    k = g.key(m["rot1"])
    ext = l + 1 + o.K
    for add_c0 in (True, False):
        out = g.array((2 * ext, o.N))
        g.eval_rotate_ext(d, l, k, m["k1"], out, add_c0=add_c0)
        assert np.array_equal(out.download(), o.eval_rotate_ext(m["ct1"], l, m["rot1"], m["k1"], add_c0))
        out.free()
    d.free(); k.free()


@pytest.fixture(scope="module")
def cfg2_material(octx2):
    o = octx2
    s = o.gen_secret(31)
    return dict(s=s, relin=o.gen_relin_key(s, 32))


@pytest.mark.parametrize("level", [24, 12])
def test_cfg2_eval_mult(level, gctx2, octx2, cfg2_material):
    g, o, m = gctx2, octx2, cfg2_material
    ids = _ids(o, level)
    ct1 = np.concatenate([o.random_poly(200 + level, ids), o.random_poly(201 + level, ids)])
    ct2 = np.concatenate([o.random_poly(202 + level, ids), o.random_poly(203 + level, ids)])
    d1, d2 = g.to_device(ct1), g.to_device(ct2)
    k = g.key(m["relin"])
    out = g.array((2 * (level + 1), o.N))
    g.eval_mult(d1, d2, level, k, out)
    assert np.array_equal(out.download(), o.eval_mult(ct1, ct2, level, m["relin"]))
    d1.free(); d2.free(); out.free(); k.free()


@pytest.mark.parametrize("n,level,square,e1,e2", [(2, 12, False, 0, 0), (3, 7, True, 0, 0), (1, 12, False, 0, 0), (2, 9, False, 3, 1),
                                                  (1, 5, False, 0, 4), (2, 6, True, 2, 2)])
def test_cfg2_eval_mult_many(n, level, square, e1, e2, gctx2, octx2, cfg2_material):
    """n products through one batched launch sequence equal the oracle's products one by one; the
    operands may hold e1 / e2 more towers than the product's level (read in place)."""
    g, o, m = gctx2, octx2, cfg2_material
    t1, t2 = level + 1 + e1, level + 1 + e2
    a = [np.concatenate([o.random_poly(300 + 4 * i, _ids(o, t1 - 1)), o.random_poly(301 + 4 * i, _ids(o, t1 - 1))]) for i in range(n)]
    b = a if square else [np.concatenate([o.random_poly(302 + 4 * i, _ids(o, t2 - 1)), o.random_poly(303 + 4 * i, _ids(o, t2 - 1))])
                          for i in range(n)]
    da = g.to_device(np.concatenate(a))
    db = da if square else g.to_device(np.concatenate(b))
    k = g.key(m["relin"])
    out = g.array((n * 2 * (level + 1), o.N))
    print(f"[eval_mult_many] n={n} level={level} square={square} towers {t1} {t2 if not square else t1}", flush=True)
    g.eval_mult_many(n, da, db, level, k, out, towers1=t1, towers2=t1 if square else t2)
    got = out.download()
    P, lq = 2 * (level + 1), level + 1

    def red(x, t):
        return np.concatenate([x[:lq], x[t:t + lq]])

    for i in range(n):
        bi = red(a[i], t1) if square else red(b[i], t2)
        assert np.array_equal(got[i * P:(i + 1) * P], o.eval_mult(red(a[i], t1), bi, level, m["relin"])), i
    da.free(); out.free(); k.free()
    if not square:
        db.free()


@pytest.mark.parametrize("rot", [1, 64])
def test_cfg2_eval_rotate_decrypts(rot, gctx2, octx2, cfg2_material):
    g, o, m = gctx2, octx2, cfg2_material
    level = 24
    kidx = o.rot_index(rot)
    rk = o.gen_rot_key(m["s"], kidx, 40 + rot)
    z = _slots(o.N // 2, 5)
    ct = o.encrypt(o.encode(z, 2.0 ** 50, level), level, m["s"], 41)
    d = g.to_device(ct)
    k = g.key(rk)
    out = g.array(ct.shape)
    g.eval_rotate(d, level, k, kidx, out)
    got = out.download()
    assert np.array_equal(got, o.eval_rotate(ct, level, rk, kidx))
    dec = o.decode(o.decrypt(got, level, m["s"]), level, o.N // 2, 2.0 ** 50)
    # noise bound: the approximate ModDown leaves a rounding term u*s with u in [0, K) per coefficient;
    # its mean (K-1)/2 concentrates near slot 0 (|sum_k zeta^k| ~ 2N/pi), about 2^-25 at N=2^16, K=9
    assert np.max(np.abs(dec - np.roll(z, -rot))) < 2.0 ** -22
    d.free(); out.free(); k.free()


@pytest.fixture(scope="module")
def cfg1_rot_keys(octx1, cfg1_material):
    o, m = octx1, cfg1_material
    steps = [1, 2, 5, -1]
    ks = {r: o.gen_rot_key(m["s"], o.rot_index(r), 300 + i) for i, r in enumerate(steps)}
    return steps, ks


@pytest.mark.parametrize("mix,fused", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_rotate_batch_matches_per_op(mix, fused, gctx1, octx1, cfg1_material, cfg1_rot_keys):
    """fh_rotate_batch over 7 ciphertexts with repeated keys == per-ciphertext oracle rotations
    (prime-major NTT job order, and with 60-bit-prime jobs spread among the others; with the
    targets' row stages fused into the key inner product, and separate)."""
    g, o, m = gctx1, octx1, cfg1_material
    g.tune_set("ntt_mix", mix)
    g.tune_set("rot_fused", fused)
    steps, ks = cfg1_rot_keys
    l = 2
    ids = list(range(l + 1))
    B = 7
    cts = [np.concatenate([o.random_poly(400 + 2 * b, ids), o.random_poly(401 + 2 * b, ids)]) for b in range(B)]
    pick = [steps[b % len(steps)] for b in (0, 1, 0, 2, 3, 0, 1)]
    dk = {r: g.key(ks[r]) for r in steps}
    d = g.to_device(np.concatenate(cts))
    out = g.array((B * 2 * (l + 1), o.N))
    try:
        g.rotate_batch(d, B, l, [dk[r] for r in pick], [o.rot_index(r) for r in pick], out)
        g.sync()
    finally:
        g.tune_set("ntt_mix", 0)
        g.tune_set("rot_fused", 0)
    got = out.download().reshape(B, 2 * (l + 1), o.N)
    for b in range(B):
        exp = o.eval_rotate(cts[b], l, ks[pick[b]], o.rot_index(pick[b]))
        assert np.array_equal(got[b], exp), b
    d.free(); out.free()
    for k in dk.values():
        k.free()


@pytest.mark.parametrize("B", [9, 8, 16])
def test_rotate_batch_two_streams(B, gctx1, octx1, cfg1_material, cfg1_rot_keys):
    """rot_streams=2 (two halves on two streams, pool blocks parked until both are issued) gives
    the oracle's rotations; repeated calls reuse the parked blocks afterwards."""
    g, o, m = gctx1, octx1, cfg1_material
    steps, ks = cfg1_rot_keys
    l = 3
    ids = list(range(l + 1))
    cts = [np.concatenate([o.random_poly(500 + 2 * b, ids), o.random_poly(501 + 2 * b, ids)]) for b in range(B)]
    pick = [steps[(3 * b) % len(steps)] for b in range(B)]
    dk = {r: g.key(ks[r]) for r in steps}
    d = g.to_device(np.concatenate(cts))
    rows = 2 * (l + 1)
    out = g.array((B * rows, o.N))
    print(f"[rotate_batch two streams] B={B}", flush=True)
    g.tune_set("rot_streams", 2)
    try:
        for rep in range(2):
            g.rotate_batch(d, B, l, [dk[r] for r in pick], [o.rot_index(r) for r in pick], out)
            g.sync()
            got = out.download().reshape(B, rows, o.N)
            for b in range(B):
                assert np.array_equal(got[b], o.eval_rotate(cts[b], l, ks[pick[b]], o.rot_index(pick[b]))), (rep, b)
    finally:
        g.tune_set("rot_streams", 1)
    d.free(); out.free()
    for k in dk.values():
        k.free()


@pytest.mark.parametrize("moddown", [True, False])
def test_rotate_hoisted(moddown, gctx1, octx1, cfg1_material, cfg1_rot_keys):
    g, o, m = gctx1, octx1, cfg1_material
    steps, ks = cfg1_rot_keys
    l = 3
    ct = m["ct1"]
    dk = {r: g.key(ks[r]) for r in steps}
    d = g.to_device(ct)
    ext = l + 1 + o.K
    per = 2 * (l + 1) if moddown else 2 * ext
    out = g.array((len(steps) * per, o.N))
    g.rotate_hoisted(d, l, [dk[r] for r in steps], [o.rot_index(r) for r in steps], out, moddown=moddown)
    got = out.download().reshape(len(steps), per, o.N)
    for i, r in enumerate(steps):
        k = o.rot_index(r)
        exp = o.eval_rotate(ct, l, ks[r], k) if moddown else o.eval_rotate_ext(ct, l, ks[r], k, True)
        assert np.array_equal(got[i], exp), r
    d.free(); out.free()
    for k in dk.values():
        k.free()


def test_fill_uniform_deterministic(gctx2, octx2):
    g, o = gctx2, octx2
    pids = [0, 1, 25, 33]
    a, b = g.array((2 * len(pids), o.N)), g.array((2 * len(pids), o.N))
    g.fill_uniform(a, pids, num_polys=2, seed=9)
    g.fill_uniform(b, pids, num_polys=2, seed=9)
    x, y = a.download(), b.download()
    assert np.array_equal(x, y)
    for i, p in enumerate(pids * 2):
        assert int(x[i].max()) < int(o.primes[p])
        assert len(np.unique(x[i])) > o.N - 8  # no visible collisions
    a.free(); b.free()
