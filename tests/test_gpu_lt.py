"""Hoisted BSGS linear transform on the GPU against the oracle's restatement of the reference's
extended-basis formulation: every limb bit-exact, decrypted result = sum_k Rot(z, k*stride + offset) D_k."""
import numpy as np
import pytest

from conftest import make_oracle_ctx

pytestmark = pytest.mark.gpu


def _case(o, s, z, l, rowSize, bStep, stride, offset, scale_pt, seed):
    n = o.N // 2
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (rowSize, n)) + 1j * rng.uniform(-1, 1, (rowSize, n))
    diags = np.concatenate([o.encode_ext(np.roll(D[k], bStep * (k // bStep) * stride + offset), scale_pt, l) for k in range(rowSize)])
    want = sum(np.roll(z, -(k * stride + offset)) * D[k] for k in range(rowSize))
    return diags, want


def _run(g, o, s, l, rowSize, bStep, stride, offset, scale_pt=2.0 ** 30, seed=0, check_decrypt=True):
    n = o.N // 2
    z = np.random.default_rng(seed + 1).uniform(-1, 1, n) + 1j * np.random.default_rng(seed + 2).uniform(-1, 1, n)
    ct = o.encrypt(o.encode(z, 2.0 ** 50, l), l, s, seed + 3)
    diags, want = _case(o, s, z, l, rowSize, bStep, stride, offset, scale_pt, seed)
    bk = [None] + [o.gen_rot_key(s, o.rot_index(i * stride), 900 + i) for i in range(1, bStep)]
    bkk = [1] + [o.rot_index(i * stride) for i in range(1, bStep)]
    gk = o.gen_rot_key(s, o.rot_index(bStep * stride), 950)
    ok = o.gen_rot_key(s, o.rot_index(offset), 960) if offset else None
    exp = o.linear_transform(ct, l, diags, bStep, stride, bk, bkk, gk, o.rot_index(bStep * stride), ok,
                             o.rot_index(offset) if offset else 1, offset)
    dct, dd = g.to_device(ct), g.to_device(diags)
    dbk = [None] + [g.key(k) for k in bk[1:]]
    dgk = g.key(gk)
    dok = g.key(ok) if offset else None
    out = g.array(ct.shape)
    g.linear_transform(dct, l, rowSize, bStep, stride, dd, dbk, bkk, out, giant_key=dgk, giant_galois=o.rot_index(bStep * stride),
                       offset_key=dok, offset_galois=o.rot_index(offset) if offset else 1, offset=offset)
    got = out.download()
    assert np.array_equal(got, exp), "linear transform differs from the oracle"
    if check_decrypt:
        dec = o.decode(o.decrypt(got, l, s), l, n, 2.0 ** 50 * scale_pt)
        assert np.max(np.abs(dec - want)) < 1e-5
    for x in [dct, dd, out] + dbk[1:] + [dgk] + ([dok] if dok else []):
        x.free()


@pytest.mark.parametrize("rowSize,bStep,offset", [(12, 4, 0), (10, 4, 3), (8, 8, 0), (16, 4, 5), (40, 16, 0), (29, 12, 2)])
def test_linear_transform_cfg1(rowSize, bStep, offset, gctx1, octx1):
    s = octx1.gen_secret(70)
    _run(gctx1, octx1, s, 3, rowSize, bStep, 1, offset, seed=rowSize + offset)


def test_linear_transform_cfg4_shape(gctx2, octx2):
    """The cfg4 shape: N=2^16, level 20 (21 limbs), 64 diagonals, bStep = gStep = 8, stride 1."""
    s = octx2.gen_secret(71)
    _run(gctx2, octx2, s, 20, 64, 8, 1, 0, seed=4, check_decrypt=True)


def _batch_case(g, o, s, l, rowSize, bStep, stride, offset, count, seed, scale_pt=2.0 ** 30):
    """count ciphertexts through fh_linear_transform_batch: each result equals the oracle's LT."""
    n = o.N // 2
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (rowSize, n)) + 1j * rng.uniform(-1, 1, (rowSize, n))
    diags = np.concatenate([o.encode_ext(np.roll(D[k], bStep * (k // bStep) * stride + offset), scale_pt, l) for k in range(rowSize)])
    bk = [None] + [o.gen_rot_key(s, o.rot_index(i * stride), 900 + i) for i in range(1, bStep)]
    bkk = [1] + [o.rot_index(i * stride) for i in range(1, bStep)]
    gk = o.gen_rot_key(s, o.rot_index(bStep * stride), 950)
    ok = o.gen_rot_key(s, o.rot_index(offset), 960) if offset else None
    cts = [o.encrypt(o.encode(rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n), 2.0 ** 50, l), l, s, seed + 10 + b) for b in range(count)]
    exp = [o.linear_transform(ct, l, diags, bStep, stride, bk, bkk, gk, o.rot_index(bStep * stride), ok,
                              o.rot_index(offset) if offset else 1, offset) for ct in cts]
    dcts, dd = g.to_device(np.concatenate(cts)), g.to_device(diags)
    dbk = [None] + [g.key(k) for k in bk[1:]]
    dgk = g.key(gk)
    dok = g.key(ok) if offset else None
    out = g.array((count * 2 * (l + 1), o.N))
    g.linear_transform_batch(dcts, count, l, rowSize, bStep, stride, dd, dbk, bkk, out, giant_key=dgk,
                             giant_galois=o.rot_index(bStep * stride), offset_key=dok,
                             offset_galois=o.rot_index(offset) if offset else 1, offset=offset)
    got = out.download().reshape(count, 2 * (l + 1), o.N)
    for b in range(count):
        assert np.array_equal(got[b], exp[b]), f"batched linear transform, ciphertext {b}, differs from the oracle"
    for x in [dcts, dd, out] + dbk[1:] + [dgk] + ([dok] if dok else []):
        x.free()


@pytest.mark.parametrize("rowSize,bStep,offset,count", [(12, 4, 0, 3), (10, 4, 3, 5), (8, 8, 0, 2), (16, 4, 5, 4)])
def test_linear_transform_batch_cfg1(rowSize, bStep, offset, count, gctx1, octx1):
    s = octx1.gen_secret(72)
    _batch_case(gctx1, octx1, s, 3, rowSize, bStep, 1, offset, count, seed=3 * rowSize + offset)


def test_linear_transform_batch_cfg4_shape(gctx2, octx2):
    """The LT workload's shape: N=2^16, level 20, 64 diagonals, bStep = gStep = 8, two ciphertexts."""
    s = octx2.gen_secret(73)
    _batch_case(gctx2, octx2, s, 20, 64, 8, 1, 0, 2, seed=5)


def _run_flat(g, o, s, l, rowSize, bStep, stride, offset, seed, scale_pt=2.0 ** 30):
    """fh_linear_transform_flat: every limb equal to the oracle's flat form, decrypting to the map."""
    n = o.N // 2
    z = np.random.default_rng(seed + 1).uniform(-1, 1, n) + 1j * np.random.default_rng(seed + 2).uniform(-1, 1, n)
    ct = o.encrypt(o.encode(z, 2.0 ** 50, l), l, s, seed + 3)
    diags, want = _case(o, s, z, l, rowSize, bStep, stride, offset, scale_pt, seed)
    bk = [None] + [o.gen_rot_key(s, o.rot_index(i * stride), 900 + i) for i in range(1, bStep)]
    bkk = [1] + [o.rot_index(i * stride) for i in range(1, bStep)]
    gStep = (rowSize + bStep - 1) // bStep
    wraps = [(j * bStep * stride) % n == 0 for j in range(1, gStep)]
    gks = [1 if w else o.rot_index(j * bStep * stride) for j, w in zip(range(1, gStep), wraps)]
    gkeys = [None if w else o.gen_rot_key(s, k, 940 + j) for j, (k, w) in enumerate(zip(gks, wraps))]
    ok = o.gen_rot_key(s, o.rot_index(offset), 960) if offset else None
    okk = o.rot_index(offset) if offset else 1
    exp = o.linear_transform_flat(ct, l, diags, bStep, stride, bk, bkk, gkeys, gks, ok, okk, offset)
    dct, dd = g.to_device(ct), g.to_device(diags)
    dbk = [None] + [g.key(k) for k in bk[1:]]
    dgk = [None if k is None else g.key(k) for k in gkeys]
    dok = g.key(ok) if offset else None
    out = g.array(ct.shape)
    g.linear_transform_flat(dct, l, rowSize, bStep, stride, dd, dbk, bkk, out, giant_keys=dgk, giant_galois=gks, offset_key=dok,
                            offset_galois=okk, offset=offset)
    got = out.download()
    assert np.array_equal(got, exp), "flat linear transform differs from the oracle"
    dec = o.decode(o.decrypt(got, l, s), l, n, 2.0 ** 50 * scale_pt)
    assert np.max(np.abs(dec - want)) < 1e-5
    for x in [dct, dd, out] + dbk[1:] + [k for k in dgk if k is not None] + ([dok] if dok else []):
        x.free()


@pytest.mark.parametrize("rowSize,bStep,stride,offset", [(12, 4, 1, 0), (10, 4, 1, 3), (16, 4, 1, 5), (8, 8, 1, 0), (12, 4, -8, 0),
                                                         (63, 16, 1, 3)])
def test_linear_transform_flat_cfg1(rowSize, bStep, stride, offset, gctx1, octx1):
    """stride -8 stands for a stride of slots/8: the second giant multiple (8 strides) wraps to 0 mod
    slots and is added without a rotation."""
    s = octx1.gen_secret(74)
    if stride < 0:
        stride = octx1.N // 2 // -stride
    print(f"[lt_flat] rowSize={rowSize} bStep={bStep} stride={stride} offset={offset}", flush=True)
    _run_flat(gctx1, octx1, s, 3, rowSize, bStep, stride, offset, seed=7 * rowSize + offset)


def test_linear_transform_flat_cfg4_shape(gctx2, octx2):
    """N=2^16, level 20, 64 diagonals, bStep = gStep = 8: seven giant rotations in one batch."""
    s = octx2.gen_secret(75)
    _run_flat(gctx2, octx2, s, 20, 64, 8, 1, 0, seed=6)
