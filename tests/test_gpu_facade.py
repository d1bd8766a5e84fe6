"""The C++ facade (fideslib::CryptoContextImpl over the C ABI) against the oracle.

tests/cpp/facade_check.cpp runs one CKKS session through the facade on the GPU (keys, encoding,
secret- and public-key encryption, EvalAdd/EvalSub/EvalMult with ciphertext, plaintext and scalar
operands, Rescale, rotations, conjugation, hoisted and batched rotation, LinearTransform) and
writes every ciphertext. The oracle recomputes the session from the same seeds; every ciphertext
must agree bit for bit, the CKKS metadata must follow the reference's rules (noise degree, scale,
level), and every decryption must be within the noise of the expected slot values.
Mirrors the reference's facade tests (test/ApiTests.cpp:95-330).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "_build", "facade_check")
SEED = 1234


@pytest.fixture(scope="module")
def session(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("facade_check is not built (run __graft_entry__.build())")
    out = tmp_path_factory.mktemp("facade")
    print(f"[facade_check] {EXE} {out} {SEED}", flush=True)
    r = subprocess.run([EXE, str(out), str(SEED)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    with open(out / "manifest.json") as f:
        man = json.load(f)
    return out, man


def _u64(out, name, shape=None):
    a = np.fromfile(out / f"{name}.u64", dtype=np.uint64)
    return a.reshape(shape) if shape else a


def _c128(out, name):
    a = np.fromfile(out / f"{name}.f64", dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


@pytest.fixture(scope="module")
def recomputed(session, oracle_mod):
    """The oracle's replay of the facade session."""
    out, man = session
    o_ = oracle_mod
    p = man["params"]
    L1, K, N = p["L1"], p["K"], 1 << p["logN"]
    moduli = np.array([int(v) for v in man["moduli"]], dtype=np.uint64)
    # the facade's moduli are OpenFHE's FIXEDMANUAL chain
    assert np.array_equal(moduli, o_.gen_primes(p["logN"], L1, 60, 50, K))
    o = o_.Ctx(p["logN"], moduli, L1, K, p["dnum"])
    seed, scale, l = p["seed"], p["scale"], L1 - 1
    s = o.gen_secret(seed + 0)
    pk = o.encrypt(o.zeros(L1), l, s, seed + 1)
    pk_b, pk_a = pk[:L1], pk[L1:]
    relin = o.gen_relin_key(s, seed + 2)
    gal = [int(v) for v in man["galois"]]
    # 5^r mod 2N, 5^(N/2 - 3) for r = -3 (5 has order N/2 in Z_2N^*)
    assert gal == [pow(5, r % (N // 2), 2 * N) for r in (1, 2, -3, 7)]
    rot = {g: o.gen_rot_key(s, g, seed + 3 + i) for i, g in enumerate(gal)}
    rot[2 * N - 1] = o.gen_rot_key(s, 2 * N - 1, seed + 7)
    for i, r in enumerate((3, 4, 8, 12)):      # the fold keys, drawn after the two encryptions
        rot[pow(5, r, 2 * N)] = o.gen_rot_key(s, pow(5, r, 2 * N), seed + 10 + i)
    x, y = _c128(out, "x"), _c128(out, "y")
    ct1 = o.encrypt(o.encode(x, scale, l), l, s, seed + 8)
    ct2 = o.encrypt_pk(o.encode(y, scale, l), l, pk_b, pk_a, seed + 9)
    return dict(o=o, s=s, relin=relin, rot=rot, x=x, y=y, ct1=ct1, ct2=ct2, l=l, scale=scale, N=N, moduli=moduli)


def _check(session, name, expected, ncomp, towers, deg, scale, tol):
    out, man = session
    m = man[name]
    assert (m["ncomp"], m["towers"], m["noise_deg"]) == (ncomp, towers, deg), (name, m)
    assert m["scale"] == pytest.approx(scale, rel=1e-15), (name, m)
    got = _u64(out, name, expected.shape)
    bad = np.count_nonzero(got != expected)
    assert bad == 0, f"{name}: {bad} residues differ from the oracle"
    assert m["err"] < tol, (name, m["err"])


def test_keys_and_encryption(session, recomputed):
    out, man = session
    R = recomputed
    np.testing.assert_array_equal(_u64(out, "secret", R["s"].shape), R["s"])
    sc = R["scale"]
    t = R["l"] + 1
    _check(session, "ct1", R["ct1"], 2, t, 1, sc, 1e-9)
    _check(session, "ct2", R["ct2"], 2, t, 1, sc, 1e-9)


def _comp(ct, t, i):
    return ct[i * t:(i + 1) * t]


def test_add_sub_negate(session, recomputed, oracle_mod):
    R = recomputed
    o, c1, c2, t, sc = R["o"], R["ct1"], R["ct2"], R["l"] + 1, R["scale"]
    both = lambda f, a, b: np.concatenate([f(_comp(a, t, i), _comp(b, t, i)) for i in range(2)])
    _check(session, "add", both(o.add, c1, c2), 2, t, 1, sc, 1e-9)
    _check(session, "sub", both(o.sub, c1, c2), 2, t, 1, sc, 1e-9)
    _check(session, "neg", np.concatenate([o.neg(_comp(c1, t, i)) for i in range(2)]), 2, t, 1, sc, 1e-9)
    pty = o.encode(R["y"], sc, R["l"])
    _check(session, "add_p", np.concatenate([o.add(_comp(c1, t, 0), pty), _comp(c1, t, 1)]), 2, t, 1, sc, 1e-9)
    q = [int(v) for v in R["moduli"][:t]]
    ea = oracle_mod.elem_for_eval_add(q, 0.25, sc, 1)
    _check(session, "add_s", np.concatenate([o.add_scalar(_comp(c1, t, 0), ea), _comp(c1, t, 1)]), 2, t, 1, sc, 1e-9)
    es = oracle_mod.elem_for_eval_add(q, 1.5, sc, 1)
    neg = np.array([(qi - int(v)) % qi for v, qi in zip(es, q)], dtype=np.uint64)
    _check(session, "sub_s", np.concatenate([o.add_scalar(_comp(c1, t, 0), neg), _comp(c1, t, 1)]), 2, t, 1, sc, 1e-9)
    _check(session, "lvl_red", np.concatenate([_comp(c1, t, i)[:t - 1] for i in range(2)]), 2, t - 1, 1, sc, 1e-9)


def test_multiplications(session, recomputed, oracle_mod):
    R = recomputed
    o, c1, c2, l, sc = R["o"], R["ct1"], R["ct2"], R["l"], R["scale"]
    t = l + 1
    m12 = o.eval_mult(c1, c2, l, R["relin"])
    _check(session, "mult", m12, 2, t, 2, sc * sc, 1e-6)
    _check(session, "mult_rs", o.rescale_ct(m12, l), 2, t - 1, 1, sc, 1e-6)
    nr = o.tensor(c1, c2, l)
    _check(session, "mult_nr", nr, 3, t, 2, sc * sc, 1e-6)
    _check(session, "relin", o.relinearize(nr, l, R["relin"]), 2, t, 2, sc * sc, 1e-6)
    pty = o.encode(R["y"], sc, l)
    mp = np.concatenate([o.mul(_comp(c1, t, i), pty) for i in range(2)])
    _check(session, "mult_p", o.rescale_ct(mp, l), 2, t - 1, 1, sc, 1e-6)
    _check(session, "square", o.rescale_ct(o.eval_mult(c2, c2, l, R["relin"]), l), 2, t - 1, 1, sc, 1e-6)
    q = [int(v) for v in R["moduli"][:t]]
    for name, c in (("mult_s", 2.5), ("mult_sn", -0.75)):
        f = oracle_mod.elem_for_eval_mult(q, c, sc)
        ms = np.concatenate([o.mul_scalar(_comp(c1, t, i), f) for i in range(2)])
        _check(session, name, o.rescale_ct(ms, l), 2, t - 1, 1, sc, 1e-6)


def test_rotations(session, recomputed):
    R = recomputed
    o, c1, c2, l, sc, N = R["o"], R["ct1"], R["ct2"], R["l"], R["scale"], R["N"]
    t = l + 1
    g = lambda r: pow(5, r % (N // 2), 2 * N)
    _check(session, "rot1", o.eval_rotate(c1, l, R["rot"][g(1)], g(1)), 2, t, 1, sc, 1e-6)
    _check(session, "rotm3", o.eval_rotate(c2, l, R["rot"][g(-3)], g(-3)), 2, t, 1, sc, 1e-6)
    _check(session, "conj", o.eval_rotate(c1, l, R["rot"][2 * N - 1], 2 * N - 1), 2, t, 1, sc, 1e-6)
    _check(session, "acc3", o.accumulate_sum(c1, l, 4, 1, 3, [R["rot"][g(1)], R["rot"][g(2)]], [g(1), g(2)]), 2, t, 1, sc, 1e-6)
    ks16 = [g(r) for r in (1, 2, 3, 4, 8, 12)]
    _check(session, "acc16", o.accumulate_sum(c1, l, 4, 1, 16, [R["rot"][k] for k in ks16], ks16), 2, t, 1, sc, 1e-5)
    ks4 = [g(r) for r in (4, 8, 12)]
    _check(session, "acc16s4", o.accumulate_sum_from(c1, l, 4, 1, 16, [R["rot"][k] for k in ks4], ks4, 4), 2, t, 1, sc, 1e-5)
    _check(session, "fast7", o.eval_rotate(c1, l, R["rot"][g(7)], g(7)), 2, t, 1, sc, 1e-6)
    # EvalFastRotationExt + KeySwitchDown from the same precomputed digits
    _check(session, "fastext7", o.eval_rotate(c1, l, R["rot"][g(7)], g(7)), 2, t, 1, sc, 1e-6)
    _check(session, "fastextm3", o.eval_rotate(c1, l, R["rot"][g(-3)], g(-3)), 2, t, 1, sc, 1e-6)
    _check(session, "many0", o.eval_rotate(c1, l, R["rot"][g(2)], g(2)), 2, t, 1, sc, 1e-6)
    _check(session, "many1", o.eval_rotate(c2, l, R["rot"][g(2)], g(2)), 2, t, 1, sc, 1e-6)


def test_linear_transform(session, recomputed):
    out, _ = session
    R = recomputed
    o, c1, l, sc, N = R["o"], R["ct1"], R["l"], R["scale"], R["N"]
    t = l + 1
    rowSize, bStep = 4, 2
    D = [_c128(out, f"diag{k}") for k in range(rowSize)]
    diags = np.concatenate([o.encode_ext(np.roll(D[k], bStep * (k // bStep)), sc, l) for k in range(rowSize)])
    g = lambda r: pow(5, r % (N // 2), 2 * N)
    res = o.linear_transform(c1, l, diags, bStep, 1, [None, R["rot"][g(1)]], [1, g(1)], R["rot"][g(2)], g(2))
    _check(session, "lt", res, 2, t, 2, sc * sc, 1e-6)
    _check(session, "lt_prepared", res, 2, t, 2, sc * sc, 1e-6)   # PrepareLinearTransform + LinearTransformInPlace


def test_decryption_matches_oracle_decode(session, recomputed):
    """The facade's Decrypt (host CRT + special FFT) equals the oracle's decode bit for bit."""
    out, man = session
    R = recomputed
    o, l, sc = R["o"], R["l"], R["scale"]
    for name, ncomp in (("ct1", 2), ("mult_nr", 3), ("mult_rs", 2)):
        m = man[name]
        t = m["towers"]
        ct = _u64(out, name, (ncomp * t, R["N"]))
        want = o.decode(o.decrypt(ct, t - 1, R["s"], ncomp), t - 1, m["slots"], m["scale"])
        np.testing.assert_array_equal(_c128(out, f"{name}_dec"), want)


def test_scalar_encodings(session, recomputed, oracle_mod):
    _, man = session
    R = recomputed
    q = [int(v) for v in R["moduli"][:man["params"]["L1"]]]
    sc = R["scale"]
    for i, c in enumerate((2.5, -0.75, 1e-9, 3.0e30, -7.25e25, 0.0)):
        assert [int(v) for v in man[f"em{i}"]] == [int(v) for v in oracle_mod.elem_for_eval_mult(q, c, sc)], c
    for i, c in enumerate((0.25, 1.5, 1e12, 3.3e20)):
        for d in (1, 2):
            assert [int(v) for v in man[f"ea{i}_{d}"]] == [int(v) for v in oracle_mod.elem_for_eval_add(q, c, sc, d)], (c, d)


def test_api_forms(session):
    """In-place forms, MakeCKKSPackedPlaintextDeviceMany, the rotation-index helpers,
    LinearTransformMany and Compress agree with the plain forms on every device word."""
    _, man = session
    ok, total = man["api_forms"]
    assert total == 8 and ok == total, man["api_forms"]


def test_error_paths(session):
    _, man = session
    assert man["errors_thrown"] == 4
