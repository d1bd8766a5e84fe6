"""FLEXIBLEAUTO through the facade (tests/cpp/flex_check.cpp) against the oracle's restatement of
the technique (tests/flex_replay.py): the per-level scale tables, and for every operation of the
session (products left at depth 2, every branch of the level / depth alignment, scalar and
plaintext operands, rotation, weighted sum) the residues bit for bit, the metadata (towers, depth,
scale) exactly, and the decryption within the noise of the expected slots."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG
from flex_replay import FCt, Flex, scale_tables

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "_build", "flex_check")
SEED = 77


@pytest.fixture(scope="module")
def session(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("flex_check is not built (run __graft_entry__.build())")
    out = tmp_path_factory.mktemp("flex")
    print(f"[flex_check] {EXE} {out} {SEED}", flush=True)
    r = subprocess.run([EXE, str(out), str(SEED)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    with open(out / "manifest.json") as f:
        return out, json.load(f)


def _c128(out, name):
    a = np.fromfile(out / f"{name}.f64", dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


@pytest.fixture(scope="module")
def replay(session, oracle_mod):
    out, man = session
    O = oracle_mod
    p = man["params"]
    L1, K = p["L1"], p["K"]
    moduli = np.array([int(v) for v in man["moduli"]], dtype=np.uint64)
    assert np.array_equal(moduli, O.gen_primes(p["logN"], L1, 60, 50, K))
    o = O.Ctx(p["logN"], moduli, L1, K, p["dnum"])
    seed = p["seed"]
    s = o.gen_secret(seed + 0)
    F = Flex(o, relin=o.gen_relin_key(s, seed + 2))
    k1 = o.rot_index(1)
    rot = o.gen_rot_key(s, k1, seed + 3)
    x, y = _c128(out, "x"), _c128(out, "y")
    l = L1 - 1
    top = F.real[l]
    mx, my = o.encode(x, top, l), o.encode(y, top, l)
    ct1 = FCt(o.encrypt(mx, l, s, seed + 4), L1, 2, 1, top)
    ct2 = FCt(o.encrypt(my, l, s, seed + 5), L1, 2, 1, top)
    px, py = F.plain(mx, L1), F.plain(my, L1)
    m = F.mult(ct1, ct2)
    m2 = F.mult(m, ct1)
    m3 = F.mult(m2, ct2)
    lr1, lr2 = F.drop(ct2, L1 - 1), F.drop(ct2, L1 - 2)
    res = {
        "ct1": ct1, "ct2": ct2, "m": m, "m_add": F.add(m, ct1), "m2": m2, "m3": m3,
        "a_d2d1_adj": F.add(m, lr1), "a_d2d1": F.add(m, lr2), "a_d2d2": F.add(m, m3), "a_d1d1": F.add(ct1, lr2),
        "a_d1d2": F.add(ct1, F.negate(m3)), "add_s": F.add_scalar(m2, 0.5), "mul_s": F.mult_scalar(m2, -0.75),
        "mul_p": F.mult_plain(m, px), "add_p": F.add_plain(m2, py), "sub_p": F.add_plain(ct1, py, sub=True),
        "rot": F.rotate(m2, rot, k1), "wsum": F.wsum([ct1, m2, ct2], [0.5, -1.25, 2.0]),
    }
    return dict(o=o, F=F, res=res, man=man, out=out, L1=L1)


def test_flex_scale_tables(replay):
    man, F, L1 = replay["man"], replay["F"], replay["L1"]
    t = man["tables"]
    # OpenFHE level numbering in the manifest (level 0 = all towers); the replay's by towers - 1
    assert t["real"] == [F.real[L1 - 1 - k] for k in range(L1)]
    assert t["big"] == [F.big[L1 - 1 - k] for k in range(L1)]
    assert t["mrf"] == [F.mrf[L1 - 1 - k] for k in range(L1)]
    assert man["pt_scale"] == F.real[L1 - 1]
    # closed form: the top factor is the top prime, each level the one above squared over its prime
    q = [int(v) for v in man["moduli"][:L1]]
    assert F.real[L1 - 1] == float(q[L1 - 1])
    for l in range(L1 - 1):
        assert abs(F.real[l] * q[l + 1] / F.real[l + 1] ** 2 - 1) < 1e-15
        assert 0.5 < F.real[l] / F.real[L1 - 1] < 2.0


@pytest.mark.parametrize("name", ["ct1", "ct2", "m", "m_add", "m2", "m3", "a_d2d1_adj", "a_d2d1", "a_d2d2", "a_d1d1", "a_d1d2", "add_s",
                                  "mul_s", "mul_p", "add_p", "sub_p", "rot", "wsum"])
def test_flex_session_bit_exact(replay, name):
    man, out, exp = replay["man"], replay["out"], replay["res"][name]
    m = man[name]
    got = np.fromfile(out / f"{name}.u64", dtype=np.uint64)
    assert (m["towers"], m["ncomp"], m["deg"]) == (exp.towers, exp.ncomp, exp.deg), name
    assert m["scale"] == exp.scale, f"{name}: scale {m['scale']} != {exp.scale}"
    assert np.array_equal(got, exp.data.reshape(-1)), f"{name}: residues differ from the oracle's FLEXIBLEAUTO replay"
    assert m["err"] < 1e-5, f"{name}: decryption error {m['err']}"
