"""Serialization of the facade's objects (tests/cpp/serial_check.cpp): a context rebuilt from its
file has the same moduli; deserialized keys and ciphertexts hold the same device words; the same
computation (EvalMult, Rescale, EvalRotate with deserialized evaluation keys) gives the same
residues and decryption in both contexts; a fresh encryption under the deserialized public key
decrypts under the deserialized secret key; JSON, files of another kind and corrupt files (a residue
not below its prime in a ciphertext or key, a slot count that is not a power of two) are refused."""
import json
import os
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

RCHECK = os.path.join(PKG, "_build", "serial_check")


def test_serialization_round_trip(tmp_path):
    if not os.path.exists(RCHECK):
        pytest.fail("serial_check is not built (run __graft_entry__.build())")
    print(f"[serial_check] {RCHECK} {tmp_path}", flush=True)
    r = subprocess.run([RCHECK, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["ok"] and res["same_moduli"] and res["same_ciphertext"]
    assert res["same_result"] and res["same_decryption"]
    assert res["fresh_encryption_err"] < 1e-6
    assert res["json_refused"] and res["wrong_kind_refused"]
    assert res["corrupt_refused"]
