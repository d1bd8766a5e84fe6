"""The C-ABI library loads and exports every function include/fideslib_hip.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fideslib_hip.h")
LIB = os.path.join(ROOT, "fideslib-fork_amd", "_build", "libfideslib_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "fh_ntt_forward" in names and "fh_eval_rotate" in names
    assert len(names) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfideslib_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfideslib_hip.so not built")
def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in declared():
        assert n in syms, n


def test_no_cpu_fallback_in_binding():
    """The Python binding only forwards to the shared library; it never imports the oracle."""
    src = open(os.path.join(ROOT, "fideslib-fork_amd", "fides_hip.py")).read()
    assert "oracle" not in src.replace("oracle/", "")
