import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fideslib-fork_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


# Parameter sets of the workloads (moduli by OpenFHE's FIXEDMANUAL rule).
#   cfg1: N=2^13, Q = 60 + 3x50 bits, dnum=2, K by OpenFHE's rule (=2)
#   cfg2: N=2^16, Q = 60 + 24x50 bits, dnum=3, K=9
CFG1 = dict(logN=13, L1=4, K=2, dnum=2)
CFG2 = dict(logN=16, L1=25, K=9, dnum=3)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle


def make_oracle_ctx(oracle, cfg):
    primes = oracle.gen_primes(cfg["logN"], cfg["L1"], 60, 50, cfg["K"])
    return oracle.Ctx(cfg["logN"], primes, cfg["L1"], cfg["K"], cfg["dnum"])


@pytest.fixture(scope="session")
def octx1(oracle_mod):
    return make_oracle_ctx(oracle_mod, CFG1)


@pytest.fixture(scope="session")
def octx2(oracle_mod):
    return make_oracle_ctx(oracle_mod, CFG2)


@pytest.fixture(scope="session")
def fh():
    import fides_hip
    return fides_hip


@pytest.fixture(scope="session")
def gctx1(fh):
    c = fh.Context(CFG1["logN"], CFG1["L1"], CFG1["K"], CFG1["dnum"])
    yield c
    c.close()


@pytest.fixture(scope="session")
def gctx2(fh):
    c = fh.Context(CFG2["logN"], CFG2["L1"], CFG2["K"], CFG2["dnum"])
    yield c
    c.close()
