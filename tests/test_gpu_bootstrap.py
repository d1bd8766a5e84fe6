"""Bootstrapping pieces on the GPU against the oracle, bit-exact: ModRaise (tower-0 lift to the
full chain)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfgname,l_in,ncomp", [("cfg1", 0, 2), ("cfg1", 2, 3), ("cfg2", 0, 2), ("cfg2", 5, 2)])
def test_mod_raise(cfgname, l_in, ncomp, gctx1, gctx2, octx1, octx2):
    g, o = (gctx1, octx1) if cfgname == "cfg1" else (gctx2, octx2)
    top = o.L1 - 1
    ct = np.concatenate([o.random_poly(90 + 7 * k + l_in, list(range(l_in + 1))) for k in range(ncomp)])
    print(f"[mod_raise] {cfgname} l_in={l_in} ncomp={ncomp}", flush=True)
    d = g.to_device(ct)
    for l in (top, top // 2):
        out = g.array((ncomp * (l + 1), o.N))
        g.mod_raise(d, l_in, ncomp, l, out)
        assert np.array_equal(out.download(), o.mod_raise(ct, l_in, l, ncomp)), l
        out.free()
    d.free()


# ---- the facade's EvalBootstrap / EvalChebyshevSeries (tests/cpp/bootstrap_check.cpp) ----
import json
import math
import os
import subprocess

import boot_replay as B
from conftest import PKG

BCHECK = os.path.join(PKG, "_build", "bootstrap_check")


# (logN, slots, CoeffsToSlots budget, SlotsToCoeffs budget, multiplicative depth, scaling bits,
# bits required, secret weight, technique 0 FIXEDMANUAL / 1 FLEXIBLEAUTO). The bounds sit 0.7-2 bits
# under what the runs measure (19.7, 21.0, 13.9, 26.9, 26.0, 25.7, 22.2 and 22.2 bits): the
# EvalMod series tracks exact scales, and with the nominal bookkeeping (each rescale taken to divide
# by the scale itself) the N=2^16 runs measured 12.3 bits. Weight 32 takes a degree-63 series and two
# double angles (23.4 bits at N=2^16 with one degree-255 series, 10 ms slower). The
# N=2^16 runs are the serving workload's bootstrap (all 32768 slots, budget {3,3}, h=32, 59-bit
# scaling under a 60-bit q_0): the reference states 11.5-15 bits for its GPU path at N=2^16
# (test/BootstrapPrecisionTests.cu:4-7); their residues are not replayed (the oracle's bootstrap at
# N=2^16 takes too long), nor those of FLEXIBLEAUTO runs (this bootstrap's own scale adjustment,
# not a reference sequence).
RUNS = [(12, 8, 1, 1, 15, 50, 18, 16, 0), (12, 64, 2, 3, 17, 50, 19, 16, 0), (12, 2048, 3, 3, 18, 50, 12, 16, 0),
        (12, 2048, 3, 3, 19, 59, 25, 16, 0), (12, 2048, 3, 3, 20, 59, 24, 32, 0), (12, 2048, 3, 3, 20, 59, 24, 32, 1),
        (16, 32768, 3, 3, 24, 59, 21.5, 32, 0), (16, 32768, 3, 3, 24, 59, 21.5, 32, 1)]


@pytest.fixture(scope="module", params=RUNS, ids=["N4096_slots8_b11", "N4096_slots64_b23", "N4096_full_b33", "N4096_full_b33_scale59",
                                                  "N4096_full_b33_scale59_h32", "N4096_full_b33_scale59_h32_flexibleauto",
                                                  "N65536_full_b33_scale59_h32", "N65536_full_b33_scale59_h32_flexibleauto"])
def boot_run(request, tmp_path_factory):
    logN, slots, b1, b2, depth, sbits, bits, hw, tech = request.param
    if not os.path.exists(BCHECK):
        pytest.fail("bootstrap_check is not built (run __graft_entry__.build())")
    out = tmp_path_factory.mktemp(f"boot{slots}")
    # the N=2^16 runs also run the two-iteration bootstrap (Meta-BTS) at precision 12
    args = [str(out), str(logN), str(slots), "7", str(b1), str(b2), str(depth), str(sbits), str(hw), str(tech), "12" if logN == 16 else "0"]
    print(f"[bootstrap_check] {BCHECK} {' '.join(args)}", flush=True)
    r = subprocess.run([BCHECK] + args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    with open(out / "meta.json") as f:
        meta = json.load(f)
    res["bits_required"] = bits
    return out, res, meta


def test_bootstrap_precision(boot_run):
    """Decrypted slots after EvalBootstrap are within the run's stated bound of the input; a second
    bootstrap of the same ciphertext gives the same residues; the result supports another
    product."""
    out, res, meta = boot_run
    print(res)
    assert res["towers_in"] == (2 if meta["correction"] or meta["flex"] else 1) and res["towers_out"] == res["L1"] - res["depth"]
    assert res["boot_err"] < 2.0 ** -res["bits_required"], res
    assert res["deterministic"]
    assert res["square_err"] < 4 * 2.0 ** -res["bits_required"]
    assert res["cheb_err_vs_series"] < 1e-7 and res["cheb_err_vs_function"] < 1e-6
    if res["meta_precision"]:
        # the second iteration refreshes the first one's error: far beyond the single bootstrap
        print({"boot_bits": res["boot_bits"], "meta_bits": res["meta_bits"]})
        assert res["meta_bits"] >= res["boot_bits"] + 10, res


def test_bootstrap_precision_at_reference_point(tmp_path):
    """The reference's own bootstrap precision point (test/BootstrapPrecisionTests.cu:1-7, 58-82):
    N=2^16, depth 25, 55-bit scaling under a 56-bit q_0, level budget {3,2}, dnum 6, FLEXIBLEAUTO,
    sparse ternary secret (OpenFHE's SPARSE_TERNARY draws weight 192), 2^15 slots of real values in
    [-1, 1] encrypted at level 21. The reference states 18.6 bits for its OpenFHE path, 11.5-15
    bits for its GPU path and 29-32 bits for the two-iteration (Meta-BTS) refresh there. This
    bootstrap (degree-63 series, three double angles) measures 15.7 bits and Meta-BTS 27.8: one
    iteration is held to the GPU path's upper bound, Meta-BTS to 27 bits."""
    if not os.path.exists(BCHECK):
        pytest.fail("bootstrap_check is not built (run __graft_entry__.build())")
    args = [str(tmp_path), "16", "32768", "7", "3", "2", "25", "55", "192", "1", "12", "56", "6", "1", "21"]
    print(f"[bootstrap_check] {BCHECK} {' '.join(args)}", flush=True)
    r = subprocess.run([BCHECK] + args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print({k: res[k] for k in ("technique", "L1", "depth", "towers_in", "towers_out", "boot_bits", "meta_bits", "boot_ms", "meta_ms")})
    assert res["technique"] == "FLEXIBLEAUTO" and res["L1"] == 26 and res["towers_in"] == 26 - 21
    assert res["boot_bits"] >= 15.0, res
    assert res["deterministic"]
    assert res["meta_bits"] >= 27.0, res


def _u64(out, name, rows, N):
    return np.fromfile(out / f"{name}.u64", dtype=np.uint64).reshape(rows, N)


def _c128(out, name):
    a = np.fromfile(out / f"{name}.f64", dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


def _groups(out, meta, name):
    n = meta["slots"]
    gs = []
    for g, (rowSize, stride, offset) in enumerate(meta[name]):
        flat = _c128(out, f"{name}{g}").reshape(rowSize, n)
        gs.append((rowSize, stride, offset, [flat[kk] for kk in range(rowSize)]))
    return gs


def test_bootstrap_matches_oracle_replay(boot_run, oracle_mod):
    """The oracle replays the session (keys and encryptions from the same seeds, the same
    bootstrap and series sequences over the dumped constants): every residue agrees."""
    out, res, meta = boot_run
    if meta["logN"] > 12 or meta["flex"]:
        pytest.skip("residue replay at N=2^12 with FIXEDMANUAL only (the precision test covers this run)")
    O = oracle_mod
    logN, n, seed, L1, K, scale = meta["logN"], meta["slots"], meta["seed"], meta["L1"], meta["K"], meta["scale"]
    N = 1 << logN
    moduli = np.array([int(v) for v in meta["moduli"]], dtype=np.uint64)
    o = O.Ctx(logN, moduli, L1, K, meta["dnum"])
    s = o.gen_secret(seed + 0, meta["hw"])
    relin = o.gen_relin_key(s, seed + 2)
    k = dict(slots=n, K=meta["Krange"], r=meta["r"], pre=meta["pre"], correction=meta["correction"],
             cheb=list(np.fromfile(out / "cheb.f64", dtype=np.float64)),
             cts=_groups(out, meta, "cts"), stc=_groups(out, meta, "stc"))
    order = []
    for r in B.rotation_steps(k, N):
        g = o.rot_index(r)
        if g not in order:
            order.append(g)
    if 2 * N - 1 not in order:
        order.append(2 * N - 1)
    rot = {g: o.gen_rot_key(s, g, seed + 3 + i) for i, g in enumerate(order)}
    enc = seed + 3 + len(order)
    z = _c128(out, "z")
    tin = 2 if meta["correction"] else 1
    ct = o.encrypt(o.encode(z, scale, tin - 1), tin - 1, s, enc)
    assert np.array_equal(_u64(out, "ct", 2 * tin, N), ct), "input encryption differs"
    ops = B.Ops(o, scale, relin, rot)
    bt = B.bootstrap(ops, (ct, tin, 1), k)
    assert bt[1] == meta["bt_towers"]
    got = _u64(out, "bt", 2 * bt[1], N)
    bad = np.count_nonzero(got != bt[0])
    assert bad == 0, f"bootstrap: {bad} residues differ from the oracle replay"
    # the Chebyshev series
    x = _c128(out, "x")
    cx = o.encrypt(o.encode(x, scale, L1 - 1), L1 - 1, s, enc + 1)
    assert np.array_equal(_u64(out, "cx", 2 * L1, N), cx)
    cy = B.chebyshev_series(ops, (cx, L1, 1), list(np.fromfile(out / "series.f64", dtype=np.float64)), -2.0, 2.0)
    assert cy[1] == meta["cy_towers"]
    assert np.array_equal(_u64(out, "cy", 2 * cy[1], N), cy[0]), "Chebyshev series differs from the oracle replay"


def test_bootstrap_constants_closed_form(boot_run):
    """The dumped transforms are the scaled special-FFT matrices and the series approximates the
    shifted cosine (constants checked against their definitions, independently of the replay)."""
    out, res, meta = boot_run
    n, N = meta["slots"], 1 << meta["logN"]
    q0, scale, Kr, r = float(meta["moduli"][0]), meta["scale"], meta["Krange"], meta["r"]
    gap = N // 2 // n
    fc = scale / (2 * gap * q0 * Kr)
    if n <= 256:
        xi = np.exp(2j * math.pi / (4 * n))
        U = np.array([[xi ** ((pow(5, j, 4 * n) * kk) % (4 * n)) for kk in range(n)] for j in range(n)])
        bits = n.bit_length() - 1
        P = np.zeros((n, n))
        for kk in range(n):
            P[int(format(kk, f"0{bits}b")[::-1], 2) if bits else 0, kk] = 1
        UP = U @ np.linalg.inv(P)   # the special FFT without its opening bit reversal

        def product(groups):
            A = np.eye(n, dtype=complex)
            for rowSize, stride, offset, dv in groups:
                G = np.zeros((n, n), dtype=complex)
                for kk in range(rowSize):
                    G[np.arange(n), (np.arange(n) + offset + kk * stride) % n] += dv[kk]
                A = G @ A
            return A

        # (the first group also carries the rescale-drift factor prod(q_r / scale))
        assert np.allclose(product(_groups(out, meta, "cts")) * meta["pre"], fc * meta["scale_fix"] * np.linalg.inv(UP), rtol=0,
                           atol=1e-12 * fc)
        assert np.allclose(product(_groups(out, meta, "stc")), q0 / (2 * math.pi * scale) * UP, rtol=0, atol=1e-9)
    assert 1.0 <= fc / meta["pre"] < 2.0 and math.log2(meta["pre"]).is_integer()
    cheb = np.fromfile(out / "cheb.f64", dtype=np.float64)
    # weight <= 21: one degree-127 series; weight 32: degree 63 and two double angles (the same
    # depth as one degree-255 series, 21 instead of 33 products)
    assert Kr == min(math.ceil(meta["hw"] / 2) + 4, 28)
    assert (r, len(cheb)) == ((0, 128) if meta["hw"] <= 21 else (2, 64))
    # the drift factor: the top primes the prescale and the CoeffsToSlots groups rescale by, over the scale
    L1 = meta["L1"]
    fix = 1.0
    for j in range(len(meta["cts"]) + 1):
        fix *= float(int(meta["moduli"][L1 - 1 - j])) / scale
    assert abs(meta["scale_fix"] / fix - 1) < 1e-15
    ys = np.linspace(-1, 1, 101)
    approx = np.polynomial.chebyshev.chebval(ys, np.concatenate([[cheb[0] / 2], cheb[1:]]))
    assert np.max(np.abs(approx - np.cos((2 * math.pi * Kr * ys - math.pi / 2) / 2 ** r))) < 1e-12
