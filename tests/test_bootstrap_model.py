"""The bootstrap and Chebyshev-series sequences of the facade (restated in boot_replay over the
oracle) decrypt to the right values: what pins the composition on the CPU. The GPU run of the
facade is compared with this replay residue by residue in test_gpu_bootstrap.py."""
import math

import numpy as np
import pytest

import boot_replay as B
from oracle import oracle as O


def _ctx(logN=10, L1=16, K=6, dnum=3, sbits=50):
    primes = O.gen_primes(logN, L1, 60, sbits, K)
    return O.Ctx(logN, primes, L1, K, dnum)


def _keys(o, s, steps, N, seed):
    relin = o.gen_relin_key(s, seed)
    rot = {}
    for i, g in enumerate(sorted({o.rot_index(r) for r in steps} | {2 * N - 1})):
        rot[g] = o.gen_rot_key(s, g, seed + 1 + i)
    return relin, rot


def test_special_fft_matrices_are_inverse():
    n, M = 16, 2 ** 11
    U, Ui = B.matrix_of(n, M, False), B.matrix_of(n, M, True)
    assert np.allclose(U @ Ui, np.eye(n), atol=1e-12)
    # U evaluates at the slot roots: U[j][k] = xi^(5^j k), xi = exp(2 pi i / 4n) (the root of the
    # subring in X^(N/2n) that n slots occupy)
    xi = np.exp(2j * math.pi / (4 * n))
    for j in range(n):
        r = pow(5, j, 4 * n)
        for k in range(n):
            assert abs(U[j, k] - xi ** ((r * k) % (4 * n))) < 1e-12


def test_chebyshev_coefficients_closed_form():
    c = B.cheb_coefficients(lambda x: 3.0 * x ** 3 - x, -1.0, 1.0, 7)   # 3x^3 - x = 0.75 T3 + 1.25 T1
    want = [0.0, 1.25, 0.0, 0.75, 0, 0, 0, 0]
    assert np.allclose(c, want, atol=1e-13)


class _LevelOps:
    """Tracks (towers, scale degree) only: the level structure of a series evaluation."""

    def rescale(self, y):
        return (None, y[1] - 1, y[2] - 1)

    def mult(self, a, b):
        return (None, min(a[1], b[1]), a[2] + b[2])

    def add(self, a, b):
        assert a[2] == b[2]
        return (None, min(a[1], b[1]), a[2])

    sub = add

    def add_scalar(self, a, c):
        return a

    def mul_scalar(self, a, c):
        return (None, a[1], a[2] + 1)

    def reduce(self, a, t):
        return (None, t, a[2])


def test_chebyshev_depth_rule():
    """The levels a series consumes equal the planned depth, never exceed ceil(log2(d + 1)) + 1,
    and are j + 1 for degree 2^j - 1 (the bootstrap's 63, 127, 255)."""
    for d in range(1, 300):
        out = B.chebyshev_series(_LevelOps(), (None, 40, 1), [1.0] * (d + 1))
        assert out[2] == 1
        assert 40 - out[1] == B.cheb_depth(d) <= (d).bit_length() + 1, d
    for j in range(3, 9):
        assert B.cheb_depth(2 ** j - 1) == j + 1
    assert B.cheb_leaf(15) and B.cheb_leaf(13) and not B.cheb_leaf(12) and not B.cheb_leaf(16)


@pytest.mark.parametrize("degree,depth", [(27, 5), (16, 5), (15, 5)])
def test_chebyshev_series_decrypts(degree, depth):
    o = _ctx()
    N, n, scale = o.N, 16, 2.0 ** 50
    s = o.gen_secret(11, 16)
    relin, rot = _keys(o, s, [], N, 20)
    ops = B.Ops(o, scale, relin, rot)
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, n)
    l = o.L1 - 1
    ct = (o.encrypt(o.encode(x, scale, l), l, s, 5), l + 1, 1)
    f = lambda v: math.tanh(1.5 * v) + 0.25 * v * v
    c = B.cheb_coefficients(f, -2.0, 2.0, degree)
    out = B.chebyshev_series(ops, ct, c, -2.0, 2.0)
    got = o.decode(o.decrypt(out[0], out[1] - 1, s), out[1] - 1, n, scale).real
    want = np.polynomial.chebyshev.chebval(np.asarray(x) / 2.0, np.concatenate([[c[0] / 2], c[1:]]))
    assert np.max(np.abs(got - want)) < 1e-6
    assert out[1] == l + 1 - 1 - depth     # input scaling 1 + the series


def test_factored_transforms_compose_to_the_special_fft():
    """The butterfly groups of every level budget multiply out to the scaled special FFT (SlotsToCoeffs)
    and its inverse up to the bit-reversal permutation that the two transforms cancel."""
    n, M = 64, 2 ** 11
    U = B.matrix_of(n, M, False)
    P = np.zeros((n, n))
    for k in range(n):
        P[int(format(k, "06b")[::-1], 2), k] = 1
    for budget in (1, 2, 3, 6):
        for inverse, f in ((False, 3.0), (True, 0.5)):
            A = np.eye(n, dtype=complex)
            for rowSize, stride, offset, dv in B.factored(n, M, budget, inverse, f):
                G = np.zeros((n, n), dtype=complex)
                for kk in range(rowSize):
                    d = (offset + kk * stride) % n
                    G[np.arange(n), (np.arange(n) + d) % n] += dv[kk]
                A = G @ A
            want = f * (np.linalg.inv(U @ np.linalg.inv(P)) if inverse else U @ np.linalg.inv(P))
            assert np.max(np.abs(A - want)) < 1e-12, (budget, inverse)


# bits: what the series with exact scales reaches with a margin of about one bit; the nominal
# bookkeeping (rescales taken to divide by the scale itself) loses 1 to 3.5 bits at 50-bit scaling
# and misses the first four bounds
@pytest.mark.parametrize("n,budget,L1,sbits,corr,bits,hw", [(8, (1, 1), 16, 50, 0, 20, 16), (64, (1, 1), 16, 50, 0, 20.5, 16),
                                                           (64, (2, 3), 18, 50, 0, 17.5, 16), (512, (3, 3), 18, 50, 0, 15.5, 16),
                                                           (512, (3, 3), 19, 59, 9, 23.5, 16), (512, (3, 3), 20, 59, 9, 23.5, 32)])
def test_bootstrap_decrypts(n, budget, L1, sbits, corr, bits, hw):
    o = _ctx(L1=L1, K=9 if sbits > 50 else 6, sbits=sbits)
    N, scale = o.N, 2.0 ** sbits
    s = o.gen_secret(7, hw)
    k = B.constants(N, n, hw, scale, float(o.q(0)), budget=budget, correction=corr)
    relin, rot = _keys(o, s, B.rotation_steps(k, N), N, 30)
    ops = B.Ops(o, scale, relin, rot)
    rng = np.random.default_rng(n)
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    tin = 2 if corr else 1
    ct = (o.encrypt(o.encode(z, scale, tin - 1), tin - 1, s, 9), tin, 1)
    out = B.bootstrap(ops, ct, k)
    l = out[1] - 1
    got = o.decode(o.decrypt(out[0], l, s), l, n, scale)
    err = np.max(np.abs(got - z))
    print(f"bootstrap n={n} budget={budget}: towers {out[1]}, max error {err:.3g} ({-math.log2(err):.1f} bits)")
    # weight 16: one degree-127 series; weight 32: two double angles after degree 63 (the same depth
    # as degree 255 with 21 instead of 33 products)
    assert (k["r"], len(k["cheb"])) == ((0, 128) if hw <= 21 else (2, 64))
    assert out[1] == o.L1 - (1 + len(k["cts"]) + (8 if hw <= 21 else 9) + len(k["stc"]))
    assert err < 2.0 ** -bits
