"""The FLEXIBLEAUTO restatement (tests/flex_replay.py) pinned on the CPU by what the technique must
achieve, independent of the facade: the scale tables satisfy their recurrence, and every
operation (each branch of the level / depth alignment included) decrypts to the expected slot
values at the scale it reports. A dropped rescale, a wrong adjustment constant or a scale copied
from the wrong operand moves the decryption by orders of magnitude beyond the tolerance."""
import numpy as np
import pytest

from flex_replay import FCt, Flex, scale_tables
from oracle import oracle as O


@pytest.fixture(scope="module")
def flex():
    logN, L1, K, dnum = 10, 6, 2, 2
    primes = O.gen_primes(logN, L1, 60, 50, K)
    o = O.Ctx(logN, primes, L1, K, dnum)
    s = o.gen_secret(5)
    F = Flex(o, relin=o.gen_relin_key(s, 6))
    n = o.N // 2
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.8, 0.8, n) + 1j * rng.uniform(-0.8, 0.8, n)
    y = rng.uniform(-0.8, 0.8, n) + 1j * rng.uniform(-0.8, 0.8, n)
    l = L1 - 1
    top = F.real[l]
    ct1 = FCt(o.encrypt(o.encode(x, top, l), l, s, 7), L1, 2, 1, top)
    ct2 = FCt(o.encrypt(o.encode(y, top, l), l, s, 8), L1, 2, 1, top)
    return dict(o=o, F=F, s=s, x=x, y=y, ct1=ct1, ct2=ct2, n=n, L1=L1, py=F.plain(o.encode(y, top, l), L1))


def test_scale_tables_recurrence():
    q = [int(v) for v in O.gen_primes(13, 12, 60, 50, 3)[:12]]
    real, big, mrf = scale_tables(q)
    assert real[11] == float(q[11])
    for l in range(11):
        assert real[l] == real[l + 1] * real[l + 1] / float(q[l + 1])
        assert big[l] == real[l] * real[l]
    assert mrf == [float(v) for v in q]
    # a chain whose factors drift past a factor 2 of the top one is refused, as OpenFHE does
    with pytest.raises(ValueError):
        scale_tables([(1 << 50) + 1, (1 << 49) + 1, (1 << 51) + 1])


def _check(F, s, c, expect, n, tol=1e-6):
    got = F.decrypt(c, s, n)
    err = float(np.max(np.abs(got - expect)))
    assert err < tol, err


def test_flex_operations_decrypt(flex):
    F, s, n, x, y, ct1, ct2, L1 = (flex[k] for k in ("F", "s", "n", "x", "y", "ct1", "ct2", "L1"))
    m = F.mult(ct1, ct2)
    assert (m.towers, m.deg, m.scale) == (L1, 2, ct1.scale * ct2.scale)   # left at depth 2
    _check(F, s, m, x * y, n)
    _check(F, s, F.add(m, ct1), x * y + x, n)
    m2 = F.mult(m, ct1)
    assert (m2.towers, m2.deg) == (L1 - 1, 2)
    _check(F, s, m2, x * y * x, n)
    m3 = F.mult(m2, ct2)
    assert (m3.towers, m3.deg) == (L1 - 2, 2)
    _check(F, s, m3, x * y * x * y, n)
    lr1, lr2 = F.drop(ct2, L1 - 1), F.drop(ct2, L1 - 2)
    for a, b, e in [(m, lr1, x * y + y), (m, lr2, x * y + y), (m, m3, x * y + (x * y) ** 2), (ct1, lr2, x + y),
                    (ct1, F.negate(m3), x - (x * y) ** 2)]:
        r = F.add(a, b)
        assert r.towers == min(a.towers, b.towers)
        _check(F, s, r, e, n, 1e-5)
    _check(F, s, F.add_scalar(m2, 0.5), x * y * x + 0.5, n)
    _check(F, s, F.add_scalar(m2, -1.5), x * y * x - 1.5, n)
    _check(F, s, F.mult_scalar(m2, -0.75), x * y * x * -0.75, n)
    _check(F, s, F.add_plain(m2, flex["py"]), x * y * x + y, n)
    _check(F, s, F.mult_plain(m, flex["py"]), x * y * y, n)
    _check(F, s, F.wsum([ct1, m2, ct2], [0.5, -1.25, 2.0]), 0.5 * x - 1.25 * x * y * x + 2.0 * y, n)


def test_flex_adjustment_is_needed(flex):
    """The alignment matters: adding a depth-2 product to a depth-1 ciphertext without the depth
    adjustment (plain residue addition) decrypts to garbage."""
    F, s, n, x, y, ct1, ct2, o = (flex[k] for k in ("F", "s", "n", "x", "y", "ct1", "ct2", "o"))
    m = F.mult(ct1, ct2)
    wrong = FCt(np.concatenate([o.add(m.comp(k), ct1.comp(k)) for k in range(2)]), m.towers, 2, 2, m.scale)
    err = float(np.max(np.abs(F.decrypt(wrong, s, n) - (x * y + x))))
    assert err > 1e3 * 1e-6
