"""Device CKKS encoding (fh_encode_slots / fh_encode_coeffs) against the oracle's encoder, bit for
bit: the special inverse FFT in fp64 with the host encoder's rounding sequence, the gap scatter,
the residues and the forward NTT; batched plaintexts; Q and Q-and-P prime sets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _values(n, seed, count=1, mag=1.0):
    rng = np.random.default_rng(seed)
    return mag * (rng.uniform(-1, 1, (count, n)) + 1j * rng.uniform(-1, 1, (count, n)))


@pytest.mark.parametrize("slots,level,scale_bits,count", [(32768, 24, 50, 1), (8, 3, 50, 2), (4096, 12, 40, 3), (1, 0, 50, 1), (2048, 24, 59, 2)])
def test_encode_slots_matches_oracle(slots, level, scale_bits, count, gctx2, octx2):
    g, o = gctx2, octx2
    z = _values(slots, slots + level, count)
    scale = 2.0 ** scale_bits
    pids = list(range(level + 1))
    out = g.array((count * (level + 1), o.N))
    print(f"[encode_slots] slots={slots} level={level} count={count}", flush=True)
    g.encode_slots(z, scale, pids, out)
    got = out.download()
    for b in range(count):
        assert np.array_equal(got[b * (level + 1):(b + 1) * (level + 1)], o.encode(z[b], scale, level)), b
    out.free()


def test_encode_slots_extended_strided(gctx2, octx2):
    """Q_l and P limbs (the LinearTransform diagonal layout), outputs strided in one buffer."""
    g, o = gctx2, octx2
    level, slots, count = 20, 32768, 2
    pids = list(range(level + 1)) + [o.L1 + k for k in range(o.K)]
    ext = len(pids)
    z = _values(slots, 77, count, mag=2.0)
    stride = (ext + 1) * o.N
    out = g.array((count * (ext + 1), o.N))
    g.encode_slots(z, 2.0 ** 50, pids, out, out_stride=stride)
    got = out.download()
    for b in range(count):
        assert np.array_equal(got[b * (ext + 1):b * (ext + 1) + ext], o.encode_ext(z[b], 2.0 ** 50, level)), b
    out.free()


def test_encode_coeffs_matches_oracle(gctx2, octx2):
    g, o = gctx2, octx2
    rng = np.random.default_rng(5)
    count, pids = 2, [0, 3, 24, o.L1, o.L1 + 8]
    m = rng.integers(-(2 ** 62), 2 ** 62, (count, o.N), dtype=np.int64)
    m[0, :4] = [0, -1, 2 ** 62 - 1, -(2 ** 62)]
    out = g.array((count * len(pids), o.N))
    g.encode_coeffs(m, pids, out)
    got = out.download()
    for b in range(count):
        res = np.stack([np.array([int(v) % int(o.primes[p]) for v in m[b]], dtype=np.uint64) for p in pids])
        assert np.array_equal(got[b * len(pids):(b + 1) * len(pids)], o.ntt(res, pids)), b
    out.free()


def test_encode_rejects_bad_input(gctx1, fh):
    g = gctx1
    out = g.array((2, g.N))
    with pytest.raises(fh.FHError):
        g.encode_slots(np.ones(6), 2.0 ** 40, [0, 1], out)             # slots not a power of two
    with pytest.raises(fh.FHError):
        g.encode_slots(np.full(4, 4.0), 2.0 ** 62, [0, 1], out)         # |value| * scale beyond 2^63
    with pytest.raises(fh.FHError, match="2\\^61"):
        g.encode_slots(np.array([8.0, 0.0]), 2.0 ** 58, [0, 1], out)    # coefficient 2^61: OpenFHE's approximate path
    with pytest.raises(fh.FHError):
        g.encode_slots(np.array([np.nan, 1.0]), 2.0 ** 40, [0, 1], out)
    out.free()


def test_encode_at_the_2_61_boundary(gctx2, octx2):
    """Coefficients just below 2^61 (OpenFHE's MAX_BITS_IN_WORD) still encode exactly, bit for bit
    with the oracle; 2^61 itself is rejected by both encoders."""
    g, o = gctx2, octx2
    pids = [0, 1, 2]
    z = np.array([5.0 - 2.999j])                                        # one slot: coefficients z * 2^58
    out = g.array((3, o.N))
    g.encode_slots(z, 2.0 ** 58, pids, out)
    assert np.array_equal(out.download(), o.encode(z, 2.0 ** 58, 2))
    with pytest.raises(ValueError):
        o.encode(np.array([8.0 + 0j]), 2.0 ** 58, 2)
    out.free()
