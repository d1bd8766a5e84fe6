"""Pins the CPU oracle (oracle/ckks_oracle.c) against the reference's known-answer vector, closed
forms computed here with Python big integers, and decryption invariants. CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def small_ctx(logN=10, L1=4, K=2, dnum=2):
    primes = O.gen_primes(logN, L1, 60, 50, K)
    return O.Ctx(logN, primes, L1, K, dnum)


@pytest.fixture(scope="module")
def sc():
    return small_ctx()


def test_mt19937_64_stream():
    # the 10000th output of std::mt19937_64 with the default seed (C++ standard, [rand.predef])
    assert O.lib().orc_mt_first(5489, 9999) == 9981545732273789042


@pytest.mark.parametrize("logN,L1,K", [(13, 4, 2), (16, 25, 9)])
def test_prime_chain_rule(logN, L1, K):
    m = 2 << logN
    p = [int(x) for x in O.gen_primes(logN, L1, 60, 50, K)]
    assert all(is_prime(x) and x % m == 1 for x in p)
    assert len(set(p)) == len(p)
    # top of the chain: the first prime = 1 mod 2N above 2^50
    top = p[L1 - 1]
    assert top > 2 ** 50 and all(not is_prime(c) for c in range(2 ** 50 + 1, top, m))
    # q_0: the last prime = 1 mod 2N below 2^60
    assert p[0] < 2 ** 60 and all(not is_prime(c) for c in range(p[0] + m, 2 ** 60, m))
    # alternation: PreviousPrime / NextPrime walking down from the top
    lo, hi = top, top
    for cnt, i in enumerate(range(L1 - 2, 0, -1)):
        if cnt % 2 == 0:
            c = lo - m
            while not is_prime(c):
                c -= m
            lo = c
        else:
            c = hi + m
            while not is_prime(c):
                c += m
            hi = c
        assert p[i] == c
    # special primes: descending primes = 1 mod 2N below 2^60, skipping the Q primes
    c = 2 ** 60 + 1
    got = []
    while len(got) < K:
        c -= m
        if is_prime(c) and c not in p[:L1]:
            got.append(c)
    assert p[L1:] == got


def test_default_special_count():
    p = O.gen_primes(13, 4, 60, 50, 0)
    assert O.lib().orc_default_num_special(np.ascontiguousarray(p[:4]), 4, 2, 60) == 2
    p2 = O.gen_primes(16, 25, 60, 50, 0)
    digit_bits = max(int(np.prod([int(x) for x in p2[d * 9:(d + 1) * 9]], dtype=object)).bit_length() for d in range(3))
    assert O.lib().orc_default_num_special(np.ascontiguousarray(p2[:25]), 25, 3, 60) == -(-digit_bits // 60)


def test_root_of_unity_is_minimal_primitive():
    for q, m in ((17, 16), (97, 32), (12289, 2048)):
        r = int(O.lib().orc_root_of_unity(m, q))
        prim = [x for x in range(2, q) if pow(x, m // 2, q) == q - 1 and pow(x, m, q) == 1]
        assert r == min(prim)
    c = small_ctx(11, 2, 1, 1)
    for i in range(3):
        q, psi, m = c.q(i), c.psi[i], 2 * c.N
        assert pow(psi, m // 2, q) == q - 1
        w = psi
        for k in range(3, 200, 2):  # no smaller root among a prefix of odd powers
            w = w * psi * psi % q
            assert pow(psi, k, q) >= psi


@pytest.mark.parametrize("logN,L1,K", [(13, 4, 2), (16, 25, 9)])
def test_root_of_unity_minimal_over_all_odd_powers(logN, L1, K):
    """At the workloads' real primes: psi has order exactly 2N, and it is the smallest of all
    primitive 2N-th roots, which are exactly psi^k for the N odd k < 2N (exhaustive)."""
    c = O.Ctx(logN, O.gen_primes(logN, L1, 60, 50, K), L1, K, 2 if logN == 13 else 3)
    N = c.N
    for i in range(L1 + K):
        q, psi = c.q(i), c.psi[i]
        assert pow(psi, N, q) == q - 1
        p2 = psi * psi % q
        w = psi
        for _ in range(N - 1):                      # psi^3, psi^5, ..., psi^(2N-1)
            w = w * p2 % q
            assert w > psi, (i, q)


def test_encode_bound_2_61(sc):
    """The encoders round exactly, which is OpenFHE's Encode only below 2^61 (MAX_BITS_IN_WORD): the
    oracle rejects slot values whose coefficients could reach 2^61 and encodes the largest ones
    below it exactly (one slot: the coefficients are re * scale and im * scale)."""
    c = sc
    with pytest.raises(ValueError):
        c.encode(np.array([8.0 + 0j]), 2.0 ** 58, 1)
    with pytest.raises(ValueError):
        c.encode(np.array([4.0 + 4.0j]), 2.0 ** 58, 1)            # |re| + |im| counts
    m = c.encode(np.array([5.0 - 2.999j]), 2.0 ** 58, 1)
    coef = c.intt(m, [0, 1])                                   # back to coefficients mod q_0, q_1
    q0 = c.q(0)
    re, im = 5 * 2 ** 58, int(round(-2.999 * 2 ** 58))
    assert int(coef[0][0]) == re % q0 and int(coef[0][c.N // 2]) == im % q0


def test_kat_q17_n8():
    """Reference known-answer case (test/NegacyclicNTTtests.cu:31-122): q=17, N=8,
    v = {1,2,3,7,5,4,1,2}: round trip, and NTT product == schoolbook negacyclic product."""
    q, N = 17, 8
    psi = int(O.lib().orc_root_of_unity(2 * N, q))
    assert psi == 3
    tw, itw = np.zeros(N, np.uint64), np.zeros(N, np.uint64)
    O.lib().orc_make_tables(N, q, psi, tw, itw)
    v = np.array([1, 2, 3, 7, 5, 4, 1, 2], dtype=np.uint64)
    a = v.copy()
    O.lib().orc_ntt_fwd_raw(a, N, q, tw)
    # closed form: A[j] = sum_i v_i psi^((2 brev(j)+1) i)
    for j in range(N):
        e = 2 * brev(j, 3) + 1
        assert int(a[j]) == sum(int(v[i]) * pow(psi, e * i, q) for i in range(N)) % q
    b = a.copy()
    O.lib().orc_ntt_inv_raw(b, N, q, itw, pow(N, q - 2, q))
    assert np.array_equal(b, v)
    w = np.array([3, 0, 16, 1, 1, 5, 2, 9], dtype=np.uint64)
    wa = w.copy()
    O.lib().orc_ntt_fwd_raw(wa, N, q, tw)
    prod = (a * wa) % q
    O.lib().orc_ntt_inv_raw(prod, N, q, itw, pow(N, q - 2, q))
    school = [0] * N
    for i in range(N):
        for k in range(N):
            t = int(v[i]) * int(w[k])
            if i + k < N:
                school[i + k] += t
            else:
                school[i + k - N] -= t
    assert [int(x) for x in prod] == [s % q for s in school]


def test_ntt_closed_form_and_convolution(sc):
    c = sc
    N = c.N
    for pid in (0, 1, c.L1):  # 60-bit, 50-bit and special prime
        q, psi = c.q(pid), c.psi[pid]
        x = c.random_poly(3 + pid, [pid])
        X = c.ntt(x, [pid])
        xi = [int(t) for t in x[0]]
        for j in (0, 1, 2, 3, N // 2, N - 1, 777):
            w = pow(psi, 2 * brev(j, c.logN) + 1, q)
            acc, p = 0, 1
            for i in range(N):
                acc += xi[i] * p
                p = p * w % q
            assert acc % q == int(X[0, j])
        assert np.array_equal(c.intt(X, [pid]), x)
    # negacyclic convolution: NTT(a) * NTT(b) -> a * b mod X^N + 1 (small operands)
    rng = np.random.default_rng(0)
    a = rng.integers(-50, 50, N)
    b = rng.integers(-50, 50, N)
    full = np.convolve(a, b)
    neg = full[:N].copy()
    neg[: N - 1] -= full[N:]
    q = c.q(1)
    A = c.ntt(np.array([a % q], dtype=np.uint64), [1])
    B = c.ntt(np.array([b % q], dtype=np.uint64), [1])
    Pm = c.intt(c.mul(np.vstack([A, A]), np.vstack([B, B]))[1:2], [1])  # limb 1 = prime 1
    assert [int(t) for t in Pm[0]] == [int(t) % q for t in neg]


def crt(res, mods):
    M = 1
    for m in mods:
        M *= m
    x = 0
    for r, m in zip(res, mods):
        Mi = M // m
        x += int(r) * Mi * pow(Mi, -1, m)
    return x % M, M


def test_modup_closed_form(sc):
    c = sc
    N, K = c.N, c.K
    for l in (3, 2, 0):
        x = c.random_poly(10 + l, list(range(l + 1)))
        dig = c.modup(x, l)
        ext = l + 1 + K
        xc = c.intt(x, list(range(l + 1)))
        pids = c.ext_pids(l)
        for d in range(c.num_digits(l)):
            start, end = d * c.alpha, min((d + 1) * c.alpha, l + 1)
            blk = dig[d * ext:(d + 1) * ext]
            assert np.array_equal(blk[start:end], x[start:end])  # own limbs untouched (EVAL)
            blkc = c.intt(blk, pids)
            qs = [c.q(i) for i in range(start, end)]
            Qd = 1
            for qq in qs:
                Qd *= qq
            for n in (0, 1, N // 3, N - 1):
                # fast base conversion: V = sum_i [x_i (Qd/q_i)^-1]_{q_i} (Qd/q_i)
                V = 0
                for i, qq in zip(range(start, end), qs):
                    Qh = Qd // qq
                    V += (int(xc[i, n]) * pow(Qh, -1, qq) % qq) * Qh
                X, _ = crt([xc[i, n] for i in range(start, end)], qs)
                assert (V - X) % Qd == 0 and 0 <= (V - X) // Qd < len(qs)
                for pos in range(ext):
                    if start <= pos < end:
                        continue
                    assert int(blkc[pos, n]) == V % c.q(pids[pos])


def test_moddown_closed_form(sc):
    c = sc
    N, K, L1 = c.N, c.K, c.L1
    for l in (3, 1):
        pids = c.ext_pids(l)
        x = c.random_poly(20 + l, pids)
        out = c.moddown(x, l)
        xc = c.intt(x, pids)
        oc = c.intt(out, list(range(l + 1)))
        ps = [c.q(L1 + k) for k in range(K)]
        P = 1
        for p in ps:
            P *= p
        for n in (0, 5, N - 1):
            V = 0
            for k, p in enumerate(ps):
                Ph = P // p
                V += (int(xc[l + 1 + k, n]) * pow(Ph, -1, p) % p) * Ph
            for i in range(l + 1):
                qi = c.q(i)
                assert int(oc[i, n]) == (int(xc[i, n]) - V) * pow(P, -1, qi) % qi
        # approximate division: ModDown(X) = round-ish(X / P) within K
        for n in (2, 9):
            X, M = crt([xc[j, n] for j in range(len(pids))], [c.q(p) for p in pids])
            Y, Mq = crt([oc[i, n] for i in range(l + 1)], [c.q(i) for i in range(l + 1)])
            diff = (Y - X // P) % Mq
            assert min(diff, Mq - diff) <= K


def test_reduced_noise_conversions_closed_form():
    """ReducedNoise: every prescaled source y above q/2 enters the conversion as y - q, so ModUp
    yields sum_i centred(y_i) Qhat_i mod p_j (and differs from the plain conversion exactly by
    -count * Q_d), and ModDown converts with centred sources too."""
    c = small_ctx()
    c.set_reduced_noise(True)
    N, K, L1 = c.N, c.K, c.L1
    cen = lambda v, q: v - q if v > q // 2 else v
    for l in (3, 1):
        x = c.random_poly(40 + l, list(range(l + 1)))
        dig = c.modup(x, l)
        ext = l + 1 + K
        xc = c.intt(x, list(range(l + 1)))
        pids = c.ext_pids(l)
        lifted_seen = 0
        for d in range(c.num_digits(l)):
            start, end = d * c.alpha, min((d + 1) * c.alpha, l + 1)
            blkc = c.intt(dig[d * ext:(d + 1) * ext], pids)
            qs = [c.q(i) for i in range(start, end)]
            Qd = 1
            for qq in qs:
                Qd *= qq
            for n in (0, 3, N // 2, N - 1):
                ys = [int(xc[i, n]) * pow(Qd // qq, -1, qq) % qq for i, qq in zip(range(start, end), qs)]
                V = sum(cen(y, qq) * (Qd // qq) for y, qq in zip(ys, qs))
                lifted_seen += sum(y > qq // 2 for y, qq in zip(ys, qs))
                X, _ = crt([xc[i, n] for i in range(start, end)], qs)
                assert (V - X) % Qd == 0 and abs((V - X) // Qd) <= len(qs)
                for pos in range(ext):
                    if start <= pos < end:
                        continue
                    assert int(blkc[pos, n]) == V % c.q(pids[pos])
        assert lifted_seen > 0
        pe = c.ext_pids(l)
        xe = c.random_poly(60 + l, pe)
        out = c.moddown(xe, l)
        xec = c.intt(xe, pe)
        oc = c.intt(out, list(range(l + 1)))
        ps = [c.q(L1 + k) for k in range(K)]
        P = 1
        for p in ps:
            P *= p
        for n in (1, 7, N - 2):
            V = sum(cen(int(xec[l + 1 + k, n]) * pow(P // p, -1, p) % p, p) * (P // p) for k, p in enumerate(ps))
            for i in range(l + 1):
                qi = c.q(i)
                assert int(oc[i, n]) == (int(xec[i, n]) - V) * pow(P, -1, qi) % qi
    # switched off again: the plain conversion
    c.set_reduced_noise(False)
    x = c.random_poly(41, list(range(4)))
    assert np.array_equal(c.modup(x, 3), small_ctx().modup(x, 3))


def test_rescale_closed_form(sc):
    c = sc
    L1 = c.L1
    for l in (3, 2, 1):
        # OpenFHE QlQlInvModqlDivqlModq: floor([Q_{l-1}^-1]_{q_l} Q_{l-1} / q_l) mod q_i
        Q = 1
        for i in range(l):
            Q *= c.q(i)
        R = (pow(Q, -1, c.q(l)) * Q) // c.q(l)
        for i in range(l):
            assert int(O.lib().orc_rescale_QlQlInv(c.h, l, i)) == R % c.q(i)
            assert int(O.lib().orc_rescale_ql_inv(c.h, l, i)) == pow(c.q(l), -1, c.q(i))
        x = c.random_poly(30 + l, list(range(l + 1)))
        out = c.rescale_poly(x, l)
        xc = c.intt(x, list(range(l + 1)))
        oc = c.intt(out, list(range(l)))
        ql = c.q(l)
        for n in (0, 1, c.N - 1, 100):
            X, M = crt([xc[i, n] for i in range(l + 1)], [c.q(i) for i in range(l + 1)])
            r = X % ql
            if r > ql // 2:
                r -= ql
            Y = (X - r) // ql
            for i in range(l):
                assert int(oc[i, n]) == Y % c.q(i)


def test_automorph_coefficient_definition(sc):
    c = sc
    N = c.N
    a = np.random.default_rng(4).integers(-5, 6, N)
    q = c.q(0)
    A = c.ntt(np.array([a % q], dtype=np.uint64), [0])
    for k in (5, c.rot_index(3), c.rot_index(-2), 2 * N - 1):
        got = c.intt(c.automorph(A, k), [0])[0]
        exp = [0] * N
        for i in range(N):
            e = i * k % (2 * N)
            if e < N:
                exp[e] += int(a[i])
            else:
                exp[e - N] -= int(a[i])
        assert [int(t) for t in got] == [t % q for t in exp]
    assert c.rot_index(0) == 1
    assert c.rot_index(1) == 5 and c.rot_index(2) == 25
    assert c.rot_index(-1) * 5 % (2 * N) == 1


def _slots(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def test_encode_decode_encrypt_roundtrip(sc):
    c = sc
    n = c.N // 2
    z = _slots(n, 1)
    l = c.L1 - 1
    m = c.encode(z, 2.0 ** 40, l)
    assert np.max(np.abs(c.decode(m, l, n, 2.0 ** 40) - z)) < 1e-9
    s = c.gen_secret(3)
    ct = c.encrypt(m, l, s, 4)
    assert np.max(np.abs(c.decode(c.decrypt(ct, l, s), l, n, 2.0 ** 40) - z)) < 1e-8
    # sparse packing
    z8 = _slots(8, 2)
    m8 = c.encode(z8, 2.0 ** 40, l)
    assert np.max(np.abs(c.decode(m8, l, 8, 2.0 ** 40) - z8)) < 1e-9


def test_keyswitch_decrypt_invariants():
    """relinearized product and hoisted rotations decrypt to the expected slots (cfg1 shape)."""
    c = small_ctx(13, 4, 2, 2)
    n, l, Dl = c.N // 2, 3, 2.0 ** 50
    s = c.gen_secret(7)
    z1, z2 = _slots(n, 8), _slots(n, 9)
    ct1 = c.encrypt(c.encode(z1, Dl, l), l, s, 10)
    ct2 = c.encrypt(c.encode(z2, Dl, l), l, s, 11)
    rk = c.gen_relin_key(s, 12)
    prod = c.eval_mult(ct1, ct2, l, rk)
    dec = c.decode(c.decrypt(prod, l, s), l, n, Dl * Dl)
    assert np.max(np.abs(dec - z1 * z2)) < 1e-9
    # the degree-2 tensor decrypts the same (relinearization only adds key-switch noise)
    t = c.tensor(ct1, ct2, l)
    dec3 = c.decode(c.decrypt(t, l, s, 3), l, n, Dl * Dl)
    assert np.max(np.abs(dec3 - z1 * z2)) < 1e-9
    rs = c.rescale_ct(prod, l)
    dec = c.decode(c.decrypt(rs, l - 1, s), l - 1, n, Dl * Dl / float(c.q(l)))
    assert np.max(np.abs(dec - z1 * z2)) < 1e-7
    for r in (1, -3, 17):
        k = c.rot_index(r)
        key = c.gen_rot_key(s, k, 100 + r)
        rot = c.eval_rotate(ct1, l, key, k)
        dec = c.decode(c.decrypt(rot, l, s), l, n, Dl)
        assert np.max(np.abs(dec - np.roll(z1, -r))) < 1e-9
    # conjugation
    kc = 2 * c.N - 1
    key = c.gen_rot_key(s, kc, 200)
    conj = c.eval_rotate(ct1, l, key, kc)
    dec = c.decode(c.decrypt(conj, l, s), l, n, Dl)
    assert np.max(np.abs(dec - np.conj(z1))) < 1e-9


def _lt_case(c, s, z, l, rowSize, bStep, stride, offset, seed=0, scale_pt=2.0 ** 30):
    n = c.N // 2
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (rowSize, n)) + 1j * rng.uniform(-1, 1, (rowSize, n))
    diags = np.concatenate([c.encode_ext(np.roll(D[k], bStep * (k // bStep) * stride + offset), scale_pt, l) for k in range(rowSize)])
    bk = [None] + [c.gen_rot_key(s, c.rot_index(i * stride), 500 + i) for i in range(1, bStep)]
    bkk = [1] + [c.rot_index(i * stride) for i in range(1, bStep)]
    gk = c.gen_rot_key(s, c.rot_index(bStep * stride), 600)
    ok = c.gen_rot_key(s, c.rot_index(offset), 700) if offset else None
    want = sum(np.roll(z, -(k * stride + offset)) * D[k] for k in range(rowSize))
    return diags, bk, bkk, gk, ok, want


@pytest.mark.parametrize("rowSize,bStep,offset", [(12, 4, 0), (10, 4, 3), (8, 8, 0)])
def test_linear_transform_decrypts(rowSize, bStep, offset):
    """BSGS linear transform (extended-basis formulation) decrypts to sum_k rot(z, k + offset) * D_k."""
    c = small_ctx(10, 4, 2, 2)
    n, l, Dl = c.N // 2, 3, 2.0 ** 50
    s = c.gen_secret(17)
    z = _slots(n, 18)
    ct = c.encrypt(c.encode(z, Dl, l), l, s, 19)
    diags, bk, bkk, gk, ok, want = _lt_case(c, s, z, l, rowSize, bStep, 1, offset)
    out = c.linear_transform(ct, l, diags, bStep, 1, bk, bkk, gk, c.rot_index(bStep), ok, c.rot_index(offset) if offset else 1, offset)
    dec = c.decode(c.decrypt(out, l, s), l, n, Dl * 2.0 ** 30)
    assert np.max(np.abs(dec - want)) < 1e-6


@pytest.mark.parametrize("rowSize,bStep,offset", [(12, 4, 0), (10, 4, 3), (16, 4, 0)])
def test_linear_transform_flat_decrypts(rowSize, bStep, offset):
    """The flat form (each giant-step partial sum rotated by its own multiple of the giant step)
    decrypts to the same sum_k rot(z, k + offset) D_k as the Horner form, with different residues;
    a giant key for the wrong multiple (every R_j rotated by the one giant step, as if it were
    Horner's) misses the map by far."""
    c = small_ctx(10, 4, 2, 2)
    n, l, Dl = c.N // 2, 3, 2.0 ** 50
    s = c.gen_secret(17)
    z = _slots(n, 18)
    ct = c.encrypt(c.encode(z, Dl, l), l, s, 19)
    diags, bk, bkk, gk, ok, want = _lt_case(c, s, z, l, rowSize, bStep, 1, offset)
    gStep = (rowSize + bStep - 1) // bStep
    gks = [c.rot_index(j * bStep) for j in range(1, gStep)]
    gkeys = [c.gen_rot_key(s, k, 800 + j) for j, k in enumerate(gks)]
    okk = c.rot_index(offset) if offset else 1
    flat = c.linear_transform_flat(ct, l, diags, bStep, 1, bk, bkk, gkeys, gks, ok, okk, offset)
    horner = c.linear_transform(ct, l, diags, bStep, 1, bk, bkk, gk, c.rot_index(bStep), ok, okk, offset)
    dec = c.decode(c.decrypt(flat, l, s), l, n, Dl * 2.0 ** 30)
    assert np.max(np.abs(dec - want)) < 1e-6
    assert not np.array_equal(flat, horner)
    wrong = c.linear_transform_flat(ct, l, diags, bStep, 1, bk, bkk, [gk] * (gStep - 1), [c.rot_index(bStep)] * (gStep - 1), ok, okk,
                                    offset)
    assert np.max(np.abs(c.decode(c.decrypt(wrong, l, s), l, n, Dl * 2.0 ** 30) - want)) > 1e-2


def test_public_key_encryption_decrypts(sc):
    """pk = (b, a) = an encryption of zero; c0 + c1 s = v e + e0 + e1 s + m, well inside the noise."""
    c = sc
    n, l, Dl = c.N // 2, c.L1 - 1, 2.0 ** 40
    s = c.gen_secret(21)
    pk = c.encrypt(c.zeros(c.L1), c.L1 - 1, s, 22)
    z = _slots(n, 23)
    ct = c.encrypt_pk(c.encode(z, Dl, l), l, pk[:c.L1], pk[c.L1:], 24)
    assert np.max(np.abs(c.decode(c.decrypt(ct, l, s), l, n, Dl) - z)) < 1e-7
    # the same message at a lower level uses the key's first towers
    ct2 = c.encrypt_pk(c.encode(z, Dl, 1), 1, pk[:c.L1], pk[c.L1:], 25)
    assert np.max(np.abs(c.decode(c.decrypt(ct2, 1, s), 1, n, Dl) - z)) < 1e-7
    # a dropped noise or message term is visible: the ciphertext is not an encryption of zero
    assert np.max(np.abs(c.decode(c.decrypt(ct, l, s), l, n, Dl))) > 0.5


def _crt_centered(res, moduli):
    """the centred integer with the given residues (Python big integers)"""
    Q = 1
    for q in moduli:
        Q *= q
    x = 0
    for r, q in zip(res, moduli):
        Qi = Q // q
        x += int(r) * Qi * pow(Qi, -1, q)
    x %= Q
    return x - Q if x > Q // 2 else x


def _round_like_double(fr):
    """trunc(fl(fr + 1/2)) for an exact rational fr: the reference's `(int)(x + 0.5)` on doubles"""
    from fractions import Fraction
    return int(float(fr + Fraction(1, 2)))


@pytest.mark.parametrize("operand", [2.5, -0.75, 1e-9, 0.1, -1234.5678, 3.0e30, -7.25e25, 0.0, 2.0 ** 70])
def test_elem_for_eval_mult_closed_form(sc, operand):
    """ElemForEvalMult: the residues are those of one integer V (CRT over all towers) with
    V = round(operand * scale) when that fits in 125 bits, else V = 2^k * round(operand*scale/2^k)
    with k = ceil(log2|operand*scale|) - 125: divisible by 2^k and within 2^k of operand*scale."""
    from fractions import Fraction
    q = [c_ for c_ in (sc.q(i) for i in range(sc.L1))]
    scale = 2.0 ** 50
    V = _crt_centered(O.elem_for_eval_mult(q, operand, scale), q)
    x = Fraction(operand) * Fraction(scale)
    if abs(x) < Fraction(2) ** 125:
        assert V == _round_like_double(x)
    else:
        import math
        k = math.ceil(math.log2(abs(float(x)))) - 125
        assert k > 0 and V % (1 << k) == 0 and abs(V - x) <= (1 << k)


@pytest.mark.parametrize("operand,deg", [(v, d) for v in (0.25, 1.5, 0.0, 1e-7, 1e12) for d in (1, 2, 3)] + [(3.3e20, 1), (3.3e20, 2)])
def test_elem_for_eval_add_closed_form(sc, operand, deg):
    """ElemForEvalAddOrSub: V = round(operand*scale) * round(scale)^(deg-1) below 2^61; above,
    round(operand*scale) is taken at 2^k granularity, k = ceil(log2(operand*scale)) - 61."""
    from fractions import Fraction
    import math
    q = [sc.q(i) for i in range(sc.L1)]
    scale = 2.0 ** 50
    x = Fraction(operand) * Fraction(scale)
    sf = (1 << 50) ** (deg - 1)
    assert x * sf * 4 < math.prod(q)  # the residues determine the integer
    V = _crt_centered(O.elem_for_eval_add(q, operand, scale, deg), q)
    assert V % sf == 0
    W = V // sf
    if x < Fraction(2) ** 61:
        assert W == _round_like_double(x)
    else:
        k = math.ceil(math.log2(float(x))) - 61
        assert k > 0 and W % (1 << k) == 0 and abs(W - x) <= (1 << k)


def test_mod_raise_closed_form(sc):
    """Bootstrapping ModRaise: after the raise, tower i of every component holds the centred lift of
    tower 0's coefficient (v - q_0 when v > q_0 / 2) mod q_i, tower 0 is untouched, and an
    encryption of m at level 0 decrypts at the top level to m + q_0 I with small integers I."""
    c = sc
    top = c.L1 - 1
    q0 = c.q(0)
    for l_in in (0, 2):
        ct = np.concatenate([c.random_poly(70 + l_in + k, list(range(l_in + 1))) for k in range(2)])
        out = c.mod_raise(ct, l_in, top)
        for k in range(2):
            x0 = c.intt(ct[k * (l_in + 1):k * (l_in + 1) + 1], [0])[0]
            oc = c.intt(out[k * (top + 1):(k + 1) * (top + 1)], list(range(top + 1)))
            assert np.array_equal(out[k * (top + 1)], ct[k * (l_in + 1)])
            for n in (0, 1, 5, c.N - 1):
                v = int(x0[n])
                cen = v - q0 if v > q0 // 2 else v
                for i in range(top + 1):
                    assert int(oc[i, n]) == cen % c.q(i), (l_in, k, n, i)
    # decryption view: m + q0 * I, |I| small (sparse secret of weight 16)
    s = c.gen_secret(5, 16)
    z = _slots(8, 3)
    m = c.encode(z, 2.0 ** 40, 0)
    ct0 = c.encrypt(m, 0, s, 9)
    up = c.mod_raise(ct0, 0, top)
    dec = c.intt(c.decrypt(up, top, s), list(range(top + 1)))
    mc = c.intt(c.decrypt(ct0, 0, s), [0])[0]
    for n in range(0, c.N, 37):
        X, M = crt([int(dec[i, n]) for i in range(top + 1)], [c.q(i) for i in range(top + 1)])
        if X > M // 2:
            X -= M
        v = int(mc[n])
        v = v - q0 if v > q0 // 2 else v
        assert (X - v) % q0 == 0
        assert abs((X - v) // q0) <= 16


def test_context_rejects_inconsistent_prime_count():
    primes = O.gen_primes(10, 4, 60, 50, 2)
    with pytest.raises(ValueError):
        O.Ctx(10, primes, 4, 3, 2)   # 6 primes cannot be L1=4 plus K=3


def test_accumulate_sum_invariants():
    """Both rotation-fold sums (the lazy fold and the cascade from a start factor) decrypt to
    sum_k roll(z, -k*start*stride) over the fold's offsets, and a one-level fold equals
    ModDown(P*ct + the extended rotations) in both forms: one ModDown of the sum, not one per rotation."""
    c = small_ctx(13, 4, 2, 2)
    n, l, Dl = c.N // 2, 3, 2.0 ** 50
    s = c.gen_secret(21)
    z = _slots(n, 22)
    ct = c.encrypt(c.encode(z, Dl, l), l, s, 23)
    cases = [((4, 1, 16, 1), range(16)), ((2, 3, 8, 1), range(0, 24, 3)), ((4, 2, 8, 1), range(0, 16, 2)),
             ((4, 1, 16, 2), range(0, 16, 2)), ((8, 1, 8, 1), range(8)), ((4, 1, 9, 1), range(12))]
    for (radix, stride, size, start), offsets in cases:
        idx = c.accumulate_indices(radix, stride, size, start)
        ks = [c.rot_index(r) for r in idx]
        keys = [c.gen_rot_key(s, k, 300 + i) for i, k in enumerate(ks)]
        outs = [c.accumulate_sum_from(ct, l, radix, stride, size, keys, ks, start)]
        if start == 1:
            outs.append(c.accumulate_sum(ct, l, radix, stride, size, keys, ks))
        exp = sum(np.roll(z, -k) for k in offsets)
        for out in outs:
            dec = c.decode(c.decrypt(out, l, s), l, n, Dl)
            assert np.max(np.abs(dec - exp)) < 1e-7, (radix, stride, size, start)
    assert c.accumulate_indices(4, 1, 16) == [1, 2, 3, 4, 8, 12]
    assert c.accumulate_indices(4, 1, 9) == [1, 2, 3, 4, 8]
    assert c.accumulate_indices(4, 2, 16, 2) == [4, 8, 12, 16]
    # one fold level by composition: P*ct over Q_l (0 over P) plus the extended rotations, one ModDown
    idx = c.accumulate_indices(4, 1, 4)
    ks = [c.rot_index(r) for r in idx]
    keys = [c.gen_rot_key(s, k, 400 + i) for i, k in enumerate(ks)]
    ext = l + 1 + c.K
    acc = _p_scaled(c, ct, l)
    for k, key in zip(ks, keys):
        acc = _add_ext(c, l, acc, c.eval_rotate_ext(ct, l, key, k))
    lazy = np.concatenate([c.moddown(acc[:ext], l), c.moddown(acc[ext:], l)])
    assert np.array_equal(c.accumulate_sum(ct, l, 4, 1, 4, keys, ks), lazy)
    assert np.array_equal(c.accumulate_sum_from(ct, l, 4, 1, 4, keys, ks, 1), lazy)
    eager = ct.copy()
    for k, key in zip(ks, keys):
        eager = c.add(eager, c.eval_rotate(ct, l, key, k))
    assert not np.array_equal(eager, lazy)       # the per-rotation ModDown rounds differently


def _p_scaled(c, ct, l, comps=(0, 1)):
    """(P*c_comp over Q_l, 0 over P) for each listed component, stacked."""
    ext = l + 1 + c.K
    P = 1
    for k in range(c.K):
        P *= c.q(c.L1 + k)
    acc = np.zeros((len(comps) * ext, c.N), dtype=np.uint64)
    for r, comp in enumerate(comps):
        for p in range(l + 1):
            q = c.q(p)
            acc[r * ext + p] = np.array([int(v) * (P % q) % q for v in ct[comp * (l + 1) + p]], dtype=np.uint64)
    return acc


def _add_ext(c, l, a, b):
    """a + b over stacked polynomials in Q_l ∪ P (exact, through Python integers)."""
    ext = l + 1 + c.K
    pids = c.ext_pids(l)
    out = np.empty_like(a)
    for row in range(a.shape[0]):
        q = c.q(pids[row % ext])
        out[row] = np.array([(int(x) + int(y)) % q for x, y in zip(a[row], b[row])], dtype=np.uint64)
    return out


@pytest.mark.parametrize("radix,size,terms", [(4, 16, 16), (2, 8, 8), (4, 9, 12)])
def test_accumulate_lazy_composed(radix, size, terms):
    """The lazy fold over two or more levels, composed from pinned pieces (the extended rotation of c1
    alone, the coefficient-domain automorphism, ModDown): c0 enters as P*c0 and then stays in Q_l∪P,
    where every later rotation adds the automorphism of the extended c0 itself; c1 takes a ModDown per
    level, c0 one at the end. It decrypts to the sum and differs in bits from the cascade, which takes
    a ModDown of c0 at every level."""
    c = small_ctx(12, 3, 2, 2)
    n, l, Dl = c.N // 2, 2, 2.0 ** 50
    s = c.gen_secret(31)
    z = _slots(n, 32)
    ct = c.encrypt(c.encode(z, Dl, l), l, s, 33)
    ext, Q = l + 1 + c.K, l + 1
    idx = c.accumulate_indices(radix, 1, size)
    ks = [c.rot_index(r) for r in idx]
    keys = [c.gen_rot_key(s, k, 600 + i) for i, k in enumerate(ks)]
    c0e, c1 = None, ct[Q:].copy()
    used, lev = 0, 1
    while lev < size:
        cur = np.concatenate([ct[:Q] if c0e is None else np.zeros((Q, c.N), dtype=np.uint64), c1])
        a0 = _p_scaled(c, cur, l, (0,)) if c0e is None else c0e.copy()
        a1 = _p_scaled(c, cur, l, (1,))
        i = 1
        while i < radix and i * lev < size:
            k, key = ks[used], keys[used]
            if c0e is None:
                r = c.eval_rotate_ext(cur, l, key, k)                  # P*c0 folded before the automorphism
            else:
                r = c.eval_rotate_ext(cur, l, key, k, add_c0=False)   # the key switch of c1 alone ...
                r[:ext] = _add_ext(c, l, r[:ext], c.automorph(c0e, k))  # ... plus the automorphism of c0e
            a0 = _add_ext(c, l, a0, r[:ext])
            a1 = _add_ext(c, l, a1, r[ext:])
            used, i = used + 1, i + 1
        c1 = c.moddown(a1, l)
        c0e = a0
        lev *= radix
    composed = np.concatenate([c.moddown(c0e, l), c1])
    lazy = c.accumulate_sum(ct, l, radix, 1, size, keys, ks)
    assert np.array_equal(lazy, composed)
    cascade = c.accumulate_sum_from(ct, l, radix, 1, size, keys, ks, 1)
    assert np.array_equal(lazy[Q:], cascade[Q:])          # c1 never depends on c0
    assert not np.array_equal(lazy[:Q], cascade[:Q])      # c0 rounds once instead of once per level
    dec = c.decode(c.decrypt(lazy, l, s), l, n, Dl)
    assert np.max(np.abs(dec - sum(np.roll(z, -k) for k in range(terms)))) < 1e-7
