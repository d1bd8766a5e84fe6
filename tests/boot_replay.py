"""The facade's CKKS bootstrap and Chebyshev series evaluation restated over the oracle's
operations (test infrastructure): the same sequence of ModRaise, rotations, linear transforms,
products, Rescales, scalar encodings and level reductions as
fideslib-fork_amd/host/bootstrap.cpp, so that a GPU run and this replay can be compared residue by
residue. Also computes the bootstrap constants the way the facade does (Chebyshev coefficients of
the shifted cosine, the special-FFT matrices) for CPU-only precision checks.
"""
import math

import numpy as np

from oracle import oracle as O



def _clog2(d):
    j = 0
    while (1 << j) < d:
        j += 1
    return j


def cheb_split_depth(d):
    """Levels a piece of degree d >= 1 consumes below T_1's when split at the largest M <= d."""
    M = 1
    while 2 * M <= d:
        M *= 2
    lm = _clog2(M)
    prod = (lm if d == M else max(cheb_depth(d - M), lm)) + 1
    return max(prod, cheb_depth(M - 1))


def cheb_leaf(d):
    """A piece of degree below 8, or below 16 where that costs no extra level, is a linear
    combination of T_1 .. T_d."""
    return d < 8 or (d < 16 and _clog2(d) + 1 <= cheb_split_depth(d))


def cheb_depth(d):
    return _clog2(d) + 1 if cheb_leaf(d) else cheb_split_depth(d)


# ---- constants (mirrors of the facade's double arithmetic) ----
def cheb_powers(d, need):
    """The T_k a series of degree d reads: its leaves' T_1..T_d and its split powers."""
    if cheb_leaf(d):
        need.update(range(1, d + 1))
        return
    M = 1
    while 2 * M <= d:
        M *= 2
    need.add(M)
    if d > M:
        cheb_powers(d - M, need)
    cheb_powers(M - 1, need)


def cheb_products(d):
    """Ciphertext products of a series of degree d: its powers T_k (k >= 2, with every T_k they
    need) and its split pieces' products."""
    need, seen = set(), set()
    cheb_powers(d, need)
    stack = list(need)
    while stack:
        k = stack.pop()
        if k <= 1 or k in seen:
            continue
        seen.add(k)
        stack += [k // 2, k // 2 + k % 2]

    def tree(d):
        if cheb_leaf(d):
            return 0
        M = 1
        while 2 * M <= d:
            M *= 2
        return (1 + tree(d - M) if d > M else 0) + tree(M - 1)

    return len(seen) + tree(d)


def cheb_coefficients(f, a, b, degree):
    n = degree + 1
    fx = [f(0.5 * (b - a) * math.cos(math.pi * (j + 0.5) / n) + 0.5 * (b + a)) for j in range(n)]
    c = []
    for k in range(n):
        s = 0.0
        for j in range(n):
            s += fx[j] * math.cos(math.pi * k * (j + 0.5) / n)
        c.append(2.0 * s / n)
    return c


def _bitrev(v):
    n = len(v)
    j = 0
    for i in range(1, n):
        bit = n >> 1
        while j & bit:
            j ^= bit
            bit >>= 1
        j ^= bit
        if i < j:
            v[i], v[j] = v[j], v[i]


def special_fft(v, M):
    v = list(v)
    n = len(v)
    _bitrev(v)
    ln = 2
    while ln <= n:
        lh, lq = ln >> 1, ln << 2
        for i in range(0, n, ln):
            g = 1
            for j in range(lh):
                ang = 2.0 * math.pi * ((g % lq) * (M // lq)) / M
                wr, wi = math.cos(ang), math.sin(ang)
                a, b = v[i + j], v[i + j + lh]
                tr, ti = b[0] * wr - b[1] * wi, b[0] * wi + b[1] * wr
                v[i + j] = (a[0] + tr, a[1] + ti)
                v[i + j + lh] = (a[0] - tr, a[1] - ti)
                g = (g * 5) % M
        ln <<= 1
    return v


def special_ifft(v, M):
    v = list(v)
    n = len(v)
    ln = n
    while ln >= 1:
        lh, lq = ln >> 1, ln << 2
        for i in range(0, n, ln):
            g = 1
            for j in range(lh):
                ang = 2.0 * math.pi * ((lq - (g % lq)) * (M // lq)) / M
                wr, wi = math.cos(ang), math.sin(ang)
                a, b = v[i + j], v[i + j + lh]
                ur, ui = a[0] + b[0], a[1] + b[1]
                dr, di = a[0] - b[0], a[1] - b[1]
                v[i + j] = (ur, ui)
                v[i + j + lh] = (dr * wr - di * wi, dr * wi + di * wr)
                g = (g * 5) % M
        ln >>= 1
    _bitrev(v)
    return [(x[0] / n, x[1] / n) for x in v]


def matrix_of(n, M, inverse):
    A = np.zeros((n, n), dtype=np.complex128)
    for k in range(n):
        e = [(0.0, 0.0)] * n
        e[k] = (1.0, 0.0)
        col = special_ifft(e, M) if inverse else special_fft(e, M)
        for j in range(n):
            A[j, k] = complex(col[j][0], col[j][1])
    return A


# ---- the special FFT factored into butterfly stages (level budgets > 1) ----
def _put(D, d, v):
    D[d] = D[d] + v if d in D else v


def fft_stages(n, M, inverse):
    """Butterfly stages of special_fft after its bit reversal in diagonal form {offset: vector}
    (row r of the stage reads column r + offset); inverse: the inverse of each stage."""
    out = []
    ln = 2
    while ln <= n:
        lh, lq = ln >> 1, ln << 2
        d0, dp, dm = np.zeros(n, complex), np.zeros(n, complex), np.zeros(n, complex)
        for i in range(0, n, ln):
            g = 1
            for j in range(lh):
                ang = 2.0 * math.pi * ((g % lq) * (M // lq)) / M
                w = complex(math.cos(ang), math.sin(ang))
                p, q = i + j, i + j + lh
                if inverse:   # x_p = (y_p + y_q) / 2, x_q = (y_p - y_q) / (2 w)
                    d0[p], dp[p] = 0.5, 0.5
                    dm[q], d0[q] = 0.5 / w, -0.5 / w
                else:         # y_p = x_p + w x_q, y_q = x_p - w x_q
                    d0[p], dp[p] = 1.0, w
                    dm[q], d0[q] = 1.0, -w
                g = (g * 5) % M
        st = {0: d0}
        _put(st, lh % n, dp)
        _put(st, (-lh) % n, dm)
        out.append(st)
        ln <<= 1
    return out


def compose(A, Bm, n):
    """A B in diagonal form: offset d1 + d2, value A_d1[r] B_d2[r + d1]."""
    C = {}
    for d1 in sorted(A):
        for d2 in sorted(Bm):
            _put(C, (d1 + d2) % n, A[d1] * np.roll(Bm[d2], -d1))
    return C


def split_levels(L, b):
    """stage counts of b groups (the first groups take the extra stages)"""
    b = max(1, min(b, L))
    return [L // b + (1 if i < L % b else 0) for i in range(b)]


def group_transform(D, n, low):
    """LinearTransform shape of a group whose smallest butterfly half-length is `low`:
    (rowSize, stride, offset, diagonals d_k for the rotations offset + k stride)."""
    span = max(abs(((d + n // 2) % n) - n // 2) for d in D) // low   # largest |offset| / stride
    if 2 * span + 1 >= n // low:
        rowSize, offset = n // low, 0
    else:
        rowSize, offset = 2 * span + 1, (-span * low) % n
    diags = []
    for k in range(rowSize):
        d = (offset + k * low) % n
        diags.append(D.get(d, np.zeros(n, complex)))
    return rowSize, low, offset, diags


def factored(n, M, budget, inverse, factor):
    """Groups of the factored CoeffsToSlots (inverse: S_1^-1 .. S_L^-1, applied from S_L^-1) or
    SlotsToCoeffs (S_L .. S_1, applied from S_1), in the order they are applied; `factor` scales
    the first group."""
    S = fft_stages(n, M, inverse)
    L = len(S)
    order = list(range(L - 1, -1, -1)) if inverse else list(range(L))
    groups = []
    pos = 0
    for m in split_levels(L, budget):
        idx = order[pos:pos + m]
        pos += m
        P = {0: np.ones(n, complex)}
        for i in idx:
            P = compose(S[i], P, n)
        if not groups:
            P = {d: v * factor for d, v in P.items()}
        groups.append(group_transform(P, n, 1 << min(idx)))
    return groups


def bstep_for(rowSize):
    """baby steps of a bootstrap transform: the power of two up to 16 minimizing
    (b - 1) + 4 (ceil(rowSize / b) - 1), a giant step (ModDown, ModUp and a key inner product in
    the flat form's batch) weighted as four hoisted baby steps (a key inner product each);
    ties go to the smaller b"""
    best = None
    for b in (1, 2, 4, 8, 16):
        cost = (b - 1) + 4 * (-(-rowSize // b) - 1)
        if best is None or cost < best[0]:
            best = (cost, b)
    return best[1]


TRACE_RADIX = 4


def trace_indices(n, N):
    """rotations of the partial trace: the lazy fold at radix 4 over the N/(2n) copies at stride n"""
    if n >= N // 2:
        return []
    out, s, size = [], 1, N // 2 // n
    while s < size:
        i = 1
        while i < TRACE_RADIX and i * s < size:
            out.append(i * s * n)
            i += 1
        s *= TRACE_RADIX
    return out


def rotation_steps(k, N):
    """slot rotations a bootstrap with constants k uses: the partial trace, then every group's
    baby steps, giant-step multiples (each rotated by its own key) and offset"""
    n = k["slots"]
    steps = list(trace_indices(n, N))
    for rowSize, stride, offset, _ in k["cts"] + k["stc"]:
        b = bstep_for(rowSize)
        steps += [(i * stride) % n for i in range(1, b)]
        steps += [(j * b * stride) % n for j in range(1, (rowSize + b - 1) // b) if (j * b * stride) % n]
        if offset:
            steps.append(offset)
    return [r for r in steps if r % (N // 2)]


def constants(N, n, hw, scale, q0, sparse=True, budget=(1, 1), correction=0, top_primes=None):
    """top_primes: the chain's primes from the top down; the raise's rescales (the prescale's and
    one per CoeffsToSlots group) divide by them, and the first group carries prod(q_r / scale)."""
    gap = (N // 2) // n
    K = min(math.ceil(hw / 2.0) + 4.0, 28.0) if sparse else 64.0
    r, degree = 5, 63
    if sparse:
        # (r, degree) with 2 pi K / 2^r within what the degree approximates: the lowest depth,
        # then the fewest products
        best = None
        for rr in range(11):
            for deg, lim in ((63, 32.0), (127, 80.0), (255, 160.0)):
                if 2.0 * math.pi * K / math.ldexp(1.0, rr) > lim:
                    continue
                key = (cheb_depth(deg) + rr, cheb_products(deg) + rr)
                if best is None or key < best[0]:
                    best = (key, rr, deg)
        r, degree = best[1], best[2]
    cheb = cheb_coefficients(lambda y: math.cos((2.0 * math.pi * K * y - 0.5 * math.pi) / math.ldexp(1.0, r)), -1.0, 1.0, degree)
    M = 2 * N
    fc = scale / (2.0 * gap * q0 * K)
    # an exact power of two 2^-e <= fc is applied as a scalar product before the trace; the
    # diagonals carry fc 2^e in [1, 2)
    e = math.ceil(-math.log2(fc))
    fs = q0 / (2.0 * math.pi * scale)
    ncts = max(1, min(budget[0], n.bit_length() - 1))
    fix = 1.0
    if top_primes is not None:
        for j in range(ncts + 1):
            fix *= float(top_primes[j]) / scale
    return dict(slots=n, gap=gap, K=K, r=r, cheb=cheb, pre=math.ldexp(1.0, -e), correction=correction, scale_fix=fix,
                cts=factored(n, M, budget[0], True, fc * math.ldexp(1.0, e) * fix), stc=factored(n, M, budget[1], False, fs))


# ---- ciphertext operations with the facade's metadata rules (towers t, noise degree) ----
class Ops:
    def __init__(self, o, scale, relin, rot):
        self.o, self.scale, self.relin, self.rot = o, scale, relin, rot
        self.q = [int(o.q(i)) for i in range(o.L1)]
        self.N = o.N
        self._mono = {}

    @staticmethod
    def comp(ct, t, i):
        return ct[i * t:(i + 1) * t]

    def reduce(self, c, t_to):
        x, t, d = c
        if t == t_to:
            return c
        assert t > t_to
        return (np.concatenate([x[:t_to], x[t:t + t_to]]), t_to, d)

    def add(self, a, b):
        assert a[2] == b[2]
        t = min(a[1], b[1])
        a, b = self.reduce(a, t), self.reduce(b, t)
        o = self.o
        return (np.concatenate([o.add(self.comp(a[0], t, i), self.comp(b[0], t, i)) for i in range(2)]), t, a[2])

    def neg(self, a):
        t = a[1]
        return (np.concatenate([self.o.neg(self.comp(a[0], t, i)) for i in range(2)]), t, a[2])

    def sub(self, a, b):
        return self.add(a, self.neg(b))

    def add_scalar(self, a, c):
        x, t, d = a
        e = O.elem_for_eval_add(self.q[:t], abs(c), self.scale, d)
        if c < 0:
            e = np.array([(qi - int(v)) % qi for v, qi in zip(e, self.q[:t])], dtype=np.uint64)
        return (np.concatenate([self.o.add_scalar(self.comp(x, t, 0), e), self.comp(x, t, 1)]), t, d)

    def mul_scalar(self, a, c):
        x, t, d = a
        f = O.elem_for_eval_mult(self.q[:t], c, self.scale)
        return (np.concatenate([self.o.mul_scalar(self.comp(x, t, i), f) for i in range(2)]), t, d + 1)

    def mult(self, a, b):
        t = min(a[1], b[1])
        a, b = self.reduce(a, t), self.reduce(b, t)
        return (self.o.eval_mult(a[0], b[0], t - 1, self.relin), t, a[2] + b[2])

    def mul_integer(self, a, v):
        x, t, d = a
        f = np.array([v % qi for qi in self.q[:t]], dtype=np.uint64)
        return (np.concatenate([self.o.mul_scalar(self.comp(x, t, i), f) for i in range(2)]), t, d)

    def rescale(self, a):
        x, t, d = a
        return (self.o.rescale_ct(x, t - 1), t - 1, max(1, d - 1))

    def accumulate(self, a, stride, size):
        """the lazy rotation-fold sum at radix 4 (metadata unchanged)"""
        x, t, d = a
        ks = [self.o.rot_index(r) for r in trace_indices(stride, 2 * stride * size)]
        return (self.o.accumulate_sum(x, t - 1, TRACE_RADIX, stride, size, [self.rot[k] for k in ks], ks), t, d)

    def rotate(self, a, k):
        x, t, d = a
        return (self.o.eval_rotate(x, t - 1, self.rot[k], k), t, d)

    def monomial(self, a, power):
        x, t, d = a
        N = self.N
        power %= 2 * N
        key = (power, t)
        if key not in self._mono:
            m = np.zeros((t, N), dtype=np.uint64)
            for i in range(t):
                m[i, power % N] = 1 if power < N else self.q[i] - 1
            self._mono[key] = self.o.ntt(m, list(range(t)))
        mm = self._mono[key]
        return (np.concatenate([self.o.mul(self.comp(x, t, i), mm) for i in range(2)]), t, d)

    def linear_transform(self, a, group, n):
        """LinearTransform over `n` slots of a (rowSize, stride, offset, diagonals) group: the
        diagonals are pre-rotated by -(bStep floor(k / bStep) stride + offset) and encoded
        extended at the ciphertext's level."""
        x, t, d = a
        o = self.o
        l = t - 1
        rowSize, stride, offset, dv = group
        bStep = bstep_for(rowSize)
        pre = [np.roll(dv[k], (bStep * (k // bStep) * stride + offset) % n) for k in range(rowSize)]
        diags = np.concatenate([o.encode_ext(v, self.scale, l) for v in pre])
        baby_keys, baby_k = [None] * bStep, [1] * bStep
        for i in range(1, bStep):
            if (i * stride) % n == 0:
                continue
            g = o.rot_index(i * stride)
            baby_keys[i], baby_k[i] = self.rot[g], g
        gStep = (rowSize + bStep - 1) // bStep
        giant_keys, giant_k = [None] * (gStep - 1), [1] * (gStep - 1)
        for j in range(1, gStep):
            r = (j * bStep * stride) % n
            if r:
                giant_k[j - 1] = o.rot_index(r)
                giant_keys[j - 1] = self.rot[giant_k[j - 1]]
        ok, og = None, 1
        if offset:
            og = o.rot_index(offset)
            ok = self.rot[og]
        return (o.linear_transform_flat(x, l, diags, bStep, stride, baby_keys, baby_k, giant_keys, giant_k, ok, og, offset=offset, slots=n),
                t, d + 1)


def chebyshev_series(ops, y, coeffs, a=-1.0, b=1.0):
    """EvalChebyshevSeries: c_0 / 2 + sum c_k T_k((2x - a - b) / (b - a))."""
    if y[2] == 2:
        y = ops.rescale(y)
    if a != -1.0 or b != 1.0:
        y = ops.rescale(ops.mul_scalar(y, 2.0 / (b - a)))
        y = ops.add_scalar(y, -(a + b) / (b - a))
    c = list(coeffs)
    c[0] *= 0.5
    T = {1: y}

    def t(k):
        if k in T:
            return T[k]
        if k % 2 == 0:
            r = ops.rescale(ops.mult(t(k // 2), t(k // 2)))
            r = ops.add_scalar(ops.add(r, r), -1.0)
        else:
            r = ops.rescale(ops.mult(t(k // 2), t(k // 2 + 1)))
            r = ops.sub(ops.add(r, r), t(1))
        T[k] = r
        return r

    def ev(c):
        d = len(c) - 1
        if cheb_leaf(d):
            towers = min(t(k)[1] for k in range(1, d + 1))
            acc = None
            for k in range(1, d + 1):
                if c[k] == 0.0:
                    continue
                term = ops.mul_scalar(ops.reduce(t(k), towers), c[k])
                acc = term if acc is None else ops.add(acc, term)
            if acc is None:
                acc = ops.mul_scalar(ops.reduce(t(1), towers), 0.0)
            return ops.add_scalar(ops.rescale(acc), c[0])
        M = 1
        while 2 * M <= d:
            M *= 2
        q = [0.0] * (d - M + 1)
        r = list(c[:M])
        q[0] = c[M]
        for j in range(1, d - M + 1):
            q[j] = 2.0 * c[M + j]
            r[M - j] -= c[M + j]
        P = ops.rescale(ops.mul_scalar(t(M), q[0])) if len(q) == 1 else ops.rescale(ops.mult(ev(q), t(M)))
        return ops.add(P, ev(r))

    return ev(c)


# ---- the bootstrap's EvalMod with exact scales (the facade's ChebExact / BootstrapOps exact forms) ----
class ExactOps:
    """Ciphertexts (residues, towers, noise degree, true scale): rescales divide the tracked scale by
    the dropped prime; constants are encoded at the true scale of what they are added to; scalar
    products at an explicit encoding scale."""

    def __init__(self, ops):
        self.b, self.o, self.q = ops, ops.o, ops.q

    def rs(self, a):
        x, t, d, s = a
        r = self.b.rescale((x, t, d))
        return (*r, s / float(self.q[t - 1]))

    def mult(self, a, b):
        r = self.b.mult(a[:3], b[:3])
        return (*r, a[3] * b[3])

    def mult_at(self, a, c, enc):
        x, t, d, s = a
        f = O.elem_for_eval_mult(self.q[:t], c, enc)
        return (np.concatenate([self.o.mul_scalar(self.b.comp(x, t, i), f) for i in range(2)]), t, d + 1, s * enc)

    def reduce(self, a, t_to):
        return (*self.b.reduce(a[:3], t_to), a[3])

    def add(self, a, b):
        assert a[3] == b[3], "exact-scale sum of different scales"
        return (*self.b.add(a[:3], b[:3]), a[3])

    def sub(self, a, b):
        assert a[3] == b[3], "exact-scale sum of different scales"
        return (*self.b.sub(a[:3], b[:3]), a[3])

    def add_const(self, a, c):
        x, t, d, s = a
        if c == 0.0:
            return a
        e = O.elem_for_eval_add(self.q[:t], abs(c), s, 1)
        if c < 0:
            e = np.array([(qi - int(v)) % qi for v, qi in zip(e, self.q[:t])], dtype=np.uint64)
        return (np.concatenate([self.o.add_scalar(self.b.comp(x, t, 0), e), self.b.comp(x, t, 1)]), t, d, s)


def chebyshev_exact(xo, y, c):
    """sum_k c_k T_k(y) (c_0 already halved), y = (residues, towers, degree, true scale)."""
    T = {1: y}

    def t(k):
        if k in T:
            return T[k]
        if k % 2 == 0:
            r = xo.rs(xo.mult(t(k // 2), t(k // 2)))
            r = xo.add_const(xo.add(r, r), -1.0)
        else:
            p = xo.rs(xo.mult(t(k // 2), t(k // 2 + 1)))
            t1 = T[1]
            a = xo.rs(xo.mult_at(t1, 1.0, p[3] * float(xo.q[t1[1] - 1]) / t1[3]))
            a = (a[0], a[1], a[2], p[3])
            r = xo.sub(xo.add(p, p), a)
        T[k] = r
        return r

    def towers_of(d):
        if cheb_leaf(d):
            return min(t(k)[1] for k in range(1, d + 1)) - 1
        M = 1
        while 2 * M <= d:
            M *= 2
        tq = t(M)[1] if d - M == 0 else towers_of(d - M)
        return min(tq, t(M)[1]) - 1

    def ev(cc, target):
        d = len(cc) - 1
        if cheb_leaf(d):
            towers = min(t(k)[1] for k in range(1, d + 1))
            S = (target if target != 0.0 else xo.b.scale) * float(xo.q[towers - 1])
            acc = None
            used = None
            for k in range(1, d + 1):
                if cc[k] == 0.0:
                    continue
                tk = t(k)
                used = tk[1] if used is None else min(used, tk[1])
                term = xo.mult_at(xo.reduce(tk, towers), cc[k], S / tk[3])
                term = (term[0], term[1], term[2], S)
                acc = term if acc is None else xo.add(acc, term)
            if acc is None or used > towers:
                t1 = t(1)
                z = xo.mult_at(xo.reduce(t1, towers), 0.0, S / t1[3])
                z = (z[0], z[1], z[2], S)
                acc = z if acc is None else xo.add(acc, z)
            return xo.add_const(xo.rs(acc), cc[0])
        M = 1
        while 2 * M <= d:
            M *= 2
        q = [0.0] * (d - M + 1)
        r = list(cc[:M])
        q[0] = cc[M]
        for j in range(1, d - M + 1):
            q[j] = 2.0 * cc[M + j]
            r[M - j] -= cc[M + j]
        tM = t(M)
        if len(q) == 1:
            enc = target * float(xo.q[tM[1] - 1]) / tM[3] if target != 0.0 else xo.b.scale
            P = xo.rs(xo.mult_at(tM, q[0], enc))
        else:
            tp = min(towers_of(d - M), tM[1])
            tq = target * float(xo.q[tp - 1]) / tM[3] if target != 0.0 else 0.0
            P = xo.rs(xo.mult(ev(q, tq), tM))
        if target != 0.0:
            P = (P[0], P[1], P[2], target)
        R = ev(r, P[3])
        R = (R[0], R[1], R[2], P[3])
        return xo.add(P, R)

    return ev(c, 0.0)


def evalmod_exact(ops, y, k):
    """EvalMod of y (slots exact at the nominal scale): the series and double angles with exact scales."""
    xo = ExactOps(ops)
    c = list(k["cheb"])
    c[0] *= 0.5
    v = chebyshev_exact(xo, (y[0], y[1], y[2], ops.scale), c)
    for _ in range(k["r"]):
        v = xo.rs(xo.mult(v, v))
        v = xo.add_const(xo.add(v, v), -1.0)
    return v[:3]


def bootstrap(ops, ct, k):
    """EvalBootstrap of ct = (residues, towers, noise degree) with constants k."""
    o, N = ops.o, ops.N
    n = k["slots"]
    corr = k.get("correction", 0)
    if ct[2] == 2:
        ct = ops.rescale(ct)
    if corr:
        ct = ops.rescale(ops.mul_scalar(ct, math.ldexp(1.0, -corr)))
    ct = ops.reduce(ct, 1)
    L1 = o.L1
    raised = (o.mod_raise(ct[0], 0, L1 - 1), L1, 1)
    raised = ops.rescale(ops.mul_scalar(raised, k["pre"]))
    if n < N // 2:
        raised = ops.accumulate(raised, n, N // 2 // n)
    t = raised
    for g in k["cts"]:
        t = ops.rescale(ops.linear_transform(t, g, n))
    cj = ops.rotate(t, 2 * N - 1)
    re = ops.add(t, cj)
    im = ops.monomial(ops.sub(t, cj), 3 * N // 2)

    u = ops.add(evalmod_exact(ops, re, k), ops.monomial(evalmod_exact(ops, im, k), N // 2))
    for g in k["stc"]:
        u = ops.rescale(ops.linear_transform(u, g, n))
    if corr:
        u = ops.mul_integer(u, 1 << corr)
    return u
