"""FLEXIBLEAUTO metadata algebra over the oracle, for replaying facade sessions.

A restatement of OpenFHE's FLEXIBLEAUTO scaling technique as the reference restates it: the
per-level scaling factors (CryptoParametersCKKSRNS; src/CKKS/openfhe-interface/RawCiphertext.cu:305-317),
the level-dependent scalar encodings (src/CKKS/Context.cu:669-863), and the level / depth
alignment before additions and products (LeveledSHECKKSRNS::AdjustLevelsAndDepthInPlace and
AdjustForMult; src/CKKS/Ciphertext.cpp:624-658, 1917-2070). Ciphertexts are (residues, towers,
ncomp, deg, scale) with towers = device level + 1; the arithmetic is the oracle's.
Test infrastructure: used by the tests only.
"""
from dataclasses import dataclass

import numpy as np

from oracle import oracle as O


@dataclass
class FCt:
    data: np.ndarray   # ncomp * towers limbs
    towers: int
    ncomp: int
    deg: int
    scale: float

    def comp(self, c):
        return self.data[c * self.towers:(c + 1) * self.towers]


def scale_tables(moduli_q):
    """(real, big, mrf) by towers - 1: the top level's factor is the top prime, each level below
    the one above squared over the prime dropped (prevSF * prevSF / q, in that order), the square
    for depth 2, and the dropped prime as a rescale's factor."""
    L1 = len(moduli_q)
    real = [0.0] * L1
    real[L1 - 1] = float(int(moduli_q[L1 - 1]))
    for l in range(L1 - 2, -1, -1):
        prev = real[l + 1]
        real[l] = prev * prev / float(int(moduli_q[l + 1]))
        r = real[l] / real[L1 - 1]
        if r <= 0.5 or r >= 2.0:
            raise ValueError("FLEXIBLEAUTO cannot support this number of levels")
    big = [v * v for v in real]
    mrf = [float(int(q)) for q in moduli_q]
    return real, big, mrf


class Flex:
    def __init__(self, o: "O.Ctx", relin=None):
        self.o = o
        self.moduli = [int(o.q(i)) for i in range(o.L1)]
        self.real, self.big, self.mrf = scale_tables(self.moduli)
        self.relin = relin

    # ---- primitives
    def mod_reduce(self, c: FCt) -> FCt:
        l = c.towers - 1
        d = self.o.rescale_ct(c.data, l, c.ncomp)
        return FCt(d, c.towers - 1, c.ncomp, max(c.deg - 1, 1), c.scale / self.mrf[l])

    def mult_core(self, c: FCt, operand: float) -> FCt:
        f = O.elem_for_eval_mult(self.moduli[:c.towers], operand, self.real[c.towers - 1])
        d = np.concatenate([self.o.mul_scalar(c.comp(k), f) for k in range(c.ncomp)])
        return FCt(d, c.towers, c.ncomp, c.deg + 1, c.scale * self.real[c.towers - 1])

    @staticmethod
    def drop(c: FCt, towers: int) -> FCt:
        d = np.concatenate([c.comp(k)[:towers] for k in range(c.ncomp)])
        return FCt(d, towers, c.ncomp, c.deg, c.scale)

    def adjust_down(self, c1: FCt, towers2: int, deg2: int, scale2: float) -> FCt:
        l1, l2 = c1.towers - 1, towers2 - 1
        if c1.deg == 2:
            if deg2 == 2:
                x = self.mod_reduce(self.mult_core(c1, scale2 / c1.scale * self.mrf[l1] / self.real[l1]))
                if x.towers > towers2:
                    x = self.drop(x, towers2)
                x.scale = scale2
            elif l1 - 1 == l2:
                x = self.mod_reduce(c1)
            else:
                x = self.mod_reduce(self.mult_core(c1, self.big[l2 + 1] / c1.scale * self.mrf[l1] / self.real[l1]))
                if x.towers > towers2 + 1:
                    x = self.drop(x, towers2 + 1)
                x = self.mod_reduce(x)
                x.scale = scale2
        else:
            if deg2 == 2:
                x = self.drop(self.mult_core(c1, scale2 / c1.scale / self.real[l1]), towers2)
                x.scale = scale2
            else:
                x = self.mult_core(c1, self.big[l2 + 1] / c1.scale / self.real[l1])
                if x.towers > towers2 + 1:
                    x = self.drop(x, towers2 + 1)
                x = self.mod_reduce(x)
                x.scale = scale2
        return x

    def adjust(self, a: FCt, b: FCt):
        if a.towers > b.towers:
            a = self.adjust_down(a, b.towers, b.deg, b.scale)
        elif a.towers < b.towers:
            b = self.adjust_down(b, a.towers, a.deg, a.scale)
        elif a.deg < b.deg:
            a = self.mult_core(a, 1.0)
        elif b.deg < a.deg:
            b = self.mult_core(b, 1.0)
        return a, b

    def adjust_to_one(self, a: FCt, b: FCt):
        a, b = self.adjust(a, b)
        if a.deg == 2:
            a, b = self.mod_reduce(a), self.mod_reduce(b)
        return a, b

    # ---- operations
    def add(self, a: FCt, b: FCt, sub=False) -> FCt:
        a, b = self.adjust(a, b)
        t = a.towers
        op = self.o.sub if sub else self.o.add
        d = np.concatenate([op(a.comp(k), b.comp(k)) for k in range(a.ncomp)])
        return FCt(d, t, a.ncomp, a.deg, a.scale)

    def mult(self, a: FCt, b: FCt) -> FCt:
        a, b = self.adjust_to_one(a, b)
        l = a.towers - 1
        d = self.o.eval_mult(a.data, b.data, l, self.relin)
        return FCt(d, a.towers, 2, a.deg + b.deg, a.scale * b.scale)

    def mult_scalar(self, a: FCt, s: float) -> FCt:
        if a.deg == 2:
            a = self.mod_reduce(a)
        return self.mult_core(a, s)

    def add_scalar(self, a: FCt, s: float) -> FCt:
        c = O.elem_for_eval_add(self.moduli[:a.towers], abs(s), self.real[a.towers - 1], a.deg)
        if s < 0:
            c = np.array([(int(q) - int(v)) % int(q) for v, q in zip(c, self.moduli[:a.towers])], dtype=np.uint64)
        d = np.concatenate([self.o.add_scalar(a.comp(0), c)] + [a.comp(k) for k in range(1, a.ncomp)])
        return FCt(d, a.towers, a.ncomp, a.deg, a.scale)

    def plain(self, m: np.ndarray, towers: int) -> FCt:
        """A plaintext encoded at the top level's factor, as a 1-component ciphertext."""
        return FCt(m, towers, 1, 1, self.real[towers - 1])

    def mult_plain(self, a: FCt, p: FCt) -> FCt:
        a, p = self.adjust_to_one(a, p)
        d = np.concatenate([self.o.mul(a.comp(k), p.data) for k in range(a.ncomp)])
        return FCt(d, a.towers, a.ncomp, a.deg + p.deg, a.scale * p.scale)

    def add_plain(self, a: FCt, p: FCt, sub=False) -> FCt:
        a, p = self.adjust(a, p)
        op = self.o.sub if sub else self.o.add
        d = np.concatenate([op(a.comp(0), p.data)] + [a.comp(k) for k in range(1, a.ncomp)])
        return FCt(d, a.towers, a.ncomp, a.deg, a.scale)

    def negate(self, a: FCt) -> FCt:
        d = np.concatenate([self.o.neg(a.comp(k)) for k in range(a.ncomp)])
        return FCt(d, a.towers, a.ncomp, a.deg, a.scale)

    def rotate(self, a: FCt, key, k) -> FCt:
        return FCt(self.o.eval_rotate(a.data, a.towers - 1, key, k), a.towers, 2, a.deg, a.scale)

    def wsum(self, cts, ws) -> FCt:
        cts = list(cts)
        mi = min(range(len(cts)), key=lambda i: (cts[i].towers, i))
        for i in range(len(cts)):
            if i != mi:
                cts[i], cts[mi] = self.adjust(cts[i], cts[mi])
        if cts[mi].deg == 2:
            cts = [self.mod_reduce(c) for c in cts]
        acc = None
        for c, w in zip(cts, ws):
            t = self.mult_core(c, w)
            if acc is None:
                acc = t
            else:
                d = np.concatenate([self.o.add(acc.comp(k), t.comp(k)) for k in range(acc.ncomp)])
                acc = FCt(d, acc.towers, acc.ncomp, acc.deg, acc.scale)
        return acc

    def decrypt(self, c: FCt, s, slots):
        if c.deg == 2 and c.towers > 1:
            c = self.mod_reduce(c)
        m = self.o.decrypt(c.data, c.towers - 1, s, c.ncomp)
        return self.o.decode(m, c.towers - 1, slots, c.scale)
