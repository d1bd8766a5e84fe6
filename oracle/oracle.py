"""ctypes binding of the CPU oracle (oracle/ckks_oracle.c) — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, always as the
checker. Arrays are numpy uint64, limb-major (one N-word row per limb), canonical residues.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libckks_oracle.so")
_lib = None

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        V, P, U, I, D, U32 = None, C.c_void_p, C.c_uint64, C.c_int, C.c_double, C.c_uint32
        sig = {
            "orc_mulmod": (U, [U, U, U]),
            "orc_powmod": (U, [U, U, U]),
            "orc_invmod": (U, [U, U]),
            "orc_is_prime": (I, [U]),
            "orc_first_prime": (U, [I, U]),
            "orc_last_prime": (U, [I, U]),
            "orc_previous_prime": (U, [U, U]),
            "orc_next_prime": (U, [U, U]),
            "orc_root_of_unity": (U, [U, U]),
            "orc_gen_primes": (I, [I, I, I, I, I, I, u64p]),
            "orc_default_num_special": (I, [u64p, I, I, I]),
            "orc_create": (P, [I, I, I, I, u64p]),
            "orc_destroy": (V, [P]),
            "orc_prime": (U, [P, I]),
            "orc_psi": (U, [P, I]),
            "orc_alpha": (I, [P]),
            "orc_twiddles": (C.POINTER(C.c_uint64), [P, I]),
            "orc_inv_twiddles": (C.POINTER(C.c_uint64), [P, I]),
            "orc_rescale_ql_inv": (U, [P, I, I]),
            "orc_rescale_QlQlInv": (U, [P, I, I]),
            "orc_P_mod": (U, [P, I]),
            "orc_P_inv_mod": (U, [P, I]),
            "orc_ntt_fwd": (V, [P, u64p, i32p, I]),
            "orc_ntt_inv": (V, [P, u64p, i32p, I]),
            "orc_ntt_fwd_raw": (V, [u64p, U, U, u64p]),
            "orc_ntt_inv_raw": (V, [u64p, U, U, u64p, U]),
            "orc_make_tables": (V, [U, U, U, u64p, u64p]),
            "orc_add": (V, [P, u64p, u64p, u64p, I]),
            "orc_sub": (V, [P, u64p, u64p, u64p, I]),
            "orc_mul": (V, [P, u64p, u64p, u64p, I]),
            "orc_neg": (V, [P, u64p, u64p, I]),
            "orc_mul_scalar": (V, [P, u64p, u64p, u64p, I]),
            "orc_add_scalar": (V, [P, u64p, u64p, u64p, I]),
            "orc_tensor": (V, [P, u64p, u64p, u64p, I]),
            "orc_square": (V, [P, u64p, u64p, I]),
            "orc_rescale_poly": (V, [P, u64p, I, u64p]),
            "orc_rescale_ct": (V, [P, u64p, I, I, u64p]),
            "orc_mod_raise": (V, [P, u64p, I, I, I, u64p]),
            "orc_num_digits": (I, [P, I]),
            "orc_modup": (V, [P, u64p, I, u64p]),
            "orc_ksk_inner": (V, [P, u64p, I, u64p, u64p, u64p]),
            "orc_moddown": (V, [P, u64p, I, u64p]),
            "orc_automorph": (V, [P, u64p, I, U32, u64p]),
            "orc_rot_index": (U32, [P, I]),
            "orc_relinearize": (V, [P, u64p, I, u64p, u64p]),
            "orc_eval_mult": (V, [P, u64p, u64p, I, u64p, u64p]),
            "orc_eval_rotate": (V, [P, u64p, I, u64p, U32, u64p]),
            "orc_eval_rotate_ext": (V, [P, u64p, I, u64p, U32, I, u64p]),
            "orc_linear_transform": (V, [P, u64p, I, I, I, I, I, I, u64p, u64p, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                         u64p, U32, u64p, U32, u64p]),
            "orc_linear_transform_flat": (V, [P, u64p, I, I, I, I, I, I, u64p, u64p,
                                              np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), u64p,
                                              np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), u64p, U32, u64p]),
            "orc_accumulate_sum": (V, [P, u64p, I, I, I, I, I, u64p, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                       u64p]),
            "orc_accumulate_lazy": (V, [P, u64p, I, I, I, u64p, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                        u64p]),
            "orc_gen_secret": (V, [P, U, I, u64p]),
            "orc_gen_switch_key": (V, [P, u64p, u64p, U, u64p]),
            "orc_gen_relin_key": (V, [P, u64p, U, u64p]),
            "orc_gen_rot_key": (V, [P, u64p, U32, U, u64p]),
            "orc_encode": (V, [P, f64p, I, D, I, u64p]),
            "orc_encode_ext": (V, [P, f64p, I, D, I, u64p]),
            "orc_decode": (V, [P, u64p, I, I, D, f64p]),
            "orc_encrypt": (V, [P, u64p, I, u64p, U, u64p]),
            "orc_decrypt": (V, [P, u64p, I, I, u64p, u64p]),
            "orc_encrypt_pk": (V, [P, u64p, I, u64p, u64p, U, u64p]),
            "orc_random_poly": (V, [P, U, i32p, I, u64p]),
            "orc_mt_first": (U, [U, I]),
            "orc_set_reduced_noise": (V, [P, I]),
            "orc_num_threads": (I, []),
            "orc_set_threads": (V, [I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---- scalar CRT encodings (src/CKKS/Context.cu:669-760 and 777-863) ----
_MAX_LOG_STEP = 60  # OpenFHE LargeScalingFactorConstants::MAX_LOG_STEP


def _log_approx(operand, sc_factor, max_bits):
    res = abs(operand * sc_factor)
    if res > 0:
        log_sf = int(math.ceil(math.log2(res)))
        return log_sf - min(log_sf, max_bits)
    return 0


def _approx_residues(log_approx, moduli):
    """2^log_approx mod q_i, built in steps of at most 2^60 (Context.cu:733-748)."""
    out = [1] * len(moduli)
    while log_approx > 0:
        step = min(log_approx, _MAX_LOG_STEP)
        out = [(o * (1 << step)) % q for o, q in zip(out, moduli)]
        log_approx -= step
    return out


def elem_for_eval_mult(moduli, operand, sc_factor):
    """ContextData::ElemForEvalMult (src/CKKS/Context.cu:669-760): the residues of
    round(operand * scFactor) for towers with the given moduli; past 2^125 the product is rounded
    at 2^logApprox and scaled back up by CRT multiplication. Python floats are IEEE doubles
    without FMA, the rounding sequence the reference documents."""
    log_approx = _log_approx(operand, sc_factor, 125)
    approx_factor = 2.0 ** log_approx
    large = int(operand / approx_factor * sc_factor + 0.5)  # C++ truncation toward zero
    f = [large % q for q in moduli]  # Python's % is already the non-negative residue
    if log_approx > 0:
        f = [(a * b) % q for a, b, q in zip(f, _approx_residues(log_approx, moduli), moduli)]
    return np.array(f, dtype=np.uint64)


def elem_for_eval_add(moduli, operand, sc_factor, noise_deg):
    """ContextData::ElemForEvalAddOrSub (src/CKKS/Context.cu:777-863) for operand >= 0: the
    residues of round(operand * scFactor) times round(scFactor)^(noise_deg - 1)."""
    if operand < 0:
        raise ValueError("the reference encodes |operand| and subtracts for negative constants")
    log_approx = _log_approx(operand, sc_factor, 61)
    approx_factor = 2.0 ** log_approx
    sc = int(operand * sc_factor / approx_factor + 0.5)
    c = [sc % q for q in moduli]
    if log_approx > 0:
        c = [(a * b) % q for a, b, q in zip(c, _approx_residues(log_approx, moduli), moduli)]
    isf = int(sc_factor + 0.5)
    for _ in range(1, noise_deg):
        c = [(a * isf) % q for a, q in zip(c, moduli)]
    return np.array(c, dtype=np.uint64)


def _check_encode_bound(z, scale):
    """The encoders round every coefficient exactly, which is OpenFHE's CKKSPackedEncoding::Encode
    only while the scaled coefficients stay below 2^61 (MAX_BITS_IN_WORD; above it OpenFHE takes an
    approximate path). A coefficient is at most max(|re| + |im|) of the slots, so inputs at or above
    2^61 (1 - 2^-40) / scale are rejected, as the device encoder does."""
    lim = 2.0 ** 61 * (1.0 - 2.0 ** -40) / scale
    if z.size and not np.all(np.abs(z.real) + np.abs(z.imag) < lim):
        raise ValueError("encode: a slot value times the scale reaches 2^61 (OpenFHE's MAX_BITS_IN_WORD)")


def gen_primes(logN, L1, first_bits, scale_bits, K, aux_bits=60):
    out = np.zeros(L1 + K, dtype=np.uint64)
    rc = lib().orc_gen_primes(logN, L1, first_bits, scale_bits, K, aux_bits, out)
    assert rc == 0
    return out


class Ctx:
    """Oracle context over the primes q_0..q_{L1-1} (Q) and p_0..p_{K-1} (P)."""

    def __init__(self, logN, primes, L1, K, dnum):
        self.logN, self.N, self.L1, self.K, self.dnum = logN, 1 << logN, L1, K, dnum
        self.primes = np.ascontiguousarray(primes, dtype=np.uint64)
        if self.primes.ndim != 1 or self.primes.size != L1 + K or L1 < 1 or K < 1 or dnum < 1:
            raise ValueError(f"oracle context: {self.primes.size} primes for L1={L1}, K={K}, dnum={dnum}")
        self.h = lib().orc_create(logN, L1, K, dnum, self.primes)
        assert self.h
        self.alpha = lib().orc_alpha(self.h)
        self.psi = [int(lib().orc_psi(self.h, i)) for i in range(L1 + K)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_reduced_noise(self, on):
        """ReducedNoise ModUp/ModDown conversions (prescaled residues above q/2 read as v - q)."""
        lib().orc_set_reduced_noise(self.h, int(on))

    # ---- helpers ----
    def q(self, i):
        return int(self.primes[i])

    def zeros(self, nlimbs, ncomp=1):
        return np.zeros((ncomp * nlimbs, self.N), dtype=np.uint64)

    def ext_pids(self, l):
        return list(range(l + 1)) + [self.L1 + k for k in range(self.K)]

    def num_digits(self, l):
        return lib().orc_num_digits(self.h, l)

    def twiddles(self, i):
        p = lib().orc_twiddles(self.h, i)
        return np.ctypeslib.as_array(p, shape=(self.N,)).copy()

    def inv_twiddles(self, i):
        p = lib().orc_inv_twiddles(self.h, i)
        return np.ctypeslib.as_array(p, shape=(self.N,)).copy()

    def random_poly(self, seed, pids):
        pids = np.ascontiguousarray(pids, dtype=np.int32)
        out = np.zeros((len(pids), self.N), dtype=np.uint64)
        lib().orc_random_poly(self.h, seed, pids, len(pids), out)
        return out

    # ---- transforms ----
    def ntt(self, a, pids):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        lib().orc_ntt_fwd(self.h, a, np.ascontiguousarray(pids, dtype=np.int32), len(pids))
        return a

    def intt(self, a, pids):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        lib().orc_ntt_inv(self.h, a, np.ascontiguousarray(pids, dtype=np.int32), len(pids))
        return a

    # ---- element-wise ----
    def _bin(self, fn, a, b):
        out = np.zeros_like(a)
        getattr(lib(), fn)(self.h, np.ascontiguousarray(a), np.ascontiguousarray(b), out, a.shape[0])
        return out

    def add(self, a, b):
        return self._bin("orc_add", a, b)

    def sub(self, a, b):
        return self._bin("orc_sub", a, b)

    def mul(self, a, b):
        return self._bin("orc_mul", a, b)

    def neg(self, a):
        out = np.zeros_like(a)
        lib().orc_neg(self.h, np.ascontiguousarray(a), out, a.shape[0])
        return out

    def mul_scalar(self, a, s):
        out = np.zeros_like(a)
        lib().orc_mul_scalar(self.h, np.ascontiguousarray(a), np.ascontiguousarray(s, dtype=np.uint64), out, a.shape[0])
        return out

    def add_scalar(self, a, s):
        out = np.zeros_like(a)
        lib().orc_add_scalar(self.h, np.ascontiguousarray(a), np.ascontiguousarray(s, dtype=np.uint64), out, a.shape[0])
        return out

    def tensor(self, ct1, ct2, l):
        out = self.zeros(l + 1, 3)
        lib().orc_tensor(self.h, np.ascontiguousarray(ct1), np.ascontiguousarray(ct2), out, l + 1)
        return out

    # ---- rescale / key switching ----
    def rescale_poly(self, x, l):
        out = self.zeros(l)
        lib().orc_rescale_poly(self.h, np.ascontiguousarray(x), l, out)
        return out

    def rescale_ct(self, ct, l, ncomp=2):
        out = self.zeros(l, ncomp)
        lib().orc_rescale_ct(self.h, np.ascontiguousarray(ct), l, ncomp, out)
        return out

    def mod_raise(self, ct, l_in, l, ncomp=2):
        """CKKS bootstrapping ModRaise (src/CKKS/Bootstrap.cu:708-930): tower 0 of each component
        lifted centred to towers 0..l."""
        out = self.zeros(l + 1, ncomp)
        lib().orc_mod_raise(self.h, np.ascontiguousarray(ct), l_in, ncomp, l, out)
        return out

    def modup(self, x, l):
        beta = self.num_digits(l)
        out = np.zeros((beta * (l + 1 + self.K), self.N), dtype=np.uint64)
        lib().orc_modup(self.h, np.ascontiguousarray(x), l, out)
        return out

    def ksk_inner(self, digits, l, key):
        ext = l + 1 + self.K
        o0 = np.zeros((ext, self.N), dtype=np.uint64)
        o1 = np.zeros((ext, self.N), dtype=np.uint64)
        lib().orc_ksk_inner(self.h, np.ascontiguousarray(digits), l, np.ascontiguousarray(key), o0, o1)
        return o0, o1

    def moddown(self, x, l):
        out = self.zeros(l + 1)
        lib().orc_moddown(self.h, np.ascontiguousarray(x), l, out)
        return out

    def automorph(self, x, k):
        out = np.zeros_like(x)
        lib().orc_automorph(self.h, np.ascontiguousarray(x), x.shape[0], k, out)
        return out

    def rot_index(self, r):
        return int(lib().orc_rot_index(self.h, r))

    def relinearize(self, ct3, l, key):
        out = self.zeros(l + 1, 2)
        lib().orc_relinearize(self.h, np.ascontiguousarray(ct3), l, np.ascontiguousarray(key), out)
        return out

    def eval_mult(self, ct1, ct2, l, key):
        out = self.zeros(l + 1, 2)
        lib().orc_eval_mult(self.h, np.ascontiguousarray(ct1), np.ascontiguousarray(ct2), l, np.ascontiguousarray(key), out)
        return out

    def eval_rotate(self, ct, l, key, k):
        out = self.zeros(l + 1, 2)
        lib().orc_eval_rotate(self.h, np.ascontiguousarray(ct), l, np.ascontiguousarray(key), k, out)
        return out

    def eval_rotate_ext(self, ct, l, key, k, add_c0=True):
        out = np.zeros((2 * (l + 1 + self.K), self.N), dtype=np.uint64)
        lib().orc_eval_rotate_ext(self.h, np.ascontiguousarray(ct), l, np.ascontiguousarray(key), k, int(add_c0), out)
        return out

    def linear_transform(self, ct, l, diags, bStep, stride, baby_keys, baby_k, giant_key=None, giant_k=1, off_key=None, off_k=1,
                         offset=0, slots=None):
        """baby_keys: list of bStep keys (None where unused); diags: rowSize x (l+1+K) limbs."""
        slots = slots or self.N // 2
        rowSize = diags.shape[0] // (l + 1 + self.K)
        zero = np.zeros(self.key_shape(), dtype=np.uint64)
        bk = np.ascontiguousarray(np.concatenate([zero if k is None else k for k in baby_keys]))
        bkk = np.ascontiguousarray(np.array([int(x) for x in baby_k], dtype=np.uint32))
        gk = zero if giant_key is None else np.ascontiguousarray(giant_key)
        ok = zero if off_key is None else np.ascontiguousarray(off_key)
        out = self.zeros(l + 1, 2)
        lib().orc_linear_transform(self.h, np.ascontiguousarray(ct), l, rowSize, bStep, stride, offset, slots,
                                   np.ascontiguousarray(diags), bk, bkk, gk, giant_k, ok, off_k, out)
        return out

    def linear_transform_flat(self, ct, l, diags, bStep, stride, baby_keys, baby_k, giant_keys=(), giant_ks=(), off_key=None, off_k=1,
                              offset=0, slots=None):
        """The same linear map with the giant steps as independent rotations: giant_keys[j-1] /
        giant_ks[j-1] rotate R_j by j*bStep*stride (this build's own sequence for its bootstrap)."""
        slots = slots or self.N // 2
        rowSize = diags.shape[0] // (l + 1 + self.K)
        gStep = (rowSize + bStep - 1) // bStep
        zero = np.zeros(self.key_shape(), dtype=np.uint64)
        bk = np.ascontiguousarray(np.concatenate([zero if k is None else k for k in baby_keys]))
        bkk = np.ascontiguousarray(np.array([int(x) for x in baby_k], dtype=np.uint32))
        gkeys = list(giant_keys) + [None] * (gStep - 1 - len(giant_keys))
        gk = np.ascontiguousarray(np.concatenate([zero if k is None else k for k in gkeys] or [zero]))
        gks = list(giant_ks) + [1] * (gStep - 1 - len(giant_ks))
        gkk = np.ascontiguousarray(np.array([int(x) for x in gks] or [1], dtype=np.uint32))
        ok = zero if off_key is None else np.ascontiguousarray(off_key)
        out = self.zeros(l + 1, 2)
        lib().orc_linear_transform_flat(self.h, np.ascontiguousarray(ct), l, rowSize, bStep, stride, offset, slots,
                                        np.ascontiguousarray(diags), bk, bkk, gk, gkk, ok, off_k, out)
        return out

    @staticmethod
    def accumulate_indices(radix, stride, size, start=1):
        """Rotation indices of the fold, in fold order (GetAccumulateRotationIndices,
        src/CKKS/AccumulateBroadcast.cu:86-96, from start instead of 1)."""
        out, s = [], start
        while start > 0 and s < size:
            i = 1
            while i < radix and i * s < size:
                out.append(i * s * stride)
                i += 1
            s *= radix
        return out

    def _fold_keys(self, n, keys, ks):
        if len(keys) != n or len(ks) != n:
            raise ValueError("one key per fold rotation")
        kk = np.ascontiguousarray(np.concatenate(keys)) if n else np.zeros(self.key_shape(), dtype=np.uint64)
        kx = np.ascontiguousarray(np.array([int(x) for x in ks] or [1], dtype=np.uint32))
        return kk, kx

    def accumulate_sum(self, ct, l, radix, stride, size, keys, ks):
        """The lazy fold (reference Accumulate(ct, bStep, stride, size), behind accumulateSum):
        keys / ks one rotation key and Galois element per index of accumulate_indices."""
        kk, kx = self._fold_keys(len(self.accumulate_indices(radix, stride, size)), keys, ks)
        out = self.zeros(l + 1, 2)
        lib().orc_accumulate_lazy(self.h, np.ascontiguousarray(ct), l, radix, size, kk, kx, out)
        return out

    def accumulate_sum_from(self, ct, l, radix, stride, size, keys, ks, start):
        """The cascade from a start factor (reference Accumulate(ct, bStep, stride, size, startFactor),
        behind accumulateSumInPlace(ct, slots, stride, start))."""
        kk, kx = self._fold_keys(len(self.accumulate_indices(radix, stride, size, start)), keys, ks)
        out = self.zeros(l + 1, 2)
        lib().orc_accumulate_sum(self.h, np.ascontiguousarray(ct), l, radix, stride, size, start, kk, kx, out)
        return out

    # ---- keys / encoding / encryption ----
    def gen_secret(self, seed, h=0):
        s = self.zeros(self.L1 + self.K)
        lib().orc_gen_secret(self.h, seed, h, s)
        return s

    def key_shape(self):
        return (self.dnum * 2 * (self.L1 + self.K), self.N)

    def gen_relin_key(self, s, seed):
        key = np.zeros(self.key_shape(), dtype=np.uint64)
        lib().orc_gen_relin_key(self.h, s, seed, key)
        return key

    def gen_rot_key(self, s, k, seed):
        key = np.zeros(self.key_shape(), dtype=np.uint64)
        lib().orc_gen_rot_key(self.h, s, k, seed, key)
        return key

    def gen_switch_key(self, s_old, s_new, seed):
        key = np.zeros(self.key_shape(), dtype=np.uint64)
        lib().orc_gen_switch_key(self.h, np.ascontiguousarray(s_old), np.ascontiguousarray(s_new), seed, key)
        return key

    def encode(self, z, scale, l):
        z = np.asarray(z, dtype=np.complex128)
        _check_encode_bound(z, scale)
        flat = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).reshape(-1))
        out = self.zeros(l + 1)
        lib().orc_encode(self.h, flat, len(z), scale, l, out)
        return out

    def encode_ext(self, z, scale, l):
        z = np.asarray(z, dtype=np.complex128)
        _check_encode_bound(z, scale)
        flat = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).reshape(-1))
        out = self.zeros(l + 1 + self.K)
        lib().orc_encode_ext(self.h, flat, len(z), scale, l, out)
        return out

    def decode(self, m, l, n, scale):
        out = np.zeros(2 * n, dtype=np.float64)
        lib().orc_decode(self.h, np.ascontiguousarray(m), l, n, scale, out)
        return out[0::2] + 1j * out[1::2]

    def encrypt(self, m, l, s, seed):
        ct = self.zeros(l + 1, 2)
        lib().orc_encrypt(self.h, np.ascontiguousarray(m), l, s, seed, ct)
        return ct

    def encrypt_pk(self, m, l, pk_b, pk_a, seed):
        ct = self.zeros(l + 1, 2)
        lib().orc_encrypt_pk(self.h, np.ascontiguousarray(m), l, np.ascontiguousarray(pk_b), np.ascontiguousarray(pk_a), seed, ct)
        return ct

    def decrypt(self, ct, l, s, ncomp=2):
        m = self.zeros(l + 1)
        lib().orc_decrypt(self.h, np.ascontiguousarray(ct), l, ncomp, s, m)
        return m
