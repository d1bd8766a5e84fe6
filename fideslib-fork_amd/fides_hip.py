"""Python binding of the C ABI in include/fideslib_hip.h (libfideslib_hip.so).

Plumbing for tests and the benchmark: numpy <-> device buffers and thin wrappers that call the
HIP kernels through ctypes. There is no CPU fallback: if the shared library is missing or no GPU
is present, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfideslib_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fideslib_hip.h")

_lib = None


class FHError(RuntimeError):
    pass


def load():
    """Load libfideslib_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FHError(f"{LIB_PATH} is missing: build it with `make -C fideslib-fork_amd`")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P, I, U, S, VP = C.c_void_p, C.c_int, C.c_uint64, C.c_size_t, C.c_void_p
    PP = C.POINTER(C.c_void_p)
    u64p, i32p, ip = C.POINTER(C.c_uint64), C.POINTER(C.c_int32), C.POINTER(C.c_int)
    sig = {
        "fh_last_error": (C.c_char_p, []),
        "fh_version": (I, []),
        "fh_context_create": (I, [I, I, I, I, u64p, PP]),
        "fh_context_create_generated": (I, [I, I, I, I, I, I, I, PP]),
        "fh_context_destroy": (I, [P]),
        "fh_context_shape": (I, [P, ip]),
        "fh_context_set_reduced_noise": (I, [P, I]),
        "fh_context_moduli": (I, [P, u64p, u64p]),
        "fh_synchronize": (I, [P]),
        "fh_device_alloc": (I, [P, S, PP]),
        "fh_device_free": (I, [P, VP]),
        "fh_copy_h2d": (I, [P, VP, VP, S]),
        "fh_copy_d2h": (I, [P, VP, VP, S]),
        "fh_copy_d2d": (I, [P, VP, VP, S]),
        "fh_ntt_forward": (I, [P, VP, i32p, I, I, S]),
        "fh_ntt_inverse": (I, [P, VP, i32p, I, I, S]),
        "fh_eltwise_binary": (I, [P, I, VP, VP, VP, i32p, I]),
        "fh_eltwise_scalar": (I, [P, I, VP, u64p, VP, i32p, I]),
        "fh_tensor": (I, [P, VP, VP, I, VP]),
        "fh_linear_wsum": (I, [P, I, VP, i32p, u64p, I, I, u64p, C.c_uint32, VP]),
        "fh_encode_slots": (I, [P, VP, I, I, C.c_double, i32p, I, VP, S]),
        "fh_encode_coeffs": (I, [P, VP, I, i32p, I, VP, S]),
        "fh_eval_mult_many": (I, [P, I, VP, I, VP, I, I, VP, VP]),
        "fh_rescale": (I, [P, VP, I, I, VP]),
        "fh_rescale_noalloc": (I, [P, VP, I, I, VP]),
        "fh_stream_copy": (I, [P, VP, VP, S]),
        "fh_mod_raise": (I, [P, VP, I, I, I, VP]),
        "fh_key_create": (I, [P, u64p, PP]),
        "fh_key_destroy": (I, [P, P]),
        "fh_num_digits": (I, [P, I]),
        "fh_modup": (I, [P, VP, I, VP]),
        "fh_key_inner_product": (I, [P, VP, I, P, VP, C.c_uint32, VP, VP]),
        "fh_moddown": (I, [P, VP, I, I, VP]),
        "fh_automorph": (I, [P, VP, I, C.c_uint32, VP]),
        "fh_galois_index": (I, [P, C.c_int32, C.POINTER(C.c_uint32)]),
        "fh_relinearize": (I, [P, VP, I, P, VP]),
        "fh_eval_mult": (I, [P, VP, VP, I, P, VP]),
        "fh_eval_rotate": (I, [P, VP, I, P, C.c_uint32, VP]),
        "fh_eval_rotate_ext": (I, [P, VP, I, P, C.c_uint32, I, VP]),
        "fh_rotate_batch": (I, [P, I, VP, S, I, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), VP]),
        "fh_rotate_hoisted": (I, [P, VP, I, I, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), I, I, VP]),
        "fh_fill_uniform": (I, [P, VP, i32p, I, I, S, U]),
        "fh_linear_transform": (I, [P, VP, I, I, I, I, I, I, VP, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), P, C.c_uint32, P,
                                    C.c_uint32, VP]),
        "fh_linear_transform_flat": (I, [P, VP, I, I, I, I, I, I, VP, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), P, C.c_uint32, VP]),
        "fh_batch_create": (I, [P, PP]),
        "fh_batch_destroy": (I, [P]),
        "fh_batch_rotate": (I, [P, VP, I, P, C.c_uint32, VP]),
        "fh_batch_mult": (I, [P, VP, VP, I, P, I, VP]),
        "fh_batch_mult_strided": (I, [P, VP, I, VP, I, I, P, I, VP]),
        "fh_batch_add": (I, [P, VP, VP, I, I, VP]),
        "fh_batch_mult_plain": (I, [P, VP, VP, I, I, VP]),
        "fh_batch_add_plain": (I, [P, VP, VP, I, VP]),
        "fh_batch_level_reduce": (I, [P, VP, I, I, I, VP]),
        "fh_batch_reset": (I, [P]),
        "fh_batch_finalize": (I, [P]),
        "fh_batch_run": (I, [P, I]),
        "fh_batch_info": (I, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "fh_serve_run": (I, [P, C.POINTER(ServeConfig), C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), P, C.POINTER(ServeResult)]),
        "fh_linear_transform_batch": (I, [P, I, VP, S, I, I, I, I, I, I, VP, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), P,
                                          C.c_uint32, P, C.c_uint32, VP]),
        "fh_accumulate_sum": (I, [P, VP, I, I, I, I, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), I, VP]),
        "fh_accumulate_sum_from": (I, [P, VP, I, I, I, I, I, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), I, VP]),
        "fh_accumulate_indices": (I, [I, I, I, I, C.POINTER(C.c_int32), I]),
        "fh_tune_set": (I, [P, C.c_char_p, I]),
        "fh_tune_get": (I, [P, C.c_char_p, C.POINTER(C.c_int)]),
        "fh_autotune": (I, [P, I, C.POINTER(AutotuneReport)]),
        "fh_profile_enable": (I, [P, I]),
        "fh_profile_reset": (I, [P]),
        "fh_profile_read": (I, [P, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


class AutotuneReport(C.Structure):
    _fields_ = [("split", C.c_int), ("threads", C.c_int * 4), ("chunk_mb", C.c_int), ("candidates", C.c_int), ("limbs", C.c_int),
                ("ms_default", C.c_double), ("ms_best", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("rate_per_s", C.c_double), ("requests", C.c_int), ("start_level", C.c_int), ("hbm_budget_gb", C.c_double),
                ("mode", C.c_int), ("inference_share", C.c_double), ("seed", C.c_uint64), ("checksum", C.c_int),
                ("bootstrap_share", C.c_double), ("bootstrap", C.c_void_p), ("bootstrap_user", C.c_void_p),
                ("request_checksums", C.POINTER(C.c_uint64)), ("request_outputs", C.POINTER(C.c_void_p)),
                ("request_output_capacity", C.c_size_t), ("request_output_words", C.POINTER(C.c_uint64)),
                ("bootstrap_concurrent", C.c_int), ("ticks_in_flight", C.c_int), ("batch_min_ops", C.c_int)]


class ServeResult(C.Structure):
    _fields_ = [("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("requests_per_s", C.c_double), ("wall_s", C.c_double), ("mean_batch_ops", C.c_double), ("ticks", C.c_int),
                ("completed", C.c_int), ("inference", C.c_int), ("polynomial", C.c_int), ("max_active", C.c_int),
                ("checksum", C.c_uint64), ("bootstraps", C.c_int), ("p50_other_ms", C.c_double), ("p99_other_ms", C.c_double)]


def _check(rc):
    if rc != 0:
        raise FHError(f"fideslib_hip error {rc}: {load().fh_last_error().decode()}")


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _i32(ids):
    if ids is None:
        return None
    arr = (C.c_int32 * len(ids))(*[int(i) for i in ids])
    return arr


class DeviceArray:
    """A device allocation of uint64 words with a numpy-like shape (freed with the context pool)."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.words = int(np.prod(self.shape))
        p = C.c_void_p()
        _check(load().fh_device_alloc(ctx.h, max(self.words, 1) * 8, C.byref(p)))
        self.ptr = p.value

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.uint64)
        assert host.size == self.words, (host.shape, self.shape)
        _check(load().fh_copy_h2d(self.ctx.h, self.ptr, host.ctypes.data, self.words * 8))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.uint64)
        _check(load().fh_copy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.words * 8))
        return out

    def addr(self, word_offset=0):
        return self.ptr + 8 * word_offset

    def free(self):
        if self.ptr:
            _check(load().fh_device_free(self.ctx.h, self.ptr))
            self.ptr = None


class Context:
    """A device context (primes, tables, stream, pool)."""

    def __init__(self, log_n, num_q, num_p, dnum, moduli=None, first_bits=60, scale_bits=50, aux_bits=60):
        L = load()
        h = C.c_void_p()
        if moduli is not None:
            m = np.ascontiguousarray(moduli, dtype=np.uint64)
            _check(L.fh_context_create(log_n, num_q, num_p, dnum, _u64p(m), C.byref(h)))
        else:
            _check(L.fh_context_create_generated(log_n, num_q, num_p, dnum, first_bits, scale_bits, aux_bits, C.byref(h)))
        self.h = h.value
        shape = (C.c_int * 5)()
        _check(L.fh_context_shape(self.h, shape))
        self.log_n, self.L1, self.K, self.dnum, self.alpha = list(shape)
        self.N = 1 << self.log_n
        self.moduli = np.zeros(self.L1 + self.K, dtype=np.uint64)
        self.psi = np.zeros(self.L1 + self.K, dtype=np.uint64)
        _check(L.fh_context_moduli(self.h, _u64p(self.moduli), _u64p(self.psi)))

    def close(self):
        if self.h:
            _check(load().fh_context_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reduced_noise(self, on):
        _check(load().fh_context_set_reduced_noise(self.h, int(on)))

    def sync(self):
        _check(load().fh_synchronize(self.h))

    def array(self, shape):
        return DeviceArray(self, shape)

    def to_device(self, host):
        host = np.ascontiguousarray(host, dtype=np.uint64)
        return DeviceArray(self, host.shape).upload(host)

    # ---- NTT ----
    def ntt_forward(self, d, prime_ids=None, num_limbs=None, num_polys=1, poly_stride=0, offset=0):
        nl = num_limbs if num_limbs is not None else (len(prime_ids) if prime_ids is not None else d.shape[0])
        _check(load().fh_ntt_forward(self.h, d.addr(offset), _i32(prime_ids), nl, num_polys, poly_stride))

    def ntt_inverse(self, d, prime_ids=None, num_limbs=None, num_polys=1, poly_stride=0, offset=0):
        nl = num_limbs if num_limbs is not None else (len(prime_ids) if prime_ids is not None else d.shape[0])
        _check(load().fh_ntt_inverse(self.h, d.addr(offset), _i32(prime_ids), nl, num_polys, poly_stride))

    # ---- element-wise ----
    def eltwise(self, op, a, b, out, prime_ids=None):
        nl = len(prime_ids) if prime_ids is not None else a.shape[0]
        _check(load().fh_eltwise_binary(self.h, op, a.ptr, b.ptr, out.ptr, _i32(prime_ids), nl))

    def scalar(self, op, a, scalars, out, prime_ids=None):
        nl = len(prime_ids) if prime_ids is not None else a.shape[0]
        s = None if scalars is None else np.ascontiguousarray(scalars, dtype=np.uint64)
        _check(load().fh_eltwise_scalar(self.h, op, a.ptr, None if s is None else _u64p(s), out.ptr, _i32(prime_ids), nl))

    def encode_slots(self, values, scale, prime_ids, out, out_stride=0):
        """Encode count x slots complex values (a 2-D array) into count polynomials of
        len(prime_ids) limbs (EVALUATION form) on the device."""
        v = np.asarray(values, dtype=np.complex128)
        if v.ndim == 1:
            v = v[None, :]
        flat = np.ascontiguousarray(np.stack([v.real, v.imag], axis=-1).reshape(-1))
        _check(load().fh_encode_slots(self.h, flat.ctypes.data, v.shape[0], v.shape[1], float(scale), _i32(list(prime_ids)), len(prime_ids),
                                      out.ptr, out_stride))

    def encode_coeffs(self, coeffs, prime_ids, out, out_stride=0):
        """Residues and forward NTT of count x N signed integer coefficients."""
        m = np.ascontiguousarray(np.atleast_2d(np.asarray(coeffs, dtype=np.int64)))
        _check(load().fh_encode_coeffs(self.h, m.ctypes.data, m.shape[0], _i32(list(prime_ids)), len(prime_ids), out.ptr, out_stride))

    def linear_wsum(self, cts, towers_in, weights, towers, ncomp, out, add=None, add_components=1):
        """out = sum_k weights[k] * cts[k] over the first `towers` limbs of each of ncomp components
        (cts[k] has towers_in[k] limbs per component), plus add on the components in the bit mask
        add_components."""
        n = len(cts)
        ptrs = (C.c_void_p * n)(*[c.ptr for c in cts])
        w = np.ascontiguousarray(weights, dtype=np.uint64).reshape(n, towers)
        a = None if add is None else np.ascontiguousarray(add, dtype=np.uint64)
        _check(load().fh_linear_wsum(self.h, n, ptrs, _i32(list(towers_in)), _u64p(w), towers, ncomp, None if a is None else _u64p(a),
                                     add_components, out.ptr))

    def eval_mult_many(self, n, ct1, ct2, level, relin_key, out, towers1=None, towers2=None):
        """n products of ciphertexts back to back (towers1 / towers2 limbs per component, default
        level + 1) with one relinearization key."""
        t1 = level + 1 if towers1 is None else towers1
        t2 = level + 1 if towers2 is None else towers2
        _check(load().fh_eval_mult_many(self.h, n, ct1.ptr, t1, ct2.ptr, t2, level, relin_key.h, out.ptr))

    def tensor(self, ct1, ct2, level, out3):
        _check(load().fh_tensor(self.h, ct1.ptr, ct2.ptr, level, out3.ptr))

    # ---- rescale / key switching ----
    def rescale(self, ct, level, ncomp, out):
        _check(load().fh_rescale(self.h, ct.ptr, level, ncomp, out.ptr))

    def rescale_noalloc(self, ct, level, ncomp, out):
        _check(load().fh_rescale_noalloc(self.h, ct.ptr, level, ncomp, out.ptr))

    def mod_raise(self, ct, level_in, ncomp, level, out):
        _check(load().fh_mod_raise(self.h, ct.ptr, level_in, ncomp, level, out.ptr))

    def key(self, host_key):
        return Key(self, host_key)

    def key_from_device(self, dev):
        """A key copied from a device buffer in the upload layout."""
        assert dev.words == self.dnum * 2 * (self.L1 + self.K) * self.N
        k = Key.__new__(Key)
        h = C.c_void_p()
        _check(load().fh_key_create(self.h, C.cast(dev.ptr, C.POINTER(C.c_uint64)), C.byref(h)))
        k.ctx, k.h = self, h.value
        return k

    def num_digits(self, level):
        r = load().fh_num_digits(self.h, level)
        if r < 0:
            _check(r)
        return r

    def modup(self, poly, level, digits, offset=0):
        _check(load().fh_modup(self.h, poly.addr(offset), level, digits.ptr))

    def key_inner_product(self, digits, level, key, out0, out1, c0=None, galois=1):
        _check(load().fh_key_inner_product(self.h, digits.ptr, level, key.h, None if c0 is None else c0.ptr,
                                           galois, out0.ptr, out1.ptr))

    def moddown(self, x, level, num_polys, out):
        _check(load().fh_moddown(self.h, x.ptr, level, num_polys, out.ptr))

    def automorph(self, x, num_limbs, galois, out):
        _check(load().fh_automorph(self.h, x.ptr, num_limbs, galois, out.ptr))

    def galois_index(self, rotation):
        v = C.c_uint32()
        _check(load().fh_galois_index(self.h, rotation, C.byref(v)))
        return v.value

    def relinearize(self, ct3, level, key, out):
        _check(load().fh_relinearize(self.h, ct3.ptr, level, key.h, out.ptr))

    def eval_mult(self, ct1, ct2, level, key, out):
        _check(load().fh_eval_mult(self.h, ct1.ptr, ct2.ptr, level, key.h, out.ptr))

    def eval_rotate(self, ct, level, key, galois, out):
        _check(load().fh_eval_rotate(self.h, ct.ptr, level, key.h, galois, out.ptr))

    def eval_rotate_ext(self, ct, level, key, galois, out_ext, add_c0=True):
        _check(load().fh_eval_rotate_ext(self.h, ct.ptr, level, key.h, galois, int(add_c0), out_ext.ptr))

    def rotate_batch(self, cts, count, level, keys, galois, out, ct_stride=0, offset=0):
        kp = (C.c_void_p * count)(*[k.h for k in keys])
        gl = (C.c_uint32 * count)(*galois)
        _check(load().fh_rotate_batch(self.h, count, cts.addr(offset), ct_stride, level, kp, gl, out.ptr))

    def rotate_hoisted(self, ct, level, keys, galois, out, add_c0=True, moddown=True):
        n = len(keys)
        kp = (C.c_void_p * n)(*[k.h for k in keys])
        gl = (C.c_uint32 * n)(*galois)
        _check(load().fh_rotate_hoisted(self.h, ct.ptr, level, n, kp, gl, int(add_c0), int(moddown), out.ptr))

    def fill_uniform(self, d, prime_ids=None, num_limbs=None, num_polys=1, poly_stride=0, seed=0, offset=0):
        nl = num_limbs if num_limbs is not None else (len(prime_ids) if prime_ids is not None else d.shape[0])
        _check(load().fh_fill_uniform(self.h, d.addr(offset), _i32(prime_ids), nl, num_polys, poly_stride, seed))

    def linear_transform(self, ct, level, row_size, b_step, stride, diags, baby_keys, baby_galois, out, giant_key=None, giant_galois=1,
                         offset_key=None, offset_galois=1, offset=0, slots=None):
        slots = slots or self.N // 2
        bk = (C.c_void_p * b_step)(*[(k.h if k is not None else None) for k in baby_keys])
        bg = (C.c_uint32 * b_step)(*[int(g) for g in baby_galois])
        _check(load().fh_linear_transform(self.h, ct.ptr, level, row_size, b_step, stride, offset, slots, diags.ptr, bk, bg,
                                          giant_key.h if giant_key is not None else None, giant_galois,
                                          offset_key.h if offset_key is not None else None, offset_galois, out.ptr))

    def linear_transform_flat(self, ct, level, row_size, b_step, stride, diags, baby_keys, baby_galois, out, giant_keys=(),
                              giant_galois=(), offset_key=None, offset_galois=1, offset=0, slots=None):
        """The transform with giant steps as independent rotations: giant_keys[j-1] rotates the
        j-th partial sum by j*b_step*stride (None where that is 0 mod slots)."""
        slots = slots or self.N // 2
        bk = (C.c_void_p * b_step)(*[(k.h if k is not None else None) for k in baby_keys])
        bg = (C.c_uint32 * b_step)(*[int(g) for g in baby_galois])
        ng = len(giant_keys)
        gk = (C.c_void_p * ng)(*[(k.h if k is not None else None) for k in giant_keys]) if ng else None
        gg = (C.c_uint32 * ng)(*[int(g) for g in giant_galois]) if ng else None
        _check(load().fh_linear_transform_flat(self.h, ct.ptr, level, row_size, b_step, stride, offset, slots, diags.ptr, bk, bg, gk, gg,
                                               offset_key.h if offset_key is not None else None, offset_galois, out.ptr))

    @staticmethod
    def accumulate_indices(radix, stride, size, start=1):
        n = load().fh_accumulate_indices(radix, stride, size, start, None, 0)
        if n < 0:
            _check(n)
        buf = (C.c_int32 * max(n, 1))()
        _check(min(0, load().fh_accumulate_indices(radix, stride, size, start, buf, n)))
        return list(buf[:n])

    def accumulate_sum(self, ct, level, radix, stride, size, keys, galois, out):
        """The lazy fold; keys / galois: one per index of accumulate_indices(radix, stride, size)."""
        n = len(keys)
        ka = (C.c_void_p * max(n, 1))(*[k.h for k in keys])
        ga = (C.c_uint32 * max(n, 1))(*[int(g) for g in galois])
        _check(load().fh_accumulate_sum(self.h, ct.ptr, level, radix, stride, size, ka, ga, n, out.ptr))

    def accumulate_sum_from(self, ct, level, radix, stride, size, start, keys, galois, out):
        """The cascade from a start factor; keys / galois: one per index of
        accumulate_indices(radix, stride, size, start)."""
        n = len(keys)
        ka = (C.c_void_p * max(n, 1))(*[k.h for k in keys])
        ga = (C.c_uint32 * max(n, 1))(*[int(g) for g in galois])
        _check(load().fh_accumulate_sum_from(self.h, ct.ptr, level, radix, stride, size, start, ka, ga, n, out.ptr))

    def linear_transform_batch(self, cts, count, level, row_size, b_step, stride, diags, baby_keys, baby_galois, out, giant_key=None,
                               giant_galois=1, offset_key=None, offset_galois=1, offset=0, slots=None, ct_stride=0):
        slots = slots or self.N // 2
        bk = (C.c_void_p * b_step)(*[(k.h if k is not None else None) for k in baby_keys])
        bg = (C.c_uint32 * b_step)(*[int(g) for g in baby_galois])
        _check(load().fh_linear_transform_batch(self.h, count, cts.ptr, ct_stride, level, row_size, b_step, stride, offset, slots,
                                                diags.ptr, bk, bg, giant_key.h if giant_key is not None else None, giant_galois,
                                                offset_key.h if offset_key is not None else None, offset_galois, out.ptr))

    # ---- serving loop ----
    def serve(self, rot_keys, rot_galois, relin_key, rate_per_s, requests, start_level=None, mode=0, hbm_budget_gb=64.0,
              inference_share=0.74, seed=1, checksum=False, bootstrap_share=0.0, request_checksums=False, request_outputs=False,
              ticks_in_flight=0, batch_min_ops=0, bootstrap_concurrent=0):
        """bootstrap_share > 0 needs a bootstrap function, which only C/C++ callers pass (the C++
        facade's EvalBootstrap, tests/cpp/serve_check.cpp); from here it is rejected. mode 0:
        batched ticks, 1: one call per operation, 2: one request at a time. request_checksums:
        also return each request's result-word sum (list indexed by request). request_outputs: also
        return each request's whole result (numpy arrays indexed by request). ticks_in_flight:
        scheduler ticks issued ahead of the GPU (0: the default, 1: wait for every tick).
        batch_min_ops: mode 0's smallest tick run as one batch (0: the default, 3).
        bootstrap_concurrent: where bootstrap requests run (0 in the tick, 1 on a worker thread, 2
        on a worker thread that yields to the ticks); validated even without bootstrap requests."""
        per = (C.c_uint64 * max(1, requests))() if request_checksums else None
        cap = 2 * self.L1 * self.N
        bufs = [np.zeros(cap, dtype=np.uint64) for _ in range(requests)] if request_outputs else None
        ptrs = (C.c_void_p * max(1, requests))(*[b.ctypes.data for b in bufs]) if bufs else None
        words = (C.c_uint64 * max(1, requests))() if bufs else None
        cfg = ServeConfig(rate_per_s, requests, self.L1 - 1 if start_level is None else start_level, hbm_budget_gb, mode,
                          inference_share, seed, int(checksum), float(bootstrap_share), None, None,
                          C.cast(per, C.POINTER(C.c_uint64)) if per is not None else None,
                          C.cast(ptrs, C.POINTER(C.c_void_p)) if ptrs is not None else None, cap if bufs else 0,
                          C.cast(words, C.POINTER(C.c_uint64)) if words is not None else None, int(bootstrap_concurrent), int(ticks_in_flight),
                          int(batch_min_ops))
        kp = (C.c_void_p * 8)(*[k.h for k in rot_keys])
        gl = (C.c_uint32 * 8)(*[int(g) for g in rot_galois])
        r = ServeResult()
        _check(load().fh_serve_run(self.h, C.byref(cfg), kp, gl, relin_key.h, C.byref(r)))
        out = {f: getattr(r, f) for f, _ in ServeResult._fields_}
        if per is not None:
            out["request_checksums"] = [int(per[i]) for i in range(requests)]
        if bufs is not None:
            out["request_outputs"] = [bufs[i][:int(words[i])] for i in range(requests)]
        return out

    # ---- launch-shape tunables ----
    def tune_set(self, name, value):
        _check(load().fh_tune_set(self.h, name.encode(), int(value)))

    def tune_get(self, name):
        v = C.c_int()
        _check(load().fh_tune_get(self.h, name.encode(), C.byref(v)))
        return v.value

    def autotune(self, polys):
        """Choose the NTT launch shape by timing candidates on `polys` synthetic polynomials."""
        r = AutotuneReport()
        _check(load().fh_autotune(self.h, int(polys), C.byref(r)))
        return {"split": r.split, "threads": {"fwd_cols": r.threads[0], "fwd_rows": r.threads[1], "inv_rows": r.threads[2],
                                              "inv_cols": r.threads[3]},
                "chunk_mb": r.chunk_mb, "candidates": r.candidates, "limbs": r.limbs, "ms_default": round(r.ms_default, 4),
                "ms_best": round(r.ms_best, 4)}

    # ---- profiling (per-kernel-family HIP event timing) ----
    def profile(self, on=True):
        _check(load().fh_profile_enable(self.h, int(on)))

    def profile_reset(self):
        _check(load().fh_profile_reset(self.h))

    def profile_read(self, name):
        n, t = C.c_int(), C.c_double()
        _check(load().fh_profile_read(self.h, name.encode(), C.byref(n), C.byref(t)))
        return n.value, t.value


class Batch:
    """A mixed-level batch of EvalRotate / EvalMult(+Rescale) / EvalAdd items (fh_batch_*).

    Items name device addresses (ints, e.g. DeviceArray.addr()); finalize() uploads the job
    tables, run() issues the whole batch, graph=True captures it once and replays it."""

    def __init__(self, ctx):
        h = C.c_void_p()
        _check(load().fh_batch_create(ctx.h, C.byref(h)))
        self.ctx, self.h = ctx, h.value

    def rotate(self, ct_addr, level, key, galois, out_addr):
        _check(load().fh_batch_rotate(self.h, ct_addr, level, key.h, galois, out_addr))

    def mult(self, a_addr, b_addr, level, relin_key, rescale, out_addr):
        _check(load().fh_batch_mult(self.h, a_addr, b_addr, level, relin_key.h, int(rescale), out_addr))

    def mult_strided(self, a_addr, towers_a, b_addr, towers_b, level, relin_key, rescale, out_addr):
        """a / b hold towers_a / towers_b limbs per component (>= level + 1; the first level + 1 are read)"""
        _check(load().fh_batch_mult_strided(self.h, a_addr, towers_a, b_addr, towers_b, level, relin_key.h, int(rescale), out_addr))

    def add(self, a_addr, b_addr, level, ncomp, out_addr):
        _check(load().fh_batch_add(self.h, a_addr, b_addr, level, ncomp, out_addr))

    def mult_plain(self, a_addr, pt_addr, level, rescale, out_addr):
        _check(load().fh_batch_mult_plain(self.h, a_addr, pt_addr, level, int(rescale), out_addr))

    def add_plain(self, a_addr, pt_addr, level, out_addr):
        _check(load().fh_batch_add_plain(self.h, a_addr, pt_addr, level, out_addr))

    def level_reduce(self, a_addr, level_in, level_out, ncomp, out_addr):
        _check(load().fh_batch_level_reduce(self.h, a_addr, level_in, level_out, ncomp, out_addr))

    def reset(self):
        _check(load().fh_batch_reset(self.h))

    def finalize(self):
        _check(load().fh_batch_finalize(self.h))

    def run(self, graph=False):
        _check(load().fh_batch_run(self.h, int(graph)))

    def info(self):
        n, b = C.c_size_t(), C.c_size_t()
        _check(load().fh_batch_info(self.h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def free(self):
        if self.h:
            _check(load().fh_batch_destroy(self.h))
            self.h = None


class Key:
    """A key-switching key resident on the device ([dnum][b, a][L1+K][N] upload layout)."""

    def __init__(self, ctx, host_key):
        host_key = np.ascontiguousarray(host_key, dtype=np.uint64)
        assert host_key.size == ctx.dnum * 2 * (ctx.L1 + ctx.K) * ctx.N
        h = C.c_void_p()
        _check(load().fh_key_create(ctx.h, _u64p(host_key), C.byref(h)))
        self.ctx, self.h = ctx, h.value

    def free(self):
        if self.h:
            _check(load().fh_key_destroy(self.ctx.h, self.h))
            self.h = None
