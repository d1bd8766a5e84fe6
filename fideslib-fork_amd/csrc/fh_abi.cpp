// extern "C" boundary of the engine (declared in include/fideslib_hip.h).
// Every call converts C++ exceptions into a status code plus a thread-local
// message, the same contract as the reference's CudaError / OPENFHE_THROW
// surfaced through an int (src/CudaUtils.cuh:706-750).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fideslib_hip.h"
#include "fh_batch.hpp"
#include "fh_keyswitch.hpp"
#include "fh_serve.hpp"

struct fh_context {
  fh::Context* c;
};

struct fh_key {
  fh::DBuf data;
  const fh::Context* owner;
};

namespace {

thread_local std::string g_err;
thread_local fh::CallGate g_gate;

template <class F>
int guard(F&& f) {
  try {
    f();
    return FH_OK;
  } catch (const fh::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return FH_E_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return FH_E_INTERNAL;
  }
}

}  // namespace

fh::CallGate& fh::thread_call_gate() { return g_gate; }

namespace {

fh::Context& C(fh_context* h) {
  if (!h || !h->c) throw fh::Error(FH_E_INVALID, "null context");
  const fh::CallGate& g = fh::thread_call_gate();
  if (g.fn) g.fn(g.user, *h->c);
  return *h->c;
}
const fh::Context& C(const fh_context* h) {
  if (!h || !h->c) throw fh::Error(FH_E_INVALID, "null context");
  return *h->c;
}

void need(const void* p, const char* what) {
  if (!p) throw fh::Error(FH_E_INVALID, std::string("null pointer: ") + what);
}

std::vector<int> pids_or_identity(const fh::Context& c, const int32_t* ids, int n) {
  if (n < 0 || n > fh::kMaxLimbs) throw fh::Error(FH_E_INVALID, "num_limbs out of range");
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) {
    v[i] = ids ? ids[i] : i;
    if (v[i] < 0 || v[i] >= c.np) throw fh::Error(FH_E_INVALID, "prime id out of range");
  }
  return v;
}

void check_level(const fh::Context& c, int level) {
  if (level < 0 || level >= c.L1) throw fh::Error(FH_E_INVALID, "level out of range");
}

// An automorphism index must be a unit modulo 2N (odd) below 2N: anything else is not a bijection of
// the slots, and the key inner product's scatter would leave outputs unwritten.
void check_galois(const fh::Context& c, uint32_t g, const char* what) {
  if ((g & 1) == 0 || g >= 2 * c.N) throw fh::Error(FH_E_INVALID, std::string(what) + ": galois index must be odd and < 2N");
}

}  // namespace

extern "C" {

const char* fh_last_error(void) { return g_err.c_str(); }
int fh_version(void) { return (1 << 16) | 0; }

int fh_context_create(int log_n, int num_q, int num_p, int dnum, const uint64_t* moduli, fh_context** out) {
  return guard([&] {
    need(moduli, "moduli");
    need(out, "out");
    std::vector<uint64_t> m(moduli, moduli + num_q + num_p);
    auto* h = new fh_context{nullptr};
    try {
      h->c = new fh::Context(log_n, num_q, num_p, dnum, m);
      fh::noalloc_init(*h->c);
    } catch (...) {
      delete h->c;
      delete h;
      throw;
    }
    *out = h;
  });
}

int fh_context_create_generated(int log_n, int num_q, int num_p, int dnum, int first_bits, int scale_bits, int aux_bits, fh_context** out) {
  return guard([&] {
    need(out, "out");
    if (log_n < 10 || log_n > 16 || num_q < 1 || num_q > fh::kMaxPrimes || dnum < 1) throw fh::Error(FH_E_PARAM, "bad parameters");
    if (first_bits < 20 || first_bits > 61 || scale_bits < 20 || scale_bits > 61 || aux_bits < 20 || aux_bits > 61)
      throw fh::Error(FH_E_PARAM, "modulus sizes must be in [20, 61] bits");
    int K = num_p;
    if (K <= 0) {
      auto q = fh::h_gen_primes(log_n, num_q, first_bits, scale_bits, 0, aux_bits);
      K = fh::h_default_num_special(q, num_q, dnum, aux_bits);
    }
    auto m = fh::h_gen_primes(log_n, num_q, first_bits, scale_bits, K, aux_bits);
    auto* h = new fh_context{nullptr};
    try {
      h->c = new fh::Context(log_n, num_q, K, dnum, m);
      fh::noalloc_init(*h->c);
    } catch (...) {
      delete h->c;
      delete h;
      throw;
    }
    *out = h;
  });
}

int fh_context_destroy(fh_context* ctx) {
  return guard([&] {
    if (!ctx) return;
    delete ctx->c;
    delete ctx;
  });
}

int fh_context_set_reduced_noise(fh_context* ctx, int on) {
  return guard([&] {
    auto& c = C(ctx);
    if (on && !c.rn_conv_ok) throw fh::Error(FH_E_UNSUPPORTED, "ReducedNoise conversions need (alpha + 1) * max prime <= 2^64");
    FH_HIP(hipStreamSynchronize(c.stream));  // work already queued keeps the mode it was issued with
    c.reduced_noise = on != 0;
  });
}

int fh_context_shape(const fh_context* ctx, int* out5) {
  return guard([&] {
    const auto& c = C(ctx);
    need(out5, "out");
    out5[0] = c.logN; out5[1] = c.L1; out5[2] = c.K; out5[3] = c.dnum; out5[4] = c.alpha;
  });
}

int fh_context_moduli(const fh_context* ctx, uint64_t* moduli, uint64_t* psi) {
  return guard([&] {
    const auto& c = C(ctx);
    for (int i = 0; i < c.np; ++i) {
      if (moduli) moduli[i] = c.q[i];
      if (psi) psi[i] = c.psi[i];
    }
  });
}

int fh_synchronize(fh_context* ctx) {
  return guard([&] {
    auto& c = C(ctx);
    FH_HIP(hipStreamSynchronize(c.stream));
    // a team NTT launch whose barrier wait ran past its bound left its results incomplete
    uint32_t err = 0;
    FH_HIP(hipMemcpy(&err, c.team_err(), sizeof(err), hipMemcpyDeviceToHost));
    if (err) {
      FH_HIP(hipMemset(c.team_err(), 0, sizeof(err)));
      throw fh::Error(FH_E_DEVICE, "a team NTT barrier wait timed out: results since the last synchronization are invalid");
    }
  });
}

int fh_device_alloc(fh_context* ctx, size_t bytes, void** out) {
  return guard([&] {
    need(out, "out");
    *out = C(ctx).pool.get(bytes);
  });
}

int fh_device_free(fh_context* ctx, void* ptr) {
  return guard([&] {
    auto& c = C(ctx);
    // no host wait: the block is reused only after the work queued on the context stream so far
    // (the last that may read or write it) has completed
    c.pool.put_after(ptr, c.stream);
  });
}

int fh_copy_h2d(fh_context* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    auto& c = C(ctx);
    need(dst, "dst"); need(src, "src");
    FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipStreamSynchronize(c.stream));
  });
}

int fh_copy_d2h(fh_context* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    auto& c = C(ctx);
    need(dst, "dst"); need(src, "src");
    FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    FH_HIP(hipStreamSynchronize(c.stream));
  });
}

int fh_copy_d2d(fh_context* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    auto& c = C(ctx);
    need(dst, "dst"); need(src, "src");
    FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
  });
}

int fh_stream_copy(fh_context* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    auto& c = C(ctx);
    need(dst, "dst"); need(src, "src");
    fh::stream_copy(c, dst, src, bytes, c.stream);
  });
}

int fh_ntt_forward(fh_context* ctx, uint64_t* polys, const int32_t* prime_ids, int num_limbs, int num_polys, size_t poly_stride) {
  return guard([&] {
    auto& c = C(ctx);
    need(polys, "polys");
    if (num_polys < 1) throw fh::Error(FH_E_INVALID, "num_polys must be >= 1");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    uint64_t stride = poly_stride ? poly_stride : (uint64_t)num_limbs * c.N;
    fh::ntt_fwd(c, fh::make_set(polys, num_limbs, p.data(), num_polys, stride), c.stream);
  });
}

int fh_encode_slots(fh_context* ctx, const double* values, int count, int slots, double scale, const int32_t* prime_ids, int num_limbs,
                    uint64_t* out, size_t out_stride) {
  return guard([&] {
    auto& c = C(ctx);
    need(values, "values"); need(out, "out");
    if (count < 1) throw fh::Error(FH_E_INVALID, "count must be >= 1");
    if (num_limbs < 1) throw fh::Error(FH_E_INVALID, "num_limbs must be >= 1");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    const uint64_t stride = out_stride ? out_stride : (uint64_t)num_limbs * c.N;
    fh::encode_slots(c, values, count, (uint32_t)slots, scale, fh::make_set(out, num_limbs, p.data(), count, stride), c.stream);
  });
}

int fh_encode_coeffs(fh_context* ctx, const int64_t* coeffs, int count, const int32_t* prime_ids, int num_limbs, uint64_t* out,
                     size_t out_stride) {
  return guard([&] {
    auto& c = C(ctx);
    need(coeffs, "coeffs"); need(out, "out");
    if (count < 1) throw fh::Error(FH_E_INVALID, "count must be >= 1");
    if (num_limbs < 1) throw fh::Error(FH_E_INVALID, "num_limbs must be >= 1");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    const uint64_t stride = out_stride ? out_stride : (uint64_t)num_limbs * c.N;
    fh::encode_coeffs(c, coeffs, count, fh::make_set(out, num_limbs, p.data(), count, stride), c.stream);
  });
}

int fh_ntt_inverse(fh_context* ctx, uint64_t* polys, const int32_t* prime_ids, int num_limbs, int num_polys, size_t poly_stride) {
  return guard([&] {
    auto& c = C(ctx);
    need(polys, "polys");
    if (num_polys < 1) throw fh::Error(FH_E_INVALID, "num_polys must be >= 1");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    uint64_t stride = poly_stride ? poly_stride : (uint64_t)num_limbs * c.N;
    fh::ntt_inv(c, fh::make_set(polys, num_limbs, p.data(), num_polys, stride), c.stream);
  });
}

int fh_eltwise_binary(fh_context* ctx, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, const int32_t* prime_ids, int num_limbs) {
  return guard([&] {
    auto& c = C(ctx);
    need(a, "a"); need(b, "b"); need(out, "out");
    if (op < 0 || op > 2) throw fh::Error(FH_E_INVALID, "unknown element-wise op");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    fh::ew_binary(c, op, fh::make_set(const_cast<uint64_t*>(a), num_limbs, p.data()),
                  fh::make_set(const_cast<uint64_t*>(b), num_limbs, p.data()), fh::make_set(out, num_limbs, p.data()), c.stream);
  });
}

int fh_eltwise_scalar(fh_context* ctx, int op, const uint64_t* a, const uint64_t* scalars, uint64_t* out, const int32_t* prime_ids, int num_limbs) {
  return guard([&] {
    auto& c = C(ctx);
    need(a, "a"); need(out, "out");
    if (op < 0 || op > 2) throw fh::Error(FH_E_INVALID, "unknown scalar op");
    if (op != 2) need(scalars, "scalars");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    fh::ew_unary(c, op, fh::make_set(const_cast<uint64_t*>(a), num_limbs, p.data()), fh::make_set(out, num_limbs, p.data()), scalars, c.stream);
  });
}

int fh_linear_wsum(fh_context* ctx, int n, const uint64_t* const* cts, const int32_t* cts_towers, const uint64_t* weights, int towers,
                   int ncomp, const uint64_t* add, uint32_t add_components, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(cts, "cts"); need(cts_towers, "cts_towers"); need(weights, "weights"); need(out, "out");
    if (n < 1 || n > fh::kMaxWSum) throw fh::Error(FH_E_INVALID, "n must be 1..16");
    std::vector<uint32_t> limbs(n);
    for (int k = 0; k < n; ++k) {
      if (cts_towers[k] < towers) throw fh::Error(FH_E_INVALID, "an input has fewer towers than the output");
      limbs[k] = (uint32_t)cts_towers[k];
    }
    fh::ew_wsum(c, n, cts, limbs.data(), weights, towers, ncomp, add, add_components, out, c.stream);
  });
}

int fh_tensor(fh_context* ctx, const uint64_t* ct1, const uint64_t* ct2, int level, uint64_t* out3) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct1, "ct1"); need(ct2, "ct2"); need(out3, "out3");
    check_level(c, level);
    fh::ew_tensor(c, ct1, ct2, out3, level + 1, c.stream);
  });
}

int fh_rescale(fh_context* ctx, const uint64_t* ct, int level, int ncomp, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level);
    if (level < 1) throw fh::Error(FH_E_INVALID, "cannot rescale at level 0");
    if (ncomp < 1 || ncomp > 8) throw fh::Error(FH_E_INVALID, "ncomp must be 1..8");
    fh::rescale(c, ct, level, ncomp, out, c.stream);
  });
}

int fh_rescale_noalloc(fh_context* ctx, const uint64_t* ct, int level, int ncomp, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level);
    if (level < 1) throw fh::Error(FH_E_INVALID, "cannot rescale at level 0");
    if (ncomp < 1 || ncomp > 8) throw fh::Error(FH_E_INVALID, "ncomp must be 1..8");
    const size_t in_words = (size_t)ncomp * (level + 1) * c.N, out_words = (size_t)ncomp * level * c.N;
    if (ct < out + out_words && out < ct + in_words) throw fh::Error(FH_E_INVALID, "fh_rescale_noalloc cannot run in place: ct and out overlap");
    fh::rescale_noalloc(c, ct, level, ncomp, out, c.stream);
  });
}

int fh_mod_raise(fh_context* ctx, const uint64_t* ct, int level_in, int ncomp, int level, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level_in);
    check_level(c, level);
    if (ncomp < 1 || ncomp > 3) throw fh::Error(FH_E_INVALID, "ncomp must be 1..3");
    fh::mod_raise(c, ct, level_in, ncomp, level, out, c.stream);
  });
}

int fh_key_create(fh_context* ctx, const uint64_t* host_key, fh_key** out) {
  return guard([&] {
    auto& c = C(ctx);
    need(host_key, "host_key"); need(out, "out");
    size_t words = (size_t)c.dnum * 2 * c.np * c.N;
    auto* k = new fh_key{fh::DBuf(c.pool, words), &c};
    // host or device source (unified addressing resolves the direction)
    hipError_t e = hipMemcpyAsync(k->data.p, host_key, words * 8, hipMemcpyDefault, c.stream);
    if (e == hipSuccess) {
      // resident keys are kept in Montgomery form: the inner products reduce with one REDC
      std::vector<int> ids(c.np);
      for (int i = 0; i < c.np; ++i) ids[i] = i;
      fh::to_montgomery(c, fh::make_set(k->data.p, c.np, ids.data(), 2 * c.dnum, (uint64_t)c.np * c.N), c.stream);
      e = hipStreamSynchronize(c.stream);
    }
    if (e != hipSuccess) {
      delete k;
      throw fh::Error(FH_E_DEVICE, std::string("key upload failed: ") + hipGetErrorString(e));
    }
    *out = k;
  });
}

int fh_key_export(fh_context* ctx, const fh_key* key, uint64_t* host_key) {
  return guard([&] {
    auto& c = C(ctx);
    need(key, "key"); need(host_key, "host_key");
    if (key->owner != &c) throw fh::Error(FH_E_INVALID, "key belongs to another context");
    const size_t words = (size_t)c.dnum * 2 * c.np * c.N;
    FH_HIP(hipStreamSynchronize(c.stream));
    FH_HIP(hipMemcpy(host_key, key->data.p, words * 8, hipMemcpyDeviceToHost));
    // out of the resident Montgomery form: v (2^64)^-1 mod q per limb
    std::vector<uint64_t> rinv(c.np);
    for (int p = 0; p < c.np; ++p) {
      const uint64_t q = c.q[p];
      const uint64_t r = (uint64_t)(((unsigned __int128)1 << 64) % q);
      rinv[p] = fh::h_invmod(r, q);
    }
    for (size_t limb = 0; limb < (size_t)c.dnum * 2 * c.np; ++limb) {
      const int p = (int)(limb % c.np);
      uint64_t* v = host_key + limb * c.N;
      for (uint32_t i = 0; i < c.N; ++i) v[i] = fh::h_mulmod(v[i], rinv[p], c.q[p]);
    }
  });
}

int fh_key_destroy(fh_context* ctx, fh_key* key) {
  return guard([&] {
    if (!key) return;
    auto& c = C(ctx);
    FH_HIP(hipStreamSynchronize(c.stream));
    delete key;
  });
}

static const uint64_t* key_of(const fh::Context& c, const fh_key* k) {
  if (!k) throw fh::Error(FH_E_INVALID, "null key");
  if (k->owner != &c) throw fh::Error(FH_E_INVALID, "key belongs to another context");
  return k->data.p;
}

int fh_num_digits(const fh_context* ctx, int level) {
  try {
    const auto& c = C(ctx);
    check_level(c, level);
    return c.num_digits(level);
  } catch (const fh::Error& e) {
    g_err = e.what();
    return e.code;
  }
}

int fh_modup(fh_context* ctx, const uint64_t* poly, int level, uint64_t* digits) {
  return guard([&] {
    auto& c = C(ctx);
    need(poly, "poly"); need(digits, "digits");
    check_level(c, level);
    fh::modup(c, poly, level, digits, c.stream);
  });
}

int fh_key_inner_product(fh_context* ctx, const uint64_t* digits, int level, const fh_key* key, const uint64_t* c0,
                         uint32_t galois, uint64_t* out0, uint64_t* out1) {
  return guard([&] {
    auto& c = C(ctx);
    need(digits, "digits"); need(out0, "out0"); need(out1, "out1");
    check_level(c, level);
    if ((galois & 1) == 0 || galois >= 2 * c.N) throw fh::Error(FH_E_INVALID, "galois index must be odd and < 2N");
    fh::ksk_inner(c, digits, level, key_of(c, key), out0, out1, c0, galois, c.stream);
  });
}

int fh_moddown(fh_context* ctx, const uint64_t* poly_ext, int level, int num_polys, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(poly_ext, "poly_ext"); need(out, "out");
    check_level(c, level);
    if (num_polys < 1) throw fh::Error(FH_E_INVALID, "num_polys must be >= 1");
    fh::moddown(c, poly_ext, level, out, num_polys, 0, 0, c.stream);
  });
}

int fh_automorph(fh_context* ctx, const uint64_t* poly, int num_limbs, uint32_t galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(poly, "poly"); need(out, "out");
    if (num_limbs < 1) throw fh::Error(FH_E_INVALID, "num_limbs must be >= 1");
    fh::automorph(c, poly, num_limbs, galois, out, c.stream);
  });
}

int fh_galois_index(const fh_context* ctx, int32_t rotation, uint32_t* out) {
  return guard([&] {
    const auto& c = C(ctx);
    need(out, "out");
    uint64_t m = 2ull * c.N;
    if (rotation == 0) { *out = 1; return; }
    uint64_t g = rotation > 0 ? 5 : fh::h_inv_odd(5, m);
    uint64_t r = rotation > 0 ? (uint64_t)rotation : (uint64_t)(-(int64_t)rotation);
    *out = (uint32_t)fh::h_powmod(g, (uint64_t)r, m);
  });
}

int fh_relinearize(fh_context* ctx, const uint64_t* ct3, int level, const fh_key* relin_key, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct3, "ct3"); need(out, "out");
    check_level(c, level);
    fh::relinearize(c, ct3, level, key_of(c, relin_key), out, c.stream);
  });
}

int fh_eval_mult(fh_context* ctx, const uint64_t* ct1, const uint64_t* ct2, int level, const fh_key* relin_key, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct1, "ct1"); need(ct2, "ct2"); need(out, "out");
    check_level(c, level);
    fh::eval_mult(c, ct1, ct2, level, key_of(c, relin_key), out, c.stream);
  });
}

int fh_eval_mult_many(fh_context* ctx, int n, const uint64_t* ct1, int towers1, const uint64_t* ct2, int towers2, int level,
                      const fh_key* relin_key, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct1, "ct1"); need(ct2, "ct2"); need(out, "out");
    check_level(c, level);
    if (n < 1 || n > 4096) throw fh::Error(FH_E_INVALID, "n must be 1..4096");
    if (towers1 < level + 1 || towers2 < level + 1 || towers1 > c.L1 || towers2 > c.L1)
      throw fh::Error(FH_E_INVALID, "operand towers must be level+1..num_q");
    fh::eval_mult_many(c, n, ct1, towers1, ct2, towers2, level, key_of(c, relin_key), out, c.stream);
  });
}

int fh_eval_rotate(fh_context* ctx, const uint64_t* ct, int level, const fh_key* rot_key, uint32_t galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level);
    if ((galois & 1) == 0 || galois >= 2 * c.N) throw fh::Error(FH_E_INVALID, "galois index must be odd and < 2N");
    fh::eval_rotate(c, ct, level, key_of(c, rot_key), galois, out, c.stream);
  });
}

int fh_eval_rotate_ext(fh_context* ctx, const uint64_t* ct, int level, const fh_key* rot_key, uint32_t galois, int add_c0,
                       uint64_t* out_ext) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out_ext, "out_ext");
    check_level(c, level);
    if ((galois & 1) == 0 || galois >= 2 * c.N) throw fh::Error(FH_E_INVALID, "galois index must be odd and < 2N");
    fh::rotate_ext(c, ct, level, key_of(c, rot_key), galois, add_c0 != 0, out_ext, c.stream);
  });
}

int fh_rotate_batch(fh_context* ctx, int count, const uint64_t* cts, size_t ct_stride, int level, const fh_key* const* keys,
                    const uint32_t* galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(cts, "cts"); need(keys, "keys"); need(galois, "galois"); need(out, "out");
    check_level(c, level);
    if (count < 0) throw fh::Error(FH_E_INVALID, "count must be >= 0");
    std::vector<const uint64_t*> kp(count);
    for (int i = 0; i < count; ++i) {
      kp[i] = key_of(c, keys[i]);
      check_galois(c, galois[i], "fh_rotate_batch");
    }
    fh::rotate_batch(c, count, cts, ct_stride, level, kp.data(), galois, true, true, out, c.stream);
  });
}

int fh_rotate_hoisted(fh_context* ctx, const uint64_t* ct, int level, int n, const fh_key* const* keys, const uint32_t* galois,
                      int add_c0, int moddown, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(keys, "keys"); need(galois, "galois"); need(out, "out");
    check_level(c, level);
    if (n < 0) throw fh::Error(FH_E_INVALID, "n must be >= 0");
    std::vector<const uint64_t*> kp(n);
    for (int i = 0; i < n; ++i) {
      kp[i] = key_of(c, keys[i]);
      check_galois(c, galois[i], "fh_rotate_hoisted");
    }
    fh::rotate_hoisted(c, ct, level, n, kp.data(), galois, add_c0 != 0, moddown != 0, out, c.stream);
  });
}

int fh_linear_transform(fh_context* ctx, const uint64_t* ct, int level, int row_size, int b_step, int stride, int offset, int slots,
                        const uint64_t* diags, const fh_key* const* baby_keys, const uint32_t* baby_galois, const fh_key* giant_key,
                        uint32_t giant_galois, const fh_key* offset_key, uint32_t offset_galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(diags, "diags"); need(out, "out");
    check_level(c, level);
    if (b_step < 1 || b_step > 16) throw fh::Error(FH_E_INVALID, "b_step must be in [1, 16]");
    if (slots < 1 || slots > (int)c.N / 2) throw fh::Error(FH_E_INVALID, "slots out of range");
    std::vector<const uint64_t*> bk(b_step, nullptr);
    std::vector<uint32_t> bg(b_step, 1);
    for (int i = 0; i < b_step; ++i) {
      if (((long long)i * stride) % slots == 0) continue;
      need(baby_keys, "baby_keys");
      need(baby_galois, "baby_galois");
      bk[i] = key_of(c, baby_keys[i]);
      bg[i] = baby_galois[i];
      check_galois(c, bg[i], "baby step");
    }
    const uint64_t* gk = giant_key ? key_of(c, giant_key) : nullptr;
    const uint64_t* ok = offset_key ? key_of(c, offset_key) : nullptr;
    if (gk) check_galois(c, giant_galois, "giant step");
    if (ok) check_galois(c, offset_galois, "offset");
    fh::linear_transform(c, ct, level, row_size, b_step, stride, offset, slots, diags, bk.data(), bg.data(), gk, giant_galois, ok,
                         offset_galois, out, c.stream);
  });
}

int fh_linear_transform_flat(fh_context* ctx, const uint64_t* ct, int level, int row_size, int b_step, int stride, int offset,
                             int slots, const uint64_t* diags, const fh_key* const* baby_keys, const uint32_t* baby_galois,
                             const fh_key* const* giant_keys, const uint32_t* giant_galois, const fh_key* offset_key,
                             uint32_t offset_galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(diags, "diags"); need(out, "out");
    check_level(c, level);
    if (b_step < 1 || b_step > 16) throw fh::Error(FH_E_INVALID, "b_step must be in [1, 16]");
    if (row_size < 1) throw fh::Error(FH_E_INVALID, "row_size must be >= 1");
    if (slots < 1 || slots > (int)c.N / 2) throw fh::Error(FH_E_INVALID, "slots out of range");
    std::vector<const uint64_t*> bk(b_step, nullptr);
    std::vector<uint32_t> bg(b_step, 1);
    for (int i = 0; i < b_step; ++i) {
      if (((long long)i * stride) % slots == 0) continue;
      need(baby_keys, "baby_keys");
      need(baby_galois, "baby_galois");
      bk[i] = key_of(c, baby_keys[i]);
      bg[i] = baby_galois[i];
      check_galois(c, bg[i], "baby step");
    }
    const int g_step = (row_size + b_step - 1) / b_step;
    std::vector<const uint64_t*> gk(g_step > 1 ? g_step - 1 : 0, nullptr);
    std::vector<uint32_t> gg(gk.size(), 1);
    for (int j = 1; j < g_step; ++j) {
      if (((long long)j * b_step * stride) % slots == 0) continue;
      need(giant_keys, "giant_keys");
      need(giant_galois, "giant_galois");
      gk[j - 1] = key_of(c, giant_keys[j - 1]);
      gg[j - 1] = giant_galois[j - 1];
      check_galois(c, gg[j - 1], "giant step");
    }
    const uint64_t* ok = offset_key ? key_of(c, offset_key) : nullptr;
    if (ok) check_galois(c, offset_galois, "offset");
    fh::linear_transform_flat(c, ct, level, row_size, b_step, stride, offset, slots, diags, bk.data(), bg.data(), gk.data(), gg.data(),
                              ok, offset_galois, out, c.stream);
  });
}

}  // extern "C"

namespace {
// fold rotation indices; throws on an invalid radix or stride
std::vector<int32_t> fold_indices(int radix, int stride, int size, int start) {
  if (radix < 2 || radix > 64 || (radix & (radix - 1))) throw fh::Error(FH_E_INVALID, "radix must be a power of two in [2, 64]");
  if (stride < 1) throw fh::Error(FH_E_INVALID, "stride must be >= 1");
  std::vector<int32_t> v;
  if (start <= 0 || size <= 0) return v;
  for (long long s = start; s < size; s *= radix)
    for (long long i = 1; i < radix && i * s < size; ++i) {
      const long long idx = i * s * stride;
      if (idx > INT32_MAX) throw fh::Error(FH_E_INVALID, "rotation index out of range");
      v.push_back((int32_t)idx);
    }
  return v;
}
}  // namespace

extern "C" {

int fh_accumulate_indices(int radix, int stride, int size, int start, int32_t* out, int capacity) {
  int n = 0;
  int rc = guard([&] {
    auto v = fold_indices(radix, stride, size, start);
    if (out)
      for (size_t i = 0; i < v.size() && (int)i < capacity; ++i) out[i] = v[i];
    n = (int)v.size();
  });
  return rc != FH_OK ? rc : n;
}

namespace {
// one key and Galois element per fold rotation; throws on a mismatch or an invalid element
void fold_keys(fh::Context& c, size_t n, const fh_key* const* keys, const uint32_t* galois, int num_keys, std::vector<const uint64_t*>& kp,
               std::vector<uint32_t>& gl) {
  if (num_keys != (int)n) throw fh::Error(FH_E_INVALID, "num_keys must equal the fold's rotation count");
  kp.resize(n);
  gl.resize(n);
  for (size_t i = 0; i < n; ++i) {
    need(keys, "keys"); need(galois, "galois");
    kp[i] = key_of(c, keys[i]);
    gl[i] = galois[i];
    if ((gl[i] & 1) == 0 || gl[i] < 3 || gl[i] >= 2 * c.N) throw fh::Error(FH_E_INVALID, "galois must be odd, not 1, and < 2N");
  }
}
}  // namespace

int fh_accumulate_sum(fh_context* ctx, const uint64_t* ct, int level, int radix, int stride, int size, const fh_key* const* keys,
                      const uint32_t* galois, int num_keys, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level);
    std::vector<const uint64_t*> kp;
    std::vector<uint32_t> gl;
    fold_keys(c, fold_indices(radix, stride, size, 1).size(), keys, galois, num_keys, kp, gl);
    fh::accumulate_sum(c, ct, level, radix, size, kp.data(), gl.data(), out, c.stream);
  });
}

int fh_accumulate_sum_from(fh_context* ctx, const uint64_t* ct, int level, int radix, int stride, int size, int start,
                           const fh_key* const* keys, const uint32_t* galois, int num_keys, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(ct, "ct"); need(out, "out");
    check_level(c, level);
    std::vector<const uint64_t*> kp;
    std::vector<uint32_t> gl;
    fold_keys(c, fold_indices(radix, stride, size, start).size(), keys, galois, num_keys, kp, gl);
    fh::accumulate_cascade(c, ct, level, radix, size, start, kp.data(), gl.data(), out, c.stream);
  });
}

int fh_linear_transform_batch(fh_context* ctx, int count, const uint64_t* cts, size_t ct_stride, int level, int row_size, int b_step,
                              int stride, int offset, int slots, const uint64_t* diags, const fh_key* const* baby_keys,
                              const uint32_t* baby_galois, const fh_key* giant_key, uint32_t giant_galois, const fh_key* offset_key,
                              uint32_t offset_galois, uint64_t* out) {
  return guard([&] {
    auto& c = C(ctx);
    need(cts, "cts"); need(diags, "diags"); need(out, "out");
    check_level(c, level);
    if (count < 0) throw fh::Error(FH_E_INVALID, "count must be >= 0");
    if (b_step < 1 || b_step > 16) throw fh::Error(FH_E_INVALID, "b_step must be in [1, 16]");
    if (slots < 1 || slots > (int)c.N / 2) throw fh::Error(FH_E_INVALID, "slots out of range");
    std::vector<const uint64_t*> bk(b_step, nullptr);
    std::vector<uint32_t> bg(b_step, 1);
    for (int i = 0; i < b_step; ++i) {
      if (((long long)i * stride) % slots == 0) continue;
      need(baby_keys, "baby_keys");
      need(baby_galois, "baby_galois");
      bk[i] = key_of(c, baby_keys[i]);
      bg[i] = baby_galois[i];
      check_galois(c, bg[i], "baby step");
    }
    const uint64_t* gk = giant_key ? key_of(c, giant_key) : nullptr;
    const uint64_t* ok = offset_key ? key_of(c, offset_key) : nullptr;
    if (gk) check_galois(c, giant_galois, "giant step");
    if (ok) check_galois(c, offset_galois, "offset");
    fh::linear_transform_batch(c, count, cts, ct_stride, level, row_size, b_step, stride, offset, slots, diags, bk.data(), bg.data(), gk,
                               giant_galois, ok, offset_galois, out, c.stream);
  });
}

int fh_fill_uniform(fh_context* ctx, uint64_t* polys, const int32_t* prime_ids, int num_limbs, int num_polys, size_t poly_stride,
                    uint64_t seed) {
  return guard([&] {
    auto& c = C(ctx);
    need(polys, "polys");
    if (num_polys < 1) throw fh::Error(FH_E_INVALID, "num_polys must be >= 1");
    auto p = pids_or_identity(c, prime_ids, num_limbs);
    uint64_t stride = poly_stride ? poly_stride : (uint64_t)num_limbs * c.N;
    fh::fill_uniform(c, fh::make_set(polys, num_limbs, p.data(), num_polys, stride), seed, c.stream);
  });
}

struct fh_batch {
  fh_context* ctx;
  fh::Batch b;
};

static fh_batch& B(fh_batch* h) {
  if (!h) throw fh::Error(FH_E_INVALID, "null batch");
  return *h;
}

int fh_batch_create(fh_context* ctx, fh_batch** out) {
  return guard([&] {
    auto& c = C(ctx);
    need(out, "out");
    *out = new fh_batch{ctx, fh::Batch(c)};
  });
}

int fh_batch_destroy(fh_batch* b) {
  return guard([&] {
    if (!b) return;
    // replays may still be in flight on the context stream
    (void)hipStreamSynchronize(C(b->ctx).stream);
    delete b;
  });
}

int fh_batch_rotate(fh_batch* b, const uint64_t* ct, int level, const fh_key* key, uint32_t galois, uint64_t* out) {
  return guard([&] {
    auto& h = B(b);
    auto& c = C(h.ctx);
    h.b.add_rotate(ct, level, key_of(c, key), galois, out);
  });
}

int fh_batch_mult(fh_batch* b, const uint64_t* a, const uint64_t* bb, int level, const fh_key* relin_key, int rescale, uint64_t* out) {
  return guard([&] {
    auto& h = B(b);
    auto& c = C(h.ctx);
    h.b.add_mult(a, bb, level, key_of(c, relin_key), rescale != 0, out);
  });
}

int fh_batch_mult_strided(fh_batch* b, const uint64_t* a, int towers_a, const uint64_t* bb, int towers_b, int level, const fh_key* relin_key,
                          int rescale, uint64_t* out) {
  return guard([&] {
    auto& h = B(b);
    auto& c = C(h.ctx);
    h.b.add_mult(a, bb, level, key_of(c, relin_key), rescale != 0, out, towers_a, towers_b);
  });
}

int fh_batch_add(fh_batch* b, const uint64_t* a, const uint64_t* bb, int level, int ncomp, uint64_t* out) {
  return guard([&] { B(b).b.add_add(a, bb, level, ncomp, out); });
}

int fh_batch_mult_plain(fh_batch* b, const uint64_t* a, const uint64_t* pt, int level, int rescale, uint64_t* out) {
  return guard([&] { B(b).b.add_mult_plain(a, pt, level, rescale != 0, out); });
}

int fh_batch_add_plain(fh_batch* b, const uint64_t* a, const uint64_t* pt, int level, uint64_t* out) {
  return guard([&] { B(b).b.add_add_plain(a, pt, level, out); });
}

int fh_batch_level_reduce(fh_batch* b, const uint64_t* a, int level_in, int level_out, int ncomp, uint64_t* out) {
  return guard([&] { B(b).b.add_level_reduce(a, level_in, level_out, ncomp, out); });
}

int fh_batch_reset(fh_batch* b) {
  return guard([&] { B(b).b.reset(); });
}

int fh_batch_finalize(fh_batch* b) {
  return guard([&] { B(b).b.finalize(); });
}

int fh_batch_run(fh_batch* b, int use_graph) {
  return guard([&] {
    auto& h = B(b);
    h.b.run(C(h.ctx).stream, use_graph != 0);
  });
}

int fh_batch_info(const fh_batch* b, size_t* items, size_t* scratch_bytes) {
  return guard([&] {
    if (!b) throw fh::Error(FH_E_INVALID, "null batch");
    if (items) *items = b->b.items();
    if (scratch_bytes) *scratch_bytes = b->b.scratch_bytes();
  });
}

int fh_serve_run(fh_context* ctx, const fh_serve_config* config, const fh_key* const* rot_keys, const uint32_t* rot_galois,
                 const fh_key* relin_key, fh_serve_result* result) {
  return guard([&] {
    auto& c = C(ctx);
    need(config, "config"); need(rot_keys, "rot_keys"); need(rot_galois, "rot_galois"); need(result, "result");
    fh::ServeConfig cfg;
    cfg.rate_per_s = config->rate_per_s;
    cfg.requests = config->requests;
    cfg.start_level = config->start_level;
    cfg.hbm_budget_gb = config->hbm_budget_gb;
    cfg.mode = config->mode;
    cfg.inference_share = config->inference_share;
    cfg.seed = config->seed;
    cfg.checksum = config->checksum != 0;
    cfg.bootstrap_share = config->bootstrap_share;
    cfg.bootstrap = config->bootstrap;
    cfg.bootstrap_user = config->bootstrap_user;
    cfg.request_checksums = config->request_checksums;
    cfg.request_outputs = config->request_outputs;
    cfg.request_output_capacity = config->request_output_capacity;
    cfg.request_output_words = config->request_output_words;
    cfg.bootstrap_concurrent = config->bootstrap_concurrent;
    cfg.ticks_in_flight = config->ticks_in_flight;
    cfg.batch_min_ops = config->batch_min_ops;
    if (cfg.mode < 0 || cfg.mode > 2) throw fh::Error(FH_E_INVALID, "mode must be 0 (batched), 1 (per operation) or 2 (one request at a time)");
    const uint64_t* rk[8];
    for (int k = 0; k < 8; ++k) rk[k] = key_of(c, rot_keys[k]);
    fh::ServeResult r = fh::serve_run(c, cfg, rk, rot_galois, key_of(c, relin_key));
    result->p50_ms = r.p50_ms;
    result->p99_ms = r.p99_ms;
    result->mean_ms = r.mean_ms;
    result->max_ms = r.max_ms;
    result->requests_per_s = r.requests_per_s;
    result->wall_s = r.wall_s;
    result->mean_batch_ops = r.mean_batch_ops;
    result->ticks = r.ticks;
    result->completed = r.completed;
    result->inference = r.inference;
    result->polynomial = r.polynomial;
    result->max_active = r.max_active;
    result->checksum = r.checksum;
    result->bootstraps = r.bootstraps;
    result->p50_other_ms = r.p50_other_ms;
    result->p99_other_ms = r.p99_other_ms;
  });
}

// tunable slot and its smallest accepted value
static int* tune_slot(fh::Tune& t, const char* name, int* min_value = nullptr) {
  if (!name) throw fh::Error(FH_E_INVALID, "null tunable name");
  int lo = 1;
  int* slot = nullptr;
  static const char* const kThreadNames[4] = {"ntt_fwd_cols_threads", "ntt_fwd_rows_threads", "ntt_inv_rows_threads",
                                               "ntt_inv_cols_threads"};
  if (!std::strcmp(name, "ntt_chunk_mb")) slot = &t.ntt_chunk_mb;
  else if (!std::strcmp(name, "ntt_mix")) slot = &t.ntt_mix, lo = 0;
  else if (!std::strcmp(name, "ntt_split")) slot = &t.ntt_split, lo = 0;
  else if (!std::strcmp(name, "rot_streams")) slot = &t.rot_streams;
  else if (!std::strcmp(name, "rot_fused")) slot = &t.rot_fused, lo = 0;
  else if (!std::strcmp(name, "ksk_group_chunk")) slot = &t.ksk_group_chunk;
  else if (!std::strcmp(name, "ntt_streams")) slot = &t.ntt_streams;
  else if (!std::strcmp(name, "ntt_team")) slot = &t.ntt_team, lo = 0;
  else if (!std::strcmp(name, "ntt_team_per_xcd")) slot = &t.ntt_team_per_xcd;
  for (int i = 0; i < 4 && !slot; ++i)
    if (!std::strcmp(name, kThreadNames[i])) slot = &t.ntt_threads[i], lo = 256;
  if (!slot) throw fh::Error(FH_E_INVALID, std::string("unknown tunable: ") + name);
  if (min_value) *min_value = lo;
  return slot;
}

int fh_tune_set(fh_context* ctx, const char* name, int value) {
  return guard([&] {
    auto& c = C(ctx);
    int lo = 1;
    int* slot = tune_slot(c.tune, name, &lo);
    if (value < lo) throw fh::Error(FH_E_INVALID, std::string("tunable ") + name + " must be >= " + std::to_string(lo));
    if (!std::strcmp(name, "ntt_mix") && value > 1) throw fh::Error(FH_E_INVALID, "ntt_mix must be 0 or 1");
    if (!std::strcmp(name, "rot_streams") && value > 2) throw fh::Error(FH_E_INVALID, "rot_streams must be 1 or 2");
    if (!std::strcmp(name, "rot_fused") && value > 1) throw fh::Error(FH_E_INVALID, "rot_fused must be 0 or 1");
    if (!std::strcmp(name, "ksk_group_chunk") && value > 256) throw fh::Error(FH_E_INVALID, "ksk_group_chunk must be 1..256");
    if (!std::strcmp(name, "ntt_streams") && value > 2) throw fh::Error(FH_E_INVALID, "ntt_streams must be 1 or 2");
    if (!std::strcmp(name, "ntt_team") && value > 1) throw fh::Error(FH_E_INVALID, "ntt_team must be 0 or 1");
    if (!std::strcmp(name, "ntt_team_per_xcd") && value > 64) throw fh::Error(FH_E_INVALID, "ntt_team_per_xcd must be 1..64");
    if (!std::strcmp(name, "ntt_split") && value != 0 && !fh::ntt_split_ok(c.logN, value))
      throw fh::Error(FH_E_INVALID, "ntt_split " + std::to_string(value) + " is not built for logN " + std::to_string(c.logN));
    if (lo == 256 && value != 256 && value != 512) throw fh::Error(FH_E_INVALID, std::string(name) + " must be 256 or 512");
    if (!std::strcmp(name, "ntt_split")) c.ensure_twists(value ? value : fh::ntt_default_split(c.logN));
    *slot = value;
  });
}

int fh_autotune(fh_context* ctx, int polys, fh_autotune_report* report) {
  return guard([&] {
    auto& c = C(ctx);
    need(report, "report");
    const fh::AutotuneResult r = fh::autotune_ntt(c, polys);
    report->split = r.split;
    for (int i = 0; i < 4; ++i) report->threads[i] = r.threads[i];
    report->chunk_mb = r.chunk_mb;
    report->candidates = r.candidates;
    report->limbs = r.limbs;
    report->ms_default = r.ms_default;
    report->ms_best = r.ms_best;
  });
}

int fh_tune_get(const fh_context* ctx, const char* name, int* value) {
  return guard([&] {
    auto& c = const_cast<fh::Context&>(C(ctx));
    need(value, "value");
    if (name && !std::strcmp(name, "ntt_team_split_teams")) {
      // diagnostic (read only): team launches so far whose members spanned XCDs (synchronizes)
      uint32_t v = 0;
      FH_HIP(hipStreamSynchronize(c.stream));
      FH_HIP(hipMemcpy(&v, c.team_err() + 1, sizeof(v), hipMemcpyDeviceToHost));
      *value = (int)v;
      return;
    }
    *value = *tune_slot(c.tune, name);
  });
}

int fh_profile_enable(fh_context* ctx, int enable) {
  return guard([&] { C(ctx).prof.on = enable != 0; });
}

int fh_profile_reset(fh_context* ctx) {
  return guard([&] {
    auto& c = C(ctx);
    FH_HIP(hipStreamSynchronize(c.stream));
    std::lock_guard<std::mutex> lk(c.prof.mu);
    for (auto& r : c.prof.recs) { c.prof.give(r.a); c.prof.give(r.b); }
    c.prof.recs.clear();
  });
}

int fh_profile_read(fh_context* ctx, const char* name, int* launches, double* total_ms) {
  return guard([&] {
    auto& c = C(ctx);
    need(name, "name"); need(launches, "launches"); need(total_ms, "total_ms");
    FH_HIP(hipStreamSynchronize(c.stream));
    std::lock_guard<std::mutex> lk(c.prof.mu);
    int n = 0;
    double t = 0;
    for (auto& r : c.prof.recs) {
      if (std::strcmp(r.name, name) != 0) continue;
      float ms = 0;
      FH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      ++n;
      t += ms;
    }
    *launches = n;
    *total_ms = t;
  });
}

}  // extern "C"
