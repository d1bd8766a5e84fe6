// fideslib host facade: CKKS metadata, key generation, encoding and encryption on the host,
// every polynomial operation through the C ABI of the MI355X engine.
#include "fideslib.hpp"
#include "special_fft.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

namespace fideslib {

namespace {
typedef unsigned __int128 u128;
typedef __int128 i128;

uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  a %= q;
  for (; e; e >>= 1) {
    if (e & 1) r = (uint64_t)((u128)r * a % q);
    a = (uint64_t)((u128)a * a % q);
  }
  return r;
}

// extended ciphertexts (EvalFastRotationExt) only go to KeySwitchDown
static void require_standard(const CiphertextImpl& c, const char* op) {
  if (c.extended) throw std::runtime_error(std::string("fideslib: ") + op + ": extended ciphertext (call KeySwitchDown first)");
}

uint64_t signed_mod(int64_t v, uint64_t q) {
  if (v >= 0) return (uint64_t)v % q;
  uint64_t r = (uint64_t)(-(v + 1)) % q;
  r = (r + 1) % q;
  return r ? q - r : 0;
}

using detail::Cx;
using detail::kPi;
using detail::special_fft;
using detail::special_ifft;

int32_t normalize_index(int32_t index, uint32_t slots) {
  int64_t r = index % (int64_t)slots;
  return (int32_t)(r < 0 ? r + slots : r);
}

// OpenFHE CKKSPackedEncoding::CRTMult on word-sized moduli
std::vector<uint64_t> crt_mult(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b, const std::vector<uint64_t>& q) {
  std::vector<uint64_t> r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = (uint64_t)((u128)(a[i] % q[i]) * (b[i] % q[i]) % q[i]);
  return r;
}
}  // namespace

// deterministic sampler: mt19937_64 per call, seeded from the context seed and a call counter
struct CryptoContextImpl::Rng {
  uint64_t base;
  uint64_t counter = 0;
  std::mt19937_64 mt;
  uint64_t next_seed() { return base + counter++; }
  void reseed(uint64_t s) { mt.seed(s); }
  uint64_t uniform_mod(uint64_t q) {
    const int bits = 64 - __builtin_clzll(q);
    for (;;) {
      uint64_t r = mt() >> (64 - bits);
      if (r < q) return r;
    }
  }
  double unit() { return (double)(mt() >> 11) * (1.0 / 9007199254740992.0); }
  int64_t gauss() {
    double u1 = unit(), u2 = unit();
    if (u1 < 1e-300) u1 = 1e-300;
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    return (int64_t)std::llround(z * 3.19);
  }
};

NativeContext::~NativeContext() {
  if (ctx) fh_context_destroy(ctx);
}

DeviceBuffer::DeviceBuffer(const std::shared_ptr<NativeContext>& c, size_t w) : owner(c), words(w) {
  void* p = nullptr;
  if (fh_device_alloc(c->ctx, std::max<size_t>(w, 1) * 8, &p) != FH_OK) throw std::runtime_error(fh_last_error());
  ptr = (uint64_t*)p;
}
DeviceBuffer::~DeviceBuffer() {
  if (ptr) fh_device_free(owner->ctx, ptr);
}

CryptoContext GenCryptoContext(const CCParams& params) { return std::make_shared<CryptoContextImpl>(params); }

CryptoContextImpl::CryptoContextImpl(const CCParams& p) : params_(p) {
  int logN = 0;
  while ((1u << logN) < p.ring_dim_) ++logN;
  if ((1u << logN) != p.ring_dim_) throw std::runtime_error("ring dimension must be a power of two");
  L1_ = p.depth_ + 1;
  check(fh_context_create_generated(logN, (int)L1_, (int)p.num_special_, (int)p.dnum_, (int)p.first_bits_, (int)p.scale_bits_, 60, &ctx_));
  native_ = std::make_shared<NativeContext>(ctx_);
  int shape[5];
  check(fh_context_shape(ctx_, shape));
  N_ = 1u << shape[0];
  K_ = shape[2];
  dnum_ = shape[3];
  alpha_ = shape[4];
  moduli_.resize(L1_ + K_);
  check(fh_context_moduli(ctx_, moduli_.data(), nullptr));
  scale_ = std::ldexp(1.0, (int)p.scale_bits_);
  pmodq_.resize(L1_);
  for (uint32_t i = 0; i < L1_; ++i) {
    uint64_t pr = 1;
    for (uint32_t k = 0; k < K_; ++k) pr = (uint64_t)((u128)pr * (moduli_[L1_ + k] % moduli_[i]) % moduli_[i]);
    pmodq_[i] = pr;
  }
  rng_ = std::make_unique<Rng>();
  rng_->base = p.seed_;
  // per-level scale tables (by towers - 1). FLEXIBLEAUTO: the top level's factor is the top Q prime,
  // level k's the previous one squared over the prime the k-th rescale drops, in OpenFHE's order of
  // operations (prevSF * prevSF / q); the technique is refused when a factor leaves (0.5, 2) times
  // the top one, as OpenFHE does.
  sfReal_.assign(L1_, scale_);
  sfBig_.assign(L1_, scale_ * scale_);
  mrf_.assign(L1_, scale_);
  if (p.tech_ == ScalingTechnique::FLEXIBLEAUTO) {
    sfReal_[L1_ - 1] = (double)moduli_[L1_ - 1];
    for (int l = (int)L1_ - 2; l >= 0; --l) {
      const double prev = sfReal_[l + 1];
      sfReal_[l] = prev * prev / (double)moduli_[l + 1];
      const double ratio = sfReal_[l] / sfReal_[L1_ - 1];
      if (ratio <= 0.5 || ratio >= 2.0)
        throw std::runtime_error("fideslib: FLEXIBLEAUTO cannot support this number of levels in this parameter setting (use FIXEDMANUAL)");
    }
    for (uint32_t l = 0; l < L1_; ++l) {
      sfBig_[l] = sfReal_[l] * sfReal_[l];
      mrf_[l] = (double)moduli_[l];
    }
  }
}

double CryptoContextImpl::GetScalingFactorReal(uint32_t level) const {
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  return sfReal_[L1_ - 1 - level];
}
double CryptoContextImpl::GetScalingFactorRealBig(uint32_t level) const {
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  return sfBig_[L1_ - 1 - level];
}
double CryptoContextImpl::GetModReduceFactor(uint32_t level) const {
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  return mrf_[L1_ - 1 - level];
}

CryptoContextImpl::~CryptoContextImpl() {
  if (!ctx_) return;
  fh_synchronize(ctx_);
  if (batch_) fh_batch_destroy(batch_);
  if (relin_) fh_key_destroy(ctx_, relin_);
  for (auto& kv : rot_) fh_key_destroy(ctx_, kv.second);
  rot_.clear();
  // the native context itself goes with the last buffer that references it
}

void CryptoContextImpl::check(int rc) const {
  if (rc != FH_OK) throw std::runtime_error(std::string("fideslib: ") + fh_last_error());
}

void CryptoContextImpl::Synchronize() const { check(fh_synchronize(ctx_)); }

Ciphertext CryptoContextImpl::newCt(uint32_t towers, uint32_t ncomp) const {
  auto c = std::make_shared<CiphertextImpl>();
  c->data = std::make_shared<DeviceBuffer>(native_, (size_t)ncomp * towers * N_);
  c->towers = towers;
  c->ncomp = ncomp;
  return c;
}

// signed small coefficients -> EVALUATION residues over pids (host reduction, device NTT)
std::shared_ptr<DeviceBuffer> CryptoContextImpl::smallToEval(const std::vector<int64_t>& v, const std::vector<int32_t>& pids) const {
  std::vector<uint64_t> h((size_t)pids.size() * N_);
  for (size_t l = 0; l < pids.size(); ++l) {
    const uint64_t q = moduli_[pids[l]];
    for (uint32_t j = 0; j < N_; ++j) h[l * N_ + j] = signed_mod(v[j], q);
  }
  auto d = std::make_shared<DeviceBuffer>(native_, h.size());
  check(fh_copy_h2d(ctx_, d->ptr, h.data(), h.size() * 8));
  check(fh_ntt_forward(ctx_, d->ptr, pids.data(), (int)pids.size(), 1, 0));
  return d;
}

std::shared_ptr<DeviceBuffer> CryptoContextImpl::uniform(const std::vector<int32_t>& pids) {
  std::vector<uint64_t> h((size_t)pids.size() * N_);
  for (size_t l = 0; l < pids.size(); ++l)
    for (uint32_t j = 0; j < N_; ++j) h[l * N_ + j] = rng_->uniform_mod(moduli_[pids[l]]);
  auto d = std::make_shared<DeviceBuffer>(native_, h.size());
  check(fh_copy_h2d(ctx_, d->ptr, h.data(), h.size() * 8));
  return d;
}

std::vector<int64_t> CryptoContextImpl::gaussian() {
  std::vector<int64_t> e(N_);
  for (auto& x : e) x = rng_->gauss();
  return e;
}

KeyPair CryptoContextImpl::KeyGen() {
  rng_->reseed(rng_->next_seed());
  std::vector<int64_t> v(N_, 0);
  if (params_.dist_ == SecretKeyDist::UNIFORM_TERNARY) {
    for (auto& x : v) x = (int64_t)(rng_->mt() % 3) - 1;
  } else {
    uint32_t placed = 0;
    while (placed < params_.hw_) {
      const uint64_t pos = rng_->mt() % N_;
      if (v[pos]) continue;
      v[pos] = (rng_->mt() & 1) ? 1 : -1;
      ++placed;
    }
  }
  std::vector<int32_t> all(L1_ + K_);
  for (uint32_t i = 0; i < L1_ + K_; ++i) all[i] = (int32_t)i;
  KeyPair kp;
  kp.secretKey = std::make_shared<PrivateKeyImpl>();
  kp.secretKey->s = smallToEval(v, all);
  // public key over Q: b = -a s + e, a uniform
  rng_->reseed(rng_->next_seed());
  std::vector<int32_t> q(all.begin(), all.begin() + L1_);
  kp.publicKey = std::make_shared<PublicKeyImpl>();
  kp.publicKey->a = uniform(q);
  auto e = smallToEval(gaussian(), q);
  kp.publicKey->b = std::make_shared<DeviceBuffer>(native_, (size_t)L1_ * N_);
  check(fh_eltwise_binary(ctx_, 2, kp.publicKey->a->ptr, kp.secretKey->s->ptr, kp.publicKey->b->ptr, nullptr, (int)L1_));
  check(fh_eltwise_binary(ctx_, 1, e->ptr, kp.publicKey->b->ptr, kp.publicKey->b->ptr, nullptr, (int)L1_));
  return kp;
}

// OpenFHE KeySwitchHYBRID::KeySwitchGenInternal: digit d: b_d = -a_d s_new + e_d (+ [P]_{q_i} s_old on
// the digit's Q primes), layout [dnum][b, a][L1+K][N].
fh_key* CryptoContextImpl::switchKey(const DeviceBuffer& s_old, const DeviceBuffer& s_new) {
  rng_->reseed(rng_->next_seed());
  const uint32_t np = L1_ + K_;
  std::vector<int32_t> all(np);
  for (uint32_t i = 0; i < np; ++i) all[i] = (int32_t)i;
  DeviceBuffer key(native_, (size_t)dnum_ * 2 * np * N_);
  for (uint32_t d = 0; d < dnum_; ++d) {
    uint64_t* b = key.ptr + (size_t)(2 * d) * np * N_;
    uint64_t* a = key.ptr + (size_t)(2 * d + 1) * np * N_;
    auto ua = uniform(all);
    check(fh_copy_d2d(ctx_, a, ua->ptr, (size_t)np * N_ * 8));
    auto e = smallToEval(gaussian(), all);
    check(fh_eltwise_binary(ctx_, 2, a, s_new.ptr, b, all.data(), (int)np));
    check(fh_eltwise_binary(ctx_, 1, e->ptr, b, b, all.data(), (int)np));
    const uint32_t start = d * alpha_, end = std::min(start + alpha_, L1_);
    std::vector<int32_t> dig;
    std::vector<uint64_t> pm;
    for (uint32_t i = start; i < end; ++i) {
      dig.push_back((int32_t)i);
      pm.push_back(pmodq_[i]);
    }
    DeviceBuffer t(native_, (size_t)dig.size() * N_);
    check(fh_eltwise_scalar(ctx_, 1, s_old.ptr + (size_t)start * N_, pm.data(), t.ptr, dig.data(), (int)dig.size()));
    check(fh_eltwise_binary(ctx_, 0, b + (size_t)start * N_, t.ptr, b + (size_t)start * N_, dig.data(), (int)dig.size()));
  }
  fh_key* k = nullptr;
  check(fh_key_create(ctx_, key.ptr, &k));
  return k;
}

void CryptoContextImpl::EvalMultKeyGen(const PrivateKey& sk) {
  DeviceBuffer s2(native_, (size_t)L1_ * N_);
  check(fh_eltwise_binary(ctx_, 2, sk->s->ptr, sk->s->ptr, s2.ptr, nullptr, (int)L1_));
  if (relin_) fh_key_destroy(ctx_, relin_);
  relin_ = switchKey(s2, *sk->s);
}

uint32_t CryptoContextImpl::FindAutomorphismIndex(int32_t index) const {
  uint32_t g = 1;
  check(fh_galois_index(ctx_, index, &g));
  return g;
}

// OpenFHE EvalAutomorphismKeyGen: switch from s to sigma_{k^-1}(s)
void CryptoContextImpl::EvalRotateKeyGen(const PrivateKey& sk, const std::vector<int32_t>& steps) {
  for (int32_t r : steps) {
    if (normalize_index(r, N_ / 2) == 0) continue;
    automorphismKey(sk, galoisFor(r));
  }
}

// OpenFHE EvalConjugateKeyGen: the automorphism X -> X^(2N-1)
void CryptoContextImpl::EvalConjugateKeyGen(const PrivateKey& sk) { automorphismKey(sk, 2 * N_ - 1); }

void CryptoContextImpl::automorphismKey(const PrivateKey& sk, uint32_t k) {
  if (rot_.count(k)) return;
  const uint32_t np = L1_ + K_, M = 2 * N_;
  // k^-1 mod 2N (odd k, power-of-two modulus: Newton's iteration doubles the correct low bits)
  uint64_t inv = k;
  for (int i = 0; i < 6; ++i) inv *= 2 - (uint64_t)k * inv;
  inv &= M - 1;
  DeviceBuffer sp(native_, (size_t)np * N_);
  check(fh_automorph(ctx_, sk->s->ptr, (int)np, (uint32_t)inv, sp.ptr));
  rot_[k] = switchKey(*sk->s, sp);
}

const fh_key* CryptoContextImpl::rotKey(uint32_t galois) const {
  auto it = rot_.find(galois);
  if (it == rot_.end()) throw std::runtime_error("fideslib: no rotation key for automorphism index " + std::to_string(galois));
  return it->second;
}

// ---- encoding ----
void CryptoContextImpl::encodeInto(const std::vector<std::complex<double>>& v, uint32_t slots, double scale,
                                   const std::vector<int32_t>& pids, uint64_t* dev) const {
  // on the device (fh_encode_slots): the special inverse FFT, rounding, residues and NTT; only
  // the slot values are uploaded
  if (slots == 0 || (slots & (slots - 1)) || slots > N_ / 2) throw std::runtime_error("fideslib: slots must be a power of two <= N/2");
  std::vector<double> z(2 * (size_t)slots, 0.0);
  for (size_t i = 0; i < std::min<size_t>(v.size(), slots); ++i) {
    z[2 * i] = v[i].real();
    z[2 * i + 1] = v[i].imag();
  }
  check(fh_encode_slots(ctx_, z.data(), 1, (int)slots, scale, pids.data(), (int)pids.size(), dev, 0));
}

// FIXEDMANUAL: the nominal scale to the degree; FLEXIBLEAUTO: the level's factor (OpenFHE's
// MakeCKKSPackedPlaintext takes GetScalingFactorReal(level)); a degree-2 plaintext would be encoded at
// the level's factor and multiplied by it again as an integer, which this encoder does not do.
double CryptoContextImpl::ptScale(uint32_t towers, size_t noiseScaleDeg) const {
  if (noiseScaleDeg < 1) throw std::runtime_error("fideslib: noiseScaleDeg must be >= 1");
  if (!flex()) return std::pow(scale_, (double)noiseScaleDeg);
  if (noiseScaleDeg != 1) throw std::runtime_error("fideslib: FLEXIBLEAUTO plaintexts are encoded at noise scale degree 1");
  return sfReal_[towers - 1];
}

Plaintext CryptoContextImpl::MakeCKKSPackedPlaintext(const std::vector<std::complex<double>>& v, size_t noiseScaleDeg, uint32_t level,
                                                     uint32_t slots) {
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  auto pt = std::make_shared<PlaintextImpl>();
  pt->towers = L1_ - level;
  pt->slots = slots ? slots : (params_.batch_ ? params_.batch_ : N_ / 2);
  pt->noiseScaleDeg = (uint32_t)noiseScaleDeg;
  pt->scalingFactor = ptScale(pt->towers, noiseScaleDeg);
  pt->values = v;
  pt->values.resize(pt->slots);
  std::vector<int32_t> pids(pt->towers);
  for (uint32_t i = 0; i < pt->towers; ++i) pids[i] = (int32_t)i;
  pt->data = std::make_shared<DeviceBuffer>(native_, (size_t)pt->towers * N_);
  encodeInto(pt->values, pt->slots, pt->scalingFactor, pids, pt->data->ptr);
  return pt;
}

Plaintext CryptoContextImpl::MakeCKKSPackedPlaintext(const std::vector<double>& v, size_t noiseScaleDeg, uint32_t level, uint32_t slots) {
  std::vector<std::complex<double>> c(v.begin(), v.end());
  return MakeCKKSPackedPlaintext(c, noiseScaleDeg, level, slots);
}

Plaintext CryptoContextImpl::MakeCKKSPackedPlaintextExtended(const std::vector<std::complex<double>>& v, size_t noiseScaleDeg,
                                                             uint32_t level, uint32_t slots) {
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  auto pt = std::make_shared<PlaintextImpl>();
  pt->towers = L1_ - level;
  pt->extended = true;
  pt->slots = slots ? slots : (params_.batch_ ? params_.batch_ : N_ / 2);
  pt->noiseScaleDeg = (uint32_t)noiseScaleDeg;
  pt->scalingFactor = ptScale(pt->towers, noiseScaleDeg);
  pt->values = v;
  pt->values.resize(pt->slots);
  std::vector<int32_t> pids;
  for (uint32_t i = 0; i < pt->towers; ++i) pids.push_back((int32_t)i);
  for (uint32_t k = 0; k < K_; ++k) pids.push_back((int32_t)(L1_ + k));
  pt->data = std::make_shared<DeviceBuffer>(native_, pids.size() * N_);
  encodeInto(pt->values, pt->slots, pt->scalingFactor, pids, pt->data->ptr);
  return pt;
}

// ---- encryption ----
Ciphertext CryptoContextImpl::Encrypt(const PrivateKey& sk, const Plaintext& pt) {
  if (pt->extended) throw std::runtime_error("fideslib: cannot encrypt an extended plaintext");
  rng_->reseed(rng_->next_seed());
  const uint32_t t = pt->towers;
  std::vector<int32_t> pids(t);
  for (uint32_t i = 0; i < t; ++i) pids[i] = (int32_t)i;
  auto ct = newCt(t, 2);
  auto a = uniform(pids);
  auto e = smallToEval(gaussian(), pids);
  uint64_t* c0 = ct->data->ptr;
  uint64_t* c1 = ct->data->ptr + (size_t)t * N_;
  check(fh_copy_d2d(ctx_, c1, a->ptr, (size_t)t * N_ * 8));
  check(fh_eltwise_binary(ctx_, 2, a->ptr, sk->s->ptr, c0, nullptr, (int)t));      // a s
  check(fh_eltwise_binary(ctx_, 1, e->ptr, c0, c0, nullptr, (int)t));              // e - a s
  check(fh_eltwise_binary(ctx_, 0, c0, pt->data->ptr, c0, nullptr, (int)t));       // + m
  ct->noiseScaleDeg = pt->noiseScaleDeg;
  ct->scalingFactor = pt->scalingFactor;
  ct->slots = pt->slots;
  return ct;
}

// public-key encryption: v ternary, c0 = b v + e0 + m, c1 = a v + e1
Ciphertext CryptoContextImpl::Encrypt(const PublicKey& pk, const Plaintext& pt) {
  if (pt->extended) throw std::runtime_error("fideslib: cannot encrypt an extended plaintext");
  rng_->reseed(rng_->next_seed());
  const uint32_t t = pt->towers;
  std::vector<int32_t> pids(t);
  for (uint32_t i = 0; i < t; ++i) pids[i] = (int32_t)i;
  std::vector<int64_t> vt(N_);
  for (auto& x : vt) x = (int64_t)(rng_->mt() % 3) - 1;
  auto v = smallToEval(vt, pids);
  auto e0 = smallToEval(gaussian(), pids);
  auto e1 = smallToEval(gaussian(), pids);
  auto ct = newCt(t, 2);
  uint64_t* c0 = ct->data->ptr;
  uint64_t* c1 = ct->data->ptr + (size_t)t * N_;
  check(fh_eltwise_binary(ctx_, 2, pk->b->ptr, v->ptr, c0, nullptr, (int)t));
  check(fh_eltwise_binary(ctx_, 0, c0, e0->ptr, c0, nullptr, (int)t));
  check(fh_eltwise_binary(ctx_, 0, c0, pt->data->ptr, c0, nullptr, (int)t));
  check(fh_eltwise_binary(ctx_, 2, pk->a->ptr, v->ptr, c1, nullptr, (int)t));
  check(fh_eltwise_binary(ctx_, 0, c1, e1->ptr, c1, nullptr, (int)t));
  ct->noiseScaleDeg = pt->noiseScaleDeg;
  ct->scalingFactor = pt->scalingFactor;
  ct->slots = pt->slots;
  return ct;
}

Plaintext CryptoContextImpl::Decrypt(const PrivateKey& sk, const Ciphertext& ct0) {
  require_standard(*ct0, "Decrypt");
  // FLEXIBLEAUTO: a depth-2 result is rescaled before decoding
  const Ciphertext ct = flex() && ct0->noiseScaleDeg == 2 && ct0->towers > 1 ? modReduce(ct0) : ct0;
  const uint32_t t = ct->towers;
  DeviceBuffer m(native_, (size_t)t * N_), tmp(native_, (size_t)t * N_);
  const uint64_t* c0 = ct->data->ptr;
  const uint64_t* c1 = ct->data->ptr + (size_t)t * N_;
  check(fh_eltwise_binary(ctx_, 2, c1, sk->s->ptr, m.ptr, nullptr, (int)t));
  check(fh_eltwise_binary(ctx_, 0, c0, m.ptr, m.ptr, nullptr, (int)t));
  if (ct->ncomp == 3) {
    const uint64_t* c2 = ct->data->ptr + (size_t)2 * t * N_;
    check(fh_eltwise_binary(ctx_, 2, sk->s->ptr, sk->s->ptr, tmp.ptr, nullptr, (int)t));
    check(fh_eltwise_binary(ctx_, 2, c2, tmp.ptr, tmp.ptr, nullptr, (int)t));
    check(fh_eltwise_binary(ctx_, 0, m.ptr, tmp.ptr, m.ptr, nullptr, (int)t));
  }
  const int nl = t >= 2 ? 2 : 1;
  std::vector<int32_t> pids = {0, 1};
  check(fh_ntt_inverse(ctx_, m.ptr, pids.data(), nl, 1, 0));
  std::vector<uint64_t> h((size_t)nl * N_);
  check(fh_copy_d2h(ctx_, h.data(), m.ptr, h.size() * 8));
  std::vector<double> coef(N_);
  const uint64_t q0 = moduli_[0];
  if (nl == 1) {
    for (uint32_t j = 0; j < N_; ++j) coef[j] = h[j] > q0 / 2 ? -(double)(q0 - h[j]) : (double)h[j];
  } else {
    const uint64_t q1 = moduli_[1];
    const uint64_t q0inv = powmod(q0 % q1, q1 - 2, q1);
    const u128 Q = (u128)q0 * q1;
    for (uint32_t j = 0; j < N_; ++j) {
      const uint64_t a = h[j], b = h[N_ + j];
      const uint64_t d = (b + q1 - a % q1) % q1;
      const uint64_t k = (uint64_t)((u128)d * q0inv % q1);
      const u128 x = (u128)a + (u128)q0 * k;
      coef[j] = x > Q / 2 ? -(double)(Q - x) : (double)x;
    }
  }
  auto pt = std::make_shared<PlaintextImpl>();
  pt->towers = t;
  pt->slots = ct->slots;
  pt->noiseScaleDeg = ct->noiseScaleDeg;
  pt->scalingFactor = ct->scalingFactor;
  std::vector<Cx> z(ct->slots);
  const uint32_t gap = (N_ / 2) / ct->slots;
  for (uint32_t i = 0; i < ct->slots; ++i) z[i] = Cx{coef[i * gap] / ct->scalingFactor, coef[N_ / 2 + i * gap] / ct->scalingFactor};
  special_fft(z, 2ull * N_);
  pt->values.resize(ct->slots);
  for (uint32_t i = 0; i < ct->slots; ++i) pt->values[i] = std::complex<double>(z[i].re, z[i].im);
  return pt;
}

// ---- raw device residues ----
Ciphertext CryptoContextImpl::LoadCiphertext(const uint64_t* dev, uint32_t towers, uint32_t ncomp, uint32_t noiseScaleDeg, double scale,
                                             uint32_t slots) {
  if (towers < 1 || towers > L1_ || ncomp < 2 || ncomp > 3) throw std::runtime_error("fideslib: LoadCiphertext: shape out of range");
  auto ct = newCt(towers, ncomp);
  check(fh_copy_d2d(ctx_, ct->data->ptr, dev, (size_t)ncomp * towers * N_ * 8));
  ct->noiseScaleDeg = noiseScaleDeg;
  ct->scalingFactor = scale;
  ct->slots = slots ? slots : N_ / 2;
  return ct;
}

void CryptoContextImpl::StoreCiphertext(const Ciphertext& ct, uint64_t* dev) const {
  check(fh_copy_d2d(ctx_, dev, ct->data->ptr, (size_t)ct->ncomp * ct->towers * N_ * 8));
}

// ---- level alignment (FIXEDMANUAL: drop towers, no rescale) ----
Ciphertext CryptoContextImpl::adjusted(const Ciphertext& ct, uint32_t towers) const {
  require_standard(*ct, "level alignment");
  if (ct->towers == towers) return ct;
  if (ct->towers < towers) throw std::runtime_error("fideslib: cannot raise a ciphertext's level");
  auto out = newCt(towers, ct->ncomp);
  for (uint32_t c = 0; c < ct->ncomp; ++c)
    check(fh_copy_d2d(ctx_, out->data->ptr + (size_t)c * towers * N_, ct->data->ptr + (size_t)c * ct->towers * N_,
                      (size_t)towers * N_ * 8));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

Ciphertext CryptoContextImpl::LevelReduce(const Ciphertext& ct, uint32_t levels) {
  if (levels >= ct->towers) throw std::runtime_error("fideslib: LevelReduce below level 0");
  return adjusted(ct, ct->towers - levels);
}

Ciphertext CryptoContextImpl::EvalNegate(const Ciphertext& ct) {
  require_standard(*ct, "EvalNegate");
  auto out = newCt(ct->towers, ct->ncomp);
  for (uint32_t c = 0; c < ct->ncomp; ++c)
    check(fh_eltwise_scalar(ctx_, 2, ct->data->ptr + (size_t)c * ct->towers * N_, nullptr, out->data->ptr + (size_t)c * ct->towers * N_,
                            nullptr, (int)ct->towers));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

static void require_same_deg(const CiphertextImpl& a, const CiphertextImpl& b, const char* op) {
  require_standard(a, op);
  require_standard(b, op);
  if (a.noiseScaleDeg != b.noiseScaleDeg)
    throw std::runtime_error(std::string("fideslib: ") + op + ": operands have different noise scale degrees (rescale first)");
}

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a1, const Ciphertext& b1) {
  Ciphertext a0 = a1, b0 = b1;
  if (flex()) {
    require_standard(*a0, "EvalAdd");
    require_standard(*b0, "EvalAdd");
    adjustLevelsAndDepth(a0, b0);   // OpenFHE AdjustForAddOrSub: one level and one depth
  }
  require_same_deg(*a0, *b0, "EvalAdd");
  // over the first t towers of each component, read in place (no level-reduced copies)
  const uint32_t t = std::min(a0->towers, b0->towers);
  const uint32_t nc = std::max(a0->ncomp, b0->ncomp);
  auto out = newCt(t, nc);
  for (uint32_t c = 0; c < nc; ++c) {
    uint64_t* o = out->data->ptr + (size_t)c * t * N_;
    if (c < a0->ncomp && c < b0->ncomp)
      check(fh_eltwise_binary(ctx_, 0, a0->data->ptr + (size_t)c * a0->towers * N_, b0->data->ptr + (size_t)c * b0->towers * N_, o, nullptr,
                              (int)t));
    else {
      const Ciphertext& x = c < a0->ncomp ? a0 : b0;
      check(fh_copy_d2d(ctx_, o, x->data->ptr + (size_t)c * x->towers * N_, (size_t)t * N_ * 8));
    }
  }
  out->noiseScaleDeg = a0->noiseScaleDeg;
  out->scalingFactor = a0->scalingFactor;
  out->slots = a0->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a0, const Ciphertext& b0) { return EvalAdd(a0, EvalNegate(b0)); }

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a0, const Plaintext& pt) {
  if (pt->extended) throw std::runtime_error("fideslib: extended plaintexts are for LinearTransform");
  if (flex()) {
    Ciphertext a = a0, p = morph(pt);
    adjustLevelsAndDepth(a, p);
    return addPlain(a, p, 0);
  }
  if (pt->noiseScaleDeg != a0->noiseScaleDeg) throw std::runtime_error("fideslib: EvalAdd: plaintext scale degree differs");
  const uint32_t t = std::min(a0->towers, pt->towers);
  auto a = adjusted(a0, t);
  auto out = newCt(t, a->ncomp);
  check(fh_copy_d2d(ctx_, out->data->ptr, a->data->ptr, (size_t)a->ncomp * t * N_ * 8));
  check(fh_eltwise_binary(ctx_, 0, a->data->ptr, pt->data->ptr, out->data->ptr, nullptr, (int)t));
  out->noiseScaleDeg = a->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor;
  out->slots = a->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a0, const Plaintext& pt) {
  if (pt->extended) throw std::runtime_error("fideslib: extended plaintexts are for LinearTransform");
  if (flex()) {
    Ciphertext a = a0, p = morph(pt);
    adjustLevelsAndDepth(a, p);
    return addPlain(a, p, 1);
  }
  if (pt->noiseScaleDeg != a0->noiseScaleDeg) throw std::runtime_error("fideslib: EvalSub: plaintext scale degree differs");
  const uint32_t t = std::min(a0->towers, pt->towers);
  auto a = adjusted(a0, t);
  auto out = newCt(t, a->ncomp);
  check(fh_copy_d2d(ctx_, out->data->ptr, a->data->ptr, (size_t)a->ncomp * t * N_ * 8));
  check(fh_eltwise_binary(ctx_, 1, a->data->ptr, pt->data->ptr, out->data->ptr, nullptr, (int)t));
  out->noiseScaleDeg = a->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor;
  out->slots = a->slots;
  return out;
}

// ElemForEvalMult (src/CKKS/Context.cu:669-767): round(operand * scFactor) as a 128-bit integer
// (OpenFHE's non-FMA form operand / approxFactor * scFactor + 0.5, truncated), reduced per tower,
// times 2^logApprox by CRT multiplication.
std::vector<uint64_t> CryptoContextImpl::ElemForEvalMult(uint32_t towers, double operand) const {
  // the level's scaling factor: FIXEDMANUAL the same at every level, FLEXIBLEAUTO the level's own
  return ElemForEvalMultAt(towers, operand, sfReal(towers));
}

std::vector<uint64_t> CryptoContextImpl::ElemForEvalMultAt(uint32_t towers, double operand, double scFactor) const {
  const int32_t MAX_BITS = 125, MAX_LOG_STEP = 60;
  int32_t logApprox = 0;
  const double res = std::fabs(operand * scFactor);
  if (res > 0) {
    const int32_t logSF = (int32_t)std::ceil(std::log2(res));
    const int32_t logValid = logSF <= MAX_BITS ? logSF : MAX_BITS;
    logApprox = logSF - logValid;
  }
  const double approxFactor = std::pow(2.0, logApprox);
  volatile double prod = operand / approxFactor * scFactor;  // rounded product, then + 0.5 (no FMA)
  const i128 large = (i128)(prod + 0.5);
  const i128 bound = (i128)1 << 63;
  const i128 large_abs = large < 0 ? -large : large;
  std::vector<uint64_t> q(moduli_.begin(), moduli_.begin() + towers);
  std::vector<uint64_t> f(towers);
  for (uint32_t i = 0; i < towers; ++i) {
    if (large_abs > bound) {
      const i128 r = large % (i128)q[i];
      f[i] = (uint64_t)(r < 0 ? r + (i128)q[i] : r);
    } else {
      const int64_t sc = (int64_t)large;
      const int64_t r = sc % (int64_t)q[i];
      f[i] = (uint64_t)(r < 0 ? r + (int64_t)q[i] : r);
    }
  }
  if (logApprox > 0) {
    int32_t step = logApprox <= MAX_LOG_STEP ? logApprox : MAX_LOG_STEP;
    std::vector<uint64_t> approx(towers, (uint64_t)1 << step);
    logApprox -= step;
    while (logApprox > 0) {
      step = logApprox <= MAX_LOG_STEP ? logApprox : MAX_LOG_STEP;
      approx = crt_mult(approx, std::vector<uint64_t>(towers, (uint64_t)1 << step), q);
      logApprox -= step;
    }
    f = crt_mult(f, approx, q);
  }
  for (uint32_t i = 0; i < towers; ++i) f[i] %= q[i];
  return f;
}

// ElemForEvalAddOrSub (src/CKKS/Context.cu:777-863) for a non-negative operand.
std::vector<uint64_t> CryptoContextImpl::ElemForEvalAddOrSub(uint32_t towers, double operand, uint32_t noiseDeg) const {
  return ElemForEvalAddOrSubAt(towers, operand, noiseDeg, sfReal(towers));
}

std::vector<uint64_t> CryptoContextImpl::ElemForEvalAddOrSubAt(uint32_t towers, double operand, uint32_t noiseDeg, double scFactor) const {
  const int32_t MAX_BITS = 61, MAX_LOG_STEP = 60;
  int32_t logApprox = 0;
  const double res = std::fabs(operand * scFactor);
  if (res > 0) {
    const int32_t logSF = (int32_t)std::ceil(std::log2(res));
    const int32_t logValid = logSF <= MAX_BITS ? logSF : MAX_BITS;
    logApprox = logSF - logValid;
  }
  const double approxFactor = std::pow(2.0, logApprox);
  volatile double prod = operand * scFactor / approxFactor;
  const uint64_t sc = (uint64_t)(prod + 0.5);
  std::vector<uint64_t> q(moduli_.begin(), moduli_.begin() + towers);
  std::vector<uint64_t> c(towers, sc);
  if (logApprox > 0) {
    int32_t step = logApprox <= MAX_LOG_STEP ? logApprox : MAX_LOG_STEP;
    std::vector<uint64_t> approx(towers, (uint64_t)1 << step);
    logApprox -= step;
    while (logApprox > 0) {
      step = logApprox <= MAX_LOG_STEP ? logApprox : MAX_LOG_STEP;
      approx = crt_mult(approx, std::vector<uint64_t>(towers, (uint64_t)1 << step), q);
      logApprox -= step;
    }
    c = crt_mult(c, approx, q);
  }
  const std::vector<uint64_t> sf(towers, (uint64_t)(scFactor + 0.5));
  for (uint32_t i = 1; i < noiseDeg; ++i) c = crt_mult(c, sf, q);
  for (uint32_t i = 0; i < towers; ++i) c[i] %= q[i];
  return c;
}

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a, double scalar) {
  require_standard(*a, "EvalAdd");
  auto out = newCt(a->towers, a->ncomp);
  check(fh_copy_d2d(ctx_, out->data->ptr, a->data->ptr, (size_t)a->ncomp * a->towers * N_ * 8));
  const auto s = ElemForEvalAddOrSub(a->towers, std::fabs(scalar), a->noiseScaleDeg);
  if (scalar >= 0) {
    check(fh_eltwise_scalar(ctx_, 0, a->data->ptr, s.data(), out->data->ptr, nullptr, (int)a->towers));
  } else {
    std::vector<uint64_t> neg(a->towers);
    for (uint32_t i = 0; i < a->towers; ++i) neg[i] = s[i] ? moduli_[i] - s[i] : 0;
    check(fh_eltwise_scalar(ctx_, 0, a->data->ptr, neg.data(), out->data->ptr, nullptr, (int)a->towers));
  }
  out->noiseScaleDeg = a->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor;
  out->slots = a->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a, double scalar) { return EvalAdd(a, -scalar); }

// FLEXIBLEAUTO rescales a degree-2 operand first (OpenFHE LeveledSHECKKSRNS::EvalMult(ct, double))
Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a0, double scalar) {
  require_standard(*a0, "EvalMult");
  const Ciphertext a = flex() && a0->noiseScaleDeg == 2 ? modReduce(a0) : a0;
  return multCore(a, scalar);
}

// EvalMultCore(ct, double): every component times the CRT encoding of operand at the level's
// scaling factor; noise degree + 1, scale times that factor
Ciphertext CryptoContextImpl::multCore(const Ciphertext& a, double operand) const {
  const auto f = ElemForEvalMult(a->towers, operand);
  auto out = newCt(a->towers, a->ncomp);
  for (uint32_t c = 0; c < a->ncomp; ++c)
    check(fh_eltwise_scalar(ctx_, 1, a->data->ptr + (size_t)c * a->towers * N_, f.data(), out->data->ptr + (size_t)c * a->towers * N_,
                            nullptr, (int)a->towers));
  out->noiseScaleDeg = a->noiseScaleDeg + 1;
  out->scalingFactor = a->scalingFactor * sfReal(a->towers);
  out->slots = a->slots;
  return out;
}

// ModReduceInternal: the top tower dropped, noise degree - 1, the scale divided by the technique's
// ModReduceFactor (FLEXIBLEAUTO: the dropped prime; FIXEDMANUAL: the nominal 2^scaleModSize, not q_l)
Ciphertext CryptoContextImpl::modReduce(const Ciphertext& ct) const {
  require_standard(*ct, "Rescale");
  if (ct->towers < 2) throw std::runtime_error("fideslib: Rescale at the last tower");
  auto out = newCt(ct->towers - 1, ct->ncomp);
  check(fh_rescale(ctx_, ct->data->ptr, (int)ct->towers - 1, (int)ct->ncomp, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg > 1 ? ct->noiseScaleDeg - 1 : 1;
  out->scalingFactor = ct->scalingFactor / (flex() ? mrf_[ct->towers - 1] : scale_);
  out->slots = ct->slots;
  return out;
}

// OpenFHE LeveledSHECKKSRNS::AdjustLevelsAndDepthInPlace for an operand c1 with more towers than
// the other (towers2, deg2, scale2), as the reference restates it (src/CKKS/Ciphertext.cpp:1917-2014):
// a scalar product that brings the scale to what a rescale (or two) will turn into the other's,
// the rescales, level drops, and the other operand's scale copied over. Device level l = towers - 1;
// OpenFHE's GetScalingFactorRealBig(c2lvl - 1) is the square of the factor one level above c2.
Ciphertext CryptoContextImpl::adjustDown(const Ciphertext& c1, uint32_t towers2, uint32_t deg2, double scale2) const {
  const uint32_t l1 = c1->towers - 1, l2 = towers2 - 1;
  Ciphertext x;
  if (c1->noiseScaleDeg == 2) {
    if (deg2 == 2) {
      const double scf1 = c1->scalingFactor, scf2 = scale2, scf = sfReal_[l1], q1 = mrf_[l1];
      x = modReduce(multCore(c1, scf2 / scf1 * q1 / scf));
      if (x->towers > towers2) x = adjusted(x, towers2);
      x->scalingFactor = scale2;
    } else if (l1 - 1 == l2) {
      x = modReduce(c1);
    } else {
      const double scf1 = c1->scalingFactor, scf2 = sfBig_[l2 + 1], scf = sfReal_[l1], q1 = mrf_[l1];
      x = modReduce(multCore(c1, scf2 / scf1 * q1 / scf));
      if (x->towers > towers2 + 1) x = adjusted(x, towers2 + 1);
      x = modReduce(x);
      x->scalingFactor = scale2;
    }
  } else {
    if (deg2 == 2) {
      const double scf1 = c1->scalingFactor, scf2 = scale2, scf = sfReal_[l1];
      x = adjusted(multCore(c1, scf2 / scf1 / scf), towers2);
      x->scalingFactor = scf2;
    } else {
      const double scf1 = c1->scalingFactor, scf2 = sfBig_[l2 + 1], scf = sfReal_[l1];
      x = multCore(c1, scf2 / scf1 / scf);
      if (x->towers > towers2 + 1) x = adjusted(x, towers2 + 1);
      x = modReduce(x);
      x->scalingFactor = scale2;
    }
  }
  return x;
}

// one level and one depth: the operand with more towers is brought down; at equal levels the
// lower-depth operand is multiplied by 1.0 (EvalMultCore), raising its depth
void CryptoContextImpl::adjustLevelsAndDepth(Ciphertext& a, Ciphertext& b) const {
  if (a->noiseScaleDeg > 2 || b->noiseScaleDeg > 2) throw std::runtime_error("fideslib: operands of noise degree above 2");
  if (a->towers > b->towers) a = adjustDown(a, b->towers, b->noiseScaleDeg, b->scalingFactor);
  else if (a->towers < b->towers) b = adjustDown(b, a->towers, a->noiseScaleDeg, a->scalingFactor);
  else if (a->noiseScaleDeg < b->noiseScaleDeg) a = multCore(a, 1.0);
  else if (b->noiseScaleDeg < a->noiseScaleDeg) b = multCore(b, 1.0);
}

// AdjustLevelsAndDepthToOneInPlace: aligned, then both rescaled when at depth 2
void CryptoContextImpl::adjustLevelsAndDepthToOne(Ciphertext& a, Ciphertext& b) const {
  adjustLevelsAndDepth(a, b);
  if (a->noiseScaleDeg == 2) {
    a = modReduce(a);
    b = modReduce(b);
  }
}

Ciphertext CryptoContextImpl::morph(const Plaintext& pt) const {
  auto c = std::make_shared<CiphertextImpl>();
  c->data = pt->data;
  c->towers = pt->towers;
  c->ncomp = 1;
  c->noiseScaleDeg = pt->noiseScaleDeg;
  c->scalingFactor = pt->scalingFactor;
  c->slots = pt->slots;
  return c;
}

// a times the (aligned) plaintext p, component by component
Ciphertext CryptoContextImpl::mulPlain(const Ciphertext& a, const Ciphertext& p) const {
  const uint32_t t = std::min(a->towers, p->towers);
  auto out = newCt(t, a->ncomp);
  for (uint32_t c = 0; c < a->ncomp; ++c)
    check(fh_eltwise_binary(ctx_, 2, a->data->ptr + (size_t)c * a->towers * N_, p->data->ptr, out->data->ptr + (size_t)c * t * N_, nullptr,
                            (int)t));
  out->noiseScaleDeg = a->noiseScaleDeg + p->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor * p->scalingFactor;
  out->slots = a->slots;
  return out;
}

// a +/- the (aligned) plaintext p on component 0 (op 0 add, 1 sub)
Ciphertext CryptoContextImpl::addPlain(const Ciphertext& a, const Ciphertext& p, int op) const {
  const uint32_t t = std::min(a->towers, p->towers);
  auto out = newCt(t, a->ncomp);
  for (uint32_t c = 1; c < a->ncomp; ++c)
    check(fh_copy_d2d(ctx_, out->data->ptr + (size_t)c * t * N_, a->data->ptr + (size_t)c * a->towers * N_, (size_t)t * N_ * 8));
  check(fh_eltwise_binary(ctx_, op, a->data->ptr, p->data->ptr, out->data->ptr, nullptr, (int)t));
  out->noiseScaleDeg = a->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor;
  out->slots = a->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a0, const Plaintext& pt) {
  if (pt->extended) throw std::runtime_error("fideslib: extended plaintexts are for LinearTransform");
  if (flex()) {
    require_standard(*a0, "EvalMult");
    Ciphertext a = a0, p = morph(pt);
    adjustLevelsAndDepthToOne(a, p);   // OpenFHE AdjustForMult: aligned, then rescaled at depth 2
    return mulPlain(a, p);
  }
  const uint32_t t = std::min(a0->towers, pt->towers);
  auto a = adjusted(a0, t);
  auto out = newCt(t, a->ncomp);
  for (uint32_t c = 0; c < a->ncomp; ++c)
    check(fh_eltwise_binary(ctx_, 2, a->data->ptr + (size_t)c * t * N_, pt->data->ptr, out->data->ptr + (size_t)c * t * N_, nullptr, (int)t));
  out->noiseScaleDeg = a->noiseScaleDeg + pt->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor * pt->scalingFactor;
  out->slots = a->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalMultNoRelin(const Ciphertext& a1, const Ciphertext& b1) {
  if (a1->ncomp != 2 || b1->ncomp != 2) throw std::runtime_error("fideslib: EvalMultNoRelin: ciphertext should be relinearized before");
  Ciphertext a0 = a1, b0 = b1;
  if (flex()) adjustLevelsAndDepthToOne(a0, b0);
  const uint32_t t = std::min(a0->towers, b0->towers);
  auto a = adjusted(a0, t), b = adjusted(b0, t);
  auto out = newCt(t, 3);
  check(fh_tensor(ctx_, a->data->ptr, b->data->ptr, (int)t - 1, out->data->ptr));
  out->noiseScaleDeg = a->noiseScaleDeg + b->noiseScaleDeg;
  out->scalingFactor = a->scalingFactor * b->scalingFactor;
  out->slots = a->slots;
  return out;
}

Ciphertext CryptoContextImpl::Relinearize(const Ciphertext& ct) {
  require_standard(*ct, "Relinearize");
  if (ct->ncomp == 2) return ct;
  if (!relin_) throw std::runtime_error("fideslib: EvalMultKeyGen was not called");
  auto out = newCt(ct->towers, 2);
  check(fh_relinearize(ctx_, ct->data->ptr, (int)ct->towers - 1, relin_, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

// RequireDegree1 (api/CryptoContext.cpp:994-998), then the key-switch-first product
Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a1, const Ciphertext& b1) {
  require_standard(*a1, "EvalMult");
  require_standard(*b1, "EvalMult");
  if (a1->ncomp != 2 || b1->ncomp != 2)
    throw std::runtime_error("fideslib: EvalMult: ciphertext should be relinearized before (has 3 elements; call Relinearize)");
  if (!relin_) throw std::runtime_error("fideslib: EvalMultKeyGen was not called");
  Ciphertext a0 = a1, b0 = b1;
  // FLEXIBLEAUTO: OpenFHE AdjustForMult (one level and depth, depth-2 operands rescaled); the product
  // is left at depth 2 and rescaled by the next operation that needs it
  if (flex()) adjustLevelsAndDepthToOne(a0, b0);
  // operands read at their own tower counts (no level-reduced copies)
  const uint32_t t = std::min(a0->towers, b0->towers);
  auto out = newCt(t, 2);
  check(fh_eval_mult_many(ctx_, 1, a0->data->ptr, (int)a0->towers, b0->data->ptr, (int)b0->towers, (int)t - 1, relin_, out->data->ptr));
  out->noiseScaleDeg = a0->noiseScaleDeg + b0->noiseScaleDeg;
  out->scalingFactor = a0->scalingFactor * b0->scalingFactor;
  out->slots = a0->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalSquare(const Ciphertext& ct) { return EvalMult(ct, ct); }

Ciphertext CryptoContextImpl::EvalLinearWSum(const std::vector<Ciphertext>& cts0, const std::vector<double>& weights) {
  if (cts0.empty() || cts0.size() != weights.size()) throw std::runtime_error("fideslib: EvalLinearWSum needs one weight per ciphertext");
  std::vector<Ciphertext> cts = cts0;
  if (flex()) {
    // OpenFHE EvalLinearWSumMutable under FLEXIBLEAUTO: every input aligned with the one of fewest
    // towers, all rescaled when that one is at depth 2
    size_t mi = 0;
    for (size_t i = 1; i < cts.size(); ++i)
      if (cts[i]->towers < cts[mi]->towers) mi = i;
    for (size_t i = 0; i < cts.size(); ++i)
      if (i != mi) adjustLevelsAndDepth(cts[i], cts[mi]);
    if (cts[mi]->noiseScaleDeg == 2)
      for (auto& c : cts) c = modReduce(c);
  }
  uint32_t t = ~0u;
  for (const auto& c : cts) {
    if (c->ncomp != cts[0]->ncomp || c->noiseScaleDeg != cts[0]->noiseScaleDeg)
      throw std::runtime_error("fideslib: EvalLinearWSum: inputs differ in size or scale degree");
    t = std::min(t, c->towers);
  }
  const uint32_t nc = cts[0]->ncomp;
  Ciphertext acc;
  for (size_t k0 = 0; k0 < cts.size(); k0 += 16) {
    const size_t n = std::min<size_t>(16, cts.size() - k0);
    std::vector<const uint64_t*> ptrs(n);
    std::vector<int32_t> towers(n);
    std::vector<uint64_t> w(n * t);
    for (size_t k = 0; k < n; ++k) {
      ptrs[k] = cts[k0 + k]->data->ptr;
      towers[k] = (int32_t)cts[k0 + k]->towers;
      const auto f = ElemForEvalMult(t, weights[k0 + k]);
      std::copy(f.begin(), f.end(), w.begin() + k * t);
    }
    auto out = newCt(t, nc);
    check(fh_linear_wsum(ctx_, (int)n, ptrs.data(), towers.data(), w.data(), (int)t, (int)nc, nullptr, 0, out->data->ptr));
    out->noiseScaleDeg = cts[0]->noiseScaleDeg + 1;
    out->scalingFactor = cts[0]->scalingFactor * sfReal(t);
    out->slots = cts[0]->slots;
    acc = acc ? EvalAdd(acc, out) : out;
  }
  return acc;
}

Ciphertext CryptoContextImpl::EvalAddMany(const std::vector<Ciphertext>& cts) {
  if (cts.empty()) throw std::runtime_error("fideslib: EvalAddMany of an empty list");
  Ciphertext acc = cts[0];
  for (size_t i = 1; i < cts.size(); ++i) acc = EvalAdd(acc, cts[i]);
  return acc;
}

// Ciphertext::rescale (src/CKKS/Ciphertext.cpp:569-607): drop the top tower, degree - 1, and divide
// the tracked scale by OpenFHE's ModReduceFactor, which the FIXED techniques set to the nominal
// scaling factor 2^scaleModSize (not q_l).
// Under FLEXIBLEAUTO the operations rescale when needed and Rescale does nothing (OpenFHE's
// ModReduceInPlace acts under FIXEDMANUAL only).
Ciphertext CryptoContextImpl::Rescale(const Ciphertext& ct) {
  require_standard(*ct, "Rescale");
  if (flex()) return ct;
  return modReduce(ct);
}

// OpenFHE FindAutomorphismIndex2nComplex: 5^index mod 2N, 5^-|index| for negative indices; 5 has
// order N/2 in Z_2N^*, so the index is reduced mod N/2 (not mod the slot count).
uint32_t CryptoContextImpl::galoisFor(int32_t index) const { return FindAutomorphismIndex(normalize_index(index, N_ / 2)); }

Ciphertext CryptoContextImpl::EvalRotate(const Ciphertext& ct, int32_t index) {
  require_standard(*ct, "EvalRotate");
  if (ct->ncomp != 2) throw std::runtime_error("fideslib: EvalRotate: ciphertext should be relinearized before");
  if (normalize_index(index, N_ / 2) == 0) return ct;
  const uint32_t g = galoisFor(index);
  auto out = newCt(ct->towers, 2);
  check(fh_eval_rotate(ctx_, ct->data->ptr, (int)ct->towers - 1, rotKey(g), g, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

Ciphertext CryptoContextImpl::EvalConjugate(const Ciphertext& ct) {
  const uint32_t g = 2 * N_ - 1;
  auto out = newCt(ct->towers, 2);
  check(fh_eval_rotate(ctx_, ct->data->ptr, (int)ct->towers - 1, rotKey(g), g, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

namespace {
struct FastRotationDigits {
  std::shared_ptr<DeviceBuffer> digits;
  uint32_t towers;
};
}  // namespace

// EvalFastRotationPrecompute: the shared ModUp of c1 (CudaEngine.cpp:85-106)
std::shared_ptr<void> CryptoContextImpl::EvalFastRotationPrecompute(const Ciphertext& ct) {
  require_standard(*ct, "EvalFastRotationPrecompute");
  const int l = (int)ct->towers - 1;
  const int beta = fh_num_digits(ctx_, l);
  if (beta < 0) check(beta);
  auto h = std::make_shared<FastRotationDigits>();
  h->towers = ct->towers;
  h->digits = std::make_shared<DeviceBuffer>(native_, (size_t)beta * (ct->towers + K_) * N_);
  check(fh_modup(ctx_, ct->data->ptr + (size_t)ct->towers * N_, l, h->digits->ptr));
  return h;
}

Ciphertext CryptoContextImpl::EvalFastRotation(const Ciphertext& ct, int32_t index, uint32_t, const std::shared_ptr<void>& precomp) {
  auto h = std::static_pointer_cast<FastRotationDigits>(precomp);
  if (!h || h->towers != ct->towers) throw std::runtime_error("fideslib: stale EvalFastRotation precomputation");
  if (normalize_index(index, N_ / 2) == 0) return ct;
  const uint32_t g = galoisFor(index);
  const uint32_t ext = ct->towers + K_;
  DeviceBuffer e(native_, (size_t)2 * ext * N_);
  check(fh_key_inner_product(ctx_, h->digits->ptr, (int)ct->towers - 1, rotKey(g), ct->data->ptr, g, e.ptr, e.ptr + (size_t)ext * N_));
  auto out = newCt(ct->towers, 2);
  check(fh_moddown(ctx_, e.ptr, (int)ct->towers - 1, 2, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

std::vector<Ciphertext> CryptoContextImpl::EvalFastRotation(const Ciphertext& ct, const std::vector<int32_t>& indices, uint32_t m,
                                                            const std::shared_ptr<void>& precomp) {
  std::vector<Ciphertext> out;
  for (int32_t i : indices) out.push_back(EvalFastRotation(ct, i, m, precomp));
  return out;
}

Ciphertext CryptoContextImpl::EvalFastRotationExt(const Ciphertext& ct, int32_t index, const std::shared_ptr<void>& precomp, bool addFirst) {
  if (ct->ncomp != 2 || ct->extended) throw std::runtime_error("fideslib: EvalFastRotationExt: ciphertext should be relinearized, not extended");
  auto h = std::static_pointer_cast<FastRotationDigits>(precomp);
  if (!h || h->towers != ct->towers) throw std::runtime_error("fideslib: stale EvalFastRotation precomputation");
  if (normalize_index(index, N_ / 2) == 0) throw std::runtime_error("fideslib: EvalFastRotationExt by a multiple of N/2 is not supported");
  const uint32_t ext = ct->towers + K_;
  const uint32_t g = galoisFor(index);
  auto out = std::make_shared<CiphertextImpl>();
  out->data = std::make_shared<DeviceBuffer>(native_, (size_t)2 * ext * N_);
  check(fh_key_inner_product(ctx_, h->digits->ptr, (int)ct->towers - 1, rotKey(g), addFirst ? ct->data->ptr : nullptr, g, out->data->ptr,
                             out->data->ptr + (size_t)ext * N_));
  out->towers = ct->towers;
  out->ncomp = 2;
  out->extended = true;
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

std::vector<Ciphertext> CryptoContextImpl::EvalFastRotationExt(const Ciphertext& ct, const std::vector<int32_t>& indices,
                                                               const std::shared_ptr<void>& precomp, bool addFirst) {
  std::vector<Ciphertext> out;
  for (int32_t i : indices) out.push_back(EvalFastRotationExt(ct, i, precomp, addFirst));
  return out;
}

Ciphertext CryptoContextImpl::KeySwitchDown(const Ciphertext& x) {
  if (!x->extended) throw std::runtime_error("fideslib: KeySwitchDown needs an extended ciphertext");
  auto out = newCt(x->towers, 2);
  check(fh_moddown(ctx_, x->data->ptr, (int)x->towers - 1, 2, out->data->ptr));
  out->noiseScaleDeg = x->noiseScaleDeg;
  out->scalingFactor = x->scalingFactor;
  out->slots = x->slots;
  return out;
}

// EvalRotateMany: one batched launch when every ciphertext shares the level (CudaEngine.cpp:656-702)
std::vector<Ciphertext> CryptoContextImpl::EvalRotateMany(const std::vector<Ciphertext>& cts, int32_t index) {
  std::vector<Ciphertext> out;
  if (cts.empty()) return out;
  bool same = true;
  for (auto& c : cts) same &= c->towers == cts[0]->towers && c->ncomp == 2 && c->slots == cts[0]->slots;
  if (!same || normalize_index(index, N_ / 2) == 0) {
    for (auto& c : cts) out.push_back(EvalRotate(c, index));
    return out;
  }
  const uint32_t g = galoisFor(index), t = cts[0]->towers;
  const size_t ctw = (size_t)2 * t * N_;
  DeviceBuffer in(native_, cts.size() * ctw), res(native_, cts.size() * ctw);
  for (size_t i = 0; i < cts.size(); ++i) check(fh_copy_d2d(ctx_, in.ptr + i * ctw, cts[i]->data->ptr, ctw * 8));
  std::vector<const fh_key*> keys(cts.size(), rotKey(g));
  std::vector<uint32_t> gal(cts.size(), g);
  check(fh_rotate_batch(ctx_, (int)cts.size(), in.ptr, 0, (int)t - 1, keys.data(), gal.data(), res.ptr));
  for (size_t i = 0; i < cts.size(); ++i) {
    auto o = newCt(t, 2);
    check(fh_copy_d2d(ctx_, o->data->ptr, res.ptr + i * ctw, ctw * 8));
    o->noiseScaleDeg = cts[i]->noiseScaleDeg;
    o->scalingFactor = cts[i]->scalingFactor;
    o->slots = cts[i]->slots;
    out.push_back(o);
  }
  return out;
}

CryptoContextImpl::PackedDiagonals CryptoContextImpl::PackDiagonals(const std::vector<Plaintext>& diags, int rowSize) const {
  if (rowSize < 1 || (int)diags.size() < rowSize) throw std::runtime_error("fideslib: LinearTransform needs rowSize diagonals");
  const uint32_t t = diags[0]->towers, ext = t + K_;
  for (int k = 0; k < rowSize; ++k)
    if (!diags[k]->extended || diags[k]->towers != t) throw std::runtime_error("fideslib: diagonals must be extended plaintexts at one level");
  PackedDiagonals p;
  p.data = std::make_shared<DeviceBuffer>(native_, (size_t)rowSize * ext * N_);
  for (int k = 0; k < rowSize; ++k) check(fh_copy_d2d(ctx_, p.data->ptr + (size_t)k * ext * N_, diags[k]->data->ptr, (size_t)ext * N_ * 8));
  p.rowSize = (uint32_t)rowSize;
  p.towers = t;
  p.noiseScaleDeg = diags[0]->noiseScaleDeg;
  p.scalingFactor = diags[0]->scalingFactor;
  return p;
}

Ciphertext CryptoContextImpl::LinearTransform(const Ciphertext& ct0, int rowSize, int bStep, const std::vector<Plaintext>& diags,
                                              int stride, int offset) {
  if ((int)diags.size() < rowSize) throw std::runtime_error("fideslib: LinearTransform needs rowSize diagonals");
  Ciphertext ct = ct0->noiseScaleDeg == 2 ? modReduce(ct0) : ct0;  // rescale a fresh product first (LinearTransform.cu:77-78)
  for (int k = 0; k < rowSize; ++k)
    if (!diags[k]->extended || diags[k]->towers != ct->towers)
      throw std::runtime_error("fideslib: diagonals must be extended plaintexts at the ciphertext's level");
  return LinearTransform(ct, bStep, PackDiagonals(diags, rowSize), stride, offset);
}

CryptoContextImpl::PreparedLinearTransform CryptoContextImpl::PrepareLinearTransform(const std::vector<Plaintext>& diagonals, int rowSize,
                                                                                     int bStep, int stride, int offset) const {
  if (bStep < 1 || bStep > rowSize) throw std::runtime_error("fideslib: PrepareLinearTransform: bStep must be 1..rowSize");
  auto p = std::make_shared<PreparedLinearTransformImpl>();
  p->rowSize = rowSize;
  p->bStep = bStep;
  p->stride = stride;
  p->offset = offset;
  p->diagonals = PackDiagonals(diagonals, rowSize);
  return p;
}

Ciphertext CryptoContextImpl::LinearTransform(const Ciphertext& ct, const PreparedLinearTransform& prepared) {
  if (!prepared) throw std::runtime_error("fideslib: LinearTransform: empty prepared transform");
  return LinearTransform(ct, prepared->bStep, prepared->diagonals, prepared->stride, prepared->offset);
}

void CryptoContextImpl::LinearTransformInPlace(Ciphertext& ct, const PreparedLinearTransform& prepared) { ct = LinearTransform(ct, prepared); }

// ---- rotation-fold sum ----
std::vector<int32_t> CryptoContextImpl::GetAccumulateSumRotationIndices(int slots, int stride, int start) {
  const int n = fh_accumulate_indices(kAccumulateSumRadix, stride, slots, start, nullptr, 0);
  if (n < 0) throw std::runtime_error(std::string("fideslib: ") + fh_last_error());
  std::vector<int32_t> v(n);
  if (n) fh_accumulate_indices(kAccumulateSumRadix, stride, slots, start, v.data(), n);
  return v;
}

// start 0: the lazy fold (CKKS::Accumulate(ct, bStep, stride, size) behind accumulateSum); start
// >= 1: the cascade from that factor (Accumulate(ct, 4, stride, size, start) behind
// accumulateSumInPlace(ct, slots, stride, start), api/engine/cuda/CudaEngine.cpp:780-784)
Ciphertext CryptoContextImpl::accumulate(const Ciphertext& ct, int slots, int stride, int start) {
  require_standard(*ct, "AccumulateSum");
  if (ct->ncomp != 2) throw std::runtime_error("fideslib: AccumulateSum: ciphertext should be relinearized before");
  const auto idx = GetAccumulateSumRotationIndices(slots, stride, start == 0 ? 1 : start);
  std::vector<const fh_key*> keys(idx.size());
  std::vector<uint32_t> gal(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    gal[i] = galoisFor(idx[i]);
    keys[i] = rotKey(gal[i]);
  }
  auto out = newCt(ct->towers, 2);
  if (start == 0)
    check(fh_accumulate_sum(ctx_, ct->data->ptr, (int)ct->towers - 1, kAccumulateSumRadix, stride, slots, keys.data(), gal.data(),
                            (int)idx.size(), out->data->ptr));
  else
    check(fh_accumulate_sum_from(ctx_, ct->data->ptr, (int)ct->towers - 1, kAccumulateSumRadix, stride, slots, start, keys.data(),
                                 gal.data(), (int)idx.size(), out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

Ciphertext CryptoContextImpl::AccumulateSum(const Ciphertext& ct, int slots, int stride) { return accumulate(ct, slots, stride, 0); }
void CryptoContextImpl::AccumulateSumInPlace(Ciphertext& ct, int slots, int stride) { ct = accumulate(ct, slots, stride, 0); }
void CryptoContextImpl::AccumulateSumInPlace(Ciphertext& ct, int slots, int stride, int start) {
  const uint32_t before = ct->slots;
  if (start <= 0) return;  // the reference's cascade returns at once for startFactor <= 0
  ct = accumulate(ct, slots, stride, start);
  if (start > 0 && slots > 0 && (long long)slots * stride == before) ct->slots = (uint32_t)(stride * start);
}

CryptoContextImpl::PreparedLinearTransform CryptoContextImpl::PrepareLinearTransformMany(const std::vector<std::vector<Plaintext>>& sets,
                                                                                         int rowSize, int bStep, int stride, int offset) const {
  if (sets.empty()) throw std::runtime_error("fideslib: PrepareLinearTransformMany needs at least one diagonal set");
  if (bStep < 1 || bStep > rowSize) throw std::runtime_error("fideslib: PrepareLinearTransformMany: bStep must be 1..rowSize");
  auto p = std::make_shared<PreparedLinearTransformImpl>();
  p->rowSize = rowSize;
  p->bStep = bStep;
  p->stride = stride;
  p->offset = offset;
  p->diagonals = PackDiagonals(sets[0], rowSize);
  for (size_t i = 1; i < sets.size(); ++i) {
    p->more.push_back(PackDiagonals(sets[i], rowSize));
    if (p->more.back().towers != p->diagonals.towers) throw std::runtime_error("fideslib: PrepareLinearTransformMany: sets at different levels");
  }
  return p;
}

std::vector<Ciphertext> CryptoContextImpl::LinearTransformMany(const Ciphertext& ct, const PreparedLinearTransform& prepared) {
  if (!prepared) throw std::runtime_error("fideslib: LinearTransformMany: empty prepared transform");
  std::vector<Ciphertext> out{LinearTransform(ct, prepared->bStep, prepared->diagonals, prepared->stride, prepared->offset)};
  for (const auto& d : prepared->more) out.push_back(LinearTransform(ct, prepared->bStep, d, prepared->stride, prepared->offset));
  return out;
}

std::vector<int> CryptoContextImpl::GetLinearTransformRotationIndices(int bStep, int stride, int offset) {
  std::vector<int> r;
  for (int i = 1; i <= bStep; ++i) r.push_back(i * stride);
  if (offset != 0) r.push_back(offset);
  return r;
}

std::vector<int> CryptoContextImpl::GetLinearTransformPlaintextRotationIndices(int rowSize, int bStep, int stride, int offset) {
  std::vector<int> r(rowSize);
  for (int k = 0; k < rowSize; ++k) r[k] = -bStep * (k / bStep) * stride - offset;
  return r;
}

std::vector<Plaintext> CryptoContextImpl::MakeCKKSPackedPlaintextDeviceMany(const std::vector<std::vector<double>>& values, size_t noiseScaleDeg,
                                                                           uint32_t level, uint32_t slots) {
  std::vector<Plaintext> out;
  if (values.empty()) return out;
  if (level >= L1_) throw std::runtime_error("fideslib: level out of range");
  const uint32_t n = slots ? slots : (params_.batch_ ? params_.batch_ : N_ / 2);
  const uint32_t towers = L1_ - level;
  const double scale = ptScale(towers, noiseScaleDeg);
  // one batched device encode into a staging buffer, then one buffer per plaintext
  std::vector<double> z(values.size() * 2 * (size_t)n, 0.0);
  for (size_t b = 0; b < values.size(); ++b)
    for (size_t i = 0; i < std::min<size_t>(values[b].size(), n); ++i) z[(b * n + i) * 2] = values[b][i];
  std::vector<int32_t> pids(towers);
  for (uint32_t i = 0; i < towers; ++i) pids[i] = (int32_t)i;
  const size_t words = (size_t)towers * N_;
  DeviceBuffer staging(native_, values.size() * words);
  check(fh_encode_slots(ctx_, z.data(), (int)values.size(), (int)n, scale, pids.data(), (int)towers, staging.ptr, words));
  for (size_t b = 0; b < values.size(); ++b) {
    auto pt = std::make_shared<PlaintextImpl>();
    pt->towers = towers;
    pt->slots = n;
    pt->noiseScaleDeg = (uint32_t)noiseScaleDeg;
    pt->scalingFactor = scale;
    pt->values.assign(n, 0.0);
    for (size_t i = 0; i < std::min<size_t>(values[b].size(), n); ++i) pt->values[i] = values[b][i];
    pt->data = std::make_shared<DeviceBuffer>(native_, words);
    check(fh_copy_d2d(ctx_, pt->data->ptr, staging.ptr + b * words, words * 8));
    out.push_back(pt);
  }
  return out;
}

Ciphertext CryptoContextImpl::Compress(const Ciphertext& ct, uint32_t towersLeft, size_t noiseScaleDeg) {
  Ciphertext c = ct;
  while (c->noiseScaleDeg > noiseScaleDeg && c->towers > 1) c = modReduce(c);
  if (c->towers > towersLeft && towersLeft >= 1) c = LevelReduce(c, c->towers - towersLeft);
  return c;
}

Ciphertext CryptoContextImpl::LinearTransform(const Ciphertext& ct0, int bStep, const PackedDiagonals& diags, int stride, int offset) {
  require_standard(*ct0, "LinearTransform");
  Ciphertext ct = ct0->noiseScaleDeg == 2 ? modReduce(ct0) : ct0;  // rescale a fresh product first (LinearTransform.cu:77-78)
  const uint32_t t = ct->towers;
  if (!diags.data || diags.towers != t) throw std::runtime_error("fideslib: diagonals must be extended plaintexts at the ciphertext's level");
  const int rowSize = (int)diags.rowSize;
  const int slots = (int)ct->slots;
  std::vector<const fh_key*> bk(bStep, nullptr);
  std::vector<uint32_t> bg(bStep, 1);
  for (int i = 1; i < bStep; ++i) {
    if (normalize_index(i * stride, slots) == 0) continue;
    bg[i] = galoisFor(i * stride);
    bk[i] = rotKey(bg[i]);
  }
  const fh_key* gk = nullptr;
  uint32_t gg = 1;
  const int gStep = (rowSize + bStep - 1) / bStep;
  if (gStep > 1 && normalize_index(bStep * stride, slots) != 0) {
    gg = galoisFor(bStep * stride);
    gk = rotKey(gg);
  }
  const fh_key* ok = nullptr;
  uint32_t og = 1;
  if (offset != 0) {
    og = galoisFor(offset);
    ok = rotKey(og);
  }
  auto out = newCt(t, 2);
  check(fh_linear_transform(ctx_, ct->data->ptr, (int)t - 1, rowSize, bStep, stride, offset, slots, diags.data->ptr, bk.data(), bg.data(), gk,
                            gg, ok, og, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg + diags.noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor * diags.scalingFactor;
  out->slots = ct->slots;
  return out;
}

Ciphertext CryptoContextImpl::LinearTransformFlat(const Ciphertext& ct0, int bStep, const PackedDiagonals& diags, int stride, int offset) {
  require_standard(*ct0, "LinearTransform");
  Ciphertext ct = ct0->noiseScaleDeg == 2 ? modReduce(ct0) : ct0;
  const uint32_t t = ct->towers;
  if (!diags.data || diags.towers != t) throw std::runtime_error("fideslib: diagonals must be extended plaintexts at the ciphertext's level");
  const int rowSize = (int)diags.rowSize;
  const int slots = (int)ct->slots;
  std::vector<const fh_key*> bk(bStep, nullptr);
  std::vector<uint32_t> bg(bStep, 1);
  for (int i = 1; i < bStep; ++i) {
    if (normalize_index(i * stride, slots) == 0) continue;
    bg[i] = galoisFor(i * stride);
    bk[i] = rotKey(bg[i]);
  }
  const int gStep = (rowSize + bStep - 1) / bStep;
  std::vector<const fh_key*> gk(std::max(gStep - 1, 1), nullptr);
  std::vector<uint32_t> gg(gk.size(), 1);
  for (int j = 1; j < gStep; ++j) {
    const int32_t r = normalize_index(j * bStep * stride, slots);  // the rotation on `slots` replicated slots
    if (r == 0) continue;
    gg[j - 1] = galoisFor(r);
    gk[j - 1] = rotKey(gg[j - 1]);
  }
  const fh_key* ok = nullptr;
  uint32_t og = 1;
  if (offset != 0) {
    og = galoisFor(offset);
    ok = rotKey(og);
  }
  auto out = newCt(t, 2);
  check(fh_linear_transform_flat(ctx_, ct->data->ptr, (int)t - 1, rowSize, bStep, stride, offset, slots, diags.data->ptr, bk.data(),
                                 bg.data(), gk.data(), gg.data(), ok, og, out->data->ptr));
  out->noiseScaleDeg = ct->noiseScaleDeg + diags.noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor * diags.scalingFactor;
  out->slots = ct->slots;
  return out;
}

}  // namespace fideslib
