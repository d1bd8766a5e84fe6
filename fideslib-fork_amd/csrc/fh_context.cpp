// Host construction of the device context: product-side number theory, twiddle
// tables and the basis-conversion constants of OpenFHE's CKKS-RNS parameters.
// Reference: src/ConstantsGPU.cu:131-698 (SetupConstants), constants copied there
// from OpenFHE CryptoParametersCKKSRNS (src/ConstantsGPU.cu:441-636).
#include "fh_context.hpp"

#include <algorithm>
#include <cstring>

namespace fh {

typedef unsigned __int128 u128;

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  a %= q;
  for (; e; e >>= 1) {
    if (e & 1) r = h_mulmod(r, a, q);
    a = h_mulmod(a, a, q);
  }
  return r;
}

uint64_t h_invmod(uint64_t a, uint64_t q) {
  // q prime in every use here: Fermat
  return h_powmod(a % q, q - 2, q);
}

uint64_t h_shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

uint64_t h_inv_odd(uint64_t g, uint64_t m) {
  // Newton iteration x <- x (2 - g x) doubles the correct low bits; g x = 1 mod 8 for x = g
  uint64_t x = g;
  for (int i = 0; i < 6; ++i) x *= 2 - g * x;
  return x & (m - 1);
}

bool h_is_prime(uint64_t n) {
  if (n < 2) return false;
  const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t b : bases) {
    if (n == b) return true;
    if (n % b == 0) return false;
  }
  uint64_t d = n - 1;
  int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (uint64_t b : bases) {
    uint64_t x = h_powmod(b, d, n);
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (int r = 1; r < s && composite; ++r) {
      x = h_mulmod(x, x, n);
      if (x == n - 1) composite = false;
    }
    if (composite) return false;
  }
  return true;
}

// Smallest primitive m-th root of unity mod q (m a power of two): OpenFHE
// RootOfUnity returns the minimum over all primitive roots.
uint64_t h_min_root_of_unity(uint64_t m, uint64_t q) {
  if ((q - 1) % m) throw Error(-2, "modulus is not 1 mod 2N");
  // m | q-1 with m a power of two: x^((q-1)/m) has order m iff x is a non-residue-like witness;
  // test candidates until the m/2-th power is -1.
  uint64_t w = 0;
  for (uint64_t g = 2;; ++g) {
    w = h_powmod(g, (q - 1) / m, q);
    if (h_powmod(w, m / 2, q) == q - 1) break;
  }
  uint64_t w2 = h_mulmod(w, w, q), best = w, cur = w;
  for (uint64_t k = 3; k < m; k += 2) {
    cur = h_mulmod(cur, w2, q);
    best = std::min(best, cur);
  }
  return best;
}

static uint64_t first_prime(int bits, uint64_t m) {
  uint64_t q = (uint64_t)1 << bits, r = q % m, c = q + 1 - r + (r ? m : 0);
  while (!h_is_prime(c)) c += m;
  return c;
}
static uint64_t last_prime(int bits, uint64_t m) {
  uint64_t q = (uint64_t)1 << bits, r = q % m, c = q + 1 - r - m;
  while (!h_is_prime(c)) c -= m;
  return c;
}
static uint64_t prev_prime(uint64_t q, uint64_t m) {
  uint64_t c = q - m;
  while (!h_is_prime(c)) c -= m;
  return c;
}
static uint64_t next_prime(uint64_t q, uint64_t m) {
  uint64_t c = q + m;
  while (!h_is_prime(c)) c += m;
  return c;
}

std::vector<uint64_t> h_gen_primes(int logN, int L1, int first_bits, int scale_bits, int K, int aux_bits) {
  uint64_t m = (uint64_t)2 << logN;
  std::vector<uint64_t> out(L1 + K);
  uint64_t top = first_prime(scale_bits, m), qn = top, qp = top;
  out[L1 - 1] = top;
  for (int i = L1 - 2, cnt = 0; i >= 1; --i, ++cnt) {
    if (cnt % 2 == 0) out[i] = qp = prev_prime(qp, m);
    else out[i] = qn = next_prime(qn, m);
  }
  if (L1 > 1) out[0] = (first_bits == scale_bits) ? next_prime(qn, m) : last_prime(first_bits, m);
  uint64_t pp = first_prime(aux_bits, m);
  for (int k = 0; k < K; ++k) {
    bool inq;
    do {
      pp = prev_prime(pp, m);
      inq = std::find(out.begin(), out.begin() + L1, pp) != out.begin() + L1;
    } while (inq);
    out[L1 + k] = pp;
  }
  return out;
}

int h_default_num_special(const std::vector<uint64_t>& q, int L1, int dnum, int aux_bits) {
  int alpha = (L1 + dnum - 1) / dnum, maxbits = 0;
  for (int d = 0; d < dnum; ++d) {
    std::vector<uint64_t> acc{1};
    for (int i = d * alpha; i < std::min((d + 1) * alpha, L1); ++i) {
      u128 carry = 0;
      for (auto& w : acc) {
        u128 v = (u128)w * q[i] + carry;
        w = (uint64_t)v;
        carry = v >> 64;
      }
      if (carry) acc.push_back((uint64_t)carry);
    }
    int bits = 64 * ((int)acc.size() - 1) + (64 - __builtin_clzll(acc.back()));
    maxbits = std::max(maxbits, bits);
  }
  return (maxbits + aux_bits - 1) / aux_bits;
}

// ---- Pool ----
Pool::~Pool() { release_all(); }

void Pool::reap(bool wait_one) {
  while (!pending_.empty()) {
    Pending& f = pending_.front();
    if (wait_one) {
      FH_HIP(hipEventSynchronize(f.ev));
      wait_one = false;
    } else {
      const hipError_t e = hipEventQuery(f.ev);
      if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // a pending event is not an error
        return;
      }
      FH_HIP(e);
    }
    free_.emplace(f.bytes, f.p);
    cached_ += f.bytes;
    pending_bytes_ -= f.bytes;
    spare_ev_.push_back(f.ev);
    pending_.pop_front();
  }
}

void Pool::put_after(void* p, hipStream_t st) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  hipEvent_t ev;
  if (!spare_ev_.empty()) {
    ev = spare_ev_.back();
    spare_ev_.pop_back();
  } else {
    FH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  FH_HIP(hipEventRecord(ev, st));
  pending_.push_back(Pending{p, it->second, ev});
  in_use_ -= it->second;
  pending_bytes_ += it->second;
}

void* Pool::get(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(mu_);
  if (!pending_.empty()) reap(false);
  auto it = free_.find(bytes);
  while (it == free_.end() && pending_bytes_ > kPendingCap) {
    reap(true);
    it = free_.find(bytes);
  }
  if (it != free_.end()) {
    void* p = it->second;
    free_.erase(it);
    cached_ -= bytes;
    in_use_ += bytes;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // wait for the pending blocks, give back every cached block and retry once
    while (!pending_.empty()) reap(true);
    for (auto& kv : free_) { (void)hipFree(kv.second); sizes_.erase(kv.second); }
    free_.clear();
    cached_ = 0;
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw Error(-4, "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  sizes_[p] = bytes;
  in_use_ += bytes;
  return p;
}

void Pool::hold() {
  std::lock_guard<std::mutex> lk(mu_);
  ++hold_;
}

void Pool::release_held() {
  std::lock_guard<std::mutex> lk(mu_);
  if (hold_ > 0 && --hold_ == 0) {
    for (void* p : held_) {
      auto it = sizes_.find(p);
      if (it == sizes_.end()) continue;
      free_.emplace(it->second, p);
      in_use_ -= it->second;
      cached_ += it->second;
    }
    held_.clear();
  }
}

void Pool::put(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  if (hold_ > 0) {
    held_.push_back(p);
    return;
  }
  free_.emplace(it->second, p);
  in_use_ -= it->second;
  cached_ += it->second;
}

void Pool::release_all() {
  std::lock_guard<std::mutex> lk(mu_);
  for (Pending& f : pending_) {
    (void)hipEventSynchronize(f.ev);
    free_.emplace(f.bytes, f.p);
    spare_ev_.push_back(f.ev);
  }
  pending_.clear();
  pending_bytes_ = 0;
  for (auto& kv : free_) { (void)hipFree(kv.second); sizes_.erase(kv.second); }
  free_.clear();
  cached_ = 0;
  for (hipEvent_t ev : spare_ev_) (void)hipEventDestroy(ev);
  spare_ev_.clear();
}

// ---- UploadRing ----
UploadRing::~UploadRing() {
  for (Slot& g : slot_) {
    if (g.ev) {
      if (g.used) (void)hipEventSynchronize(g.ev);
      (void)hipEventDestroy(g.ev);
    }
    if (g.h) (void)hipHostFree(g.h);
  }
}

hipEvent_t UploadRing::upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
  std::lock_guard<std::mutex> lk(mu_);
  Slot& g = slot_[next_];
  next_ = (next_ + 1) % kSlots;
  if (!g.ev) FH_HIP(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
  if (g.used) FH_HIP(hipEventSynchronize(g.ev));  // the copy that last read this slot
  if (g.cap < bytes) {
    if (g.h) FH_HIP(hipHostFree(g.h));
    g.h = nullptr;
    g.cap = 0;
    const size_t cap = std::max<size_t>(2 * bytes, (size_t)1 << 20);  // rare regrowth: a slot serves tables of any size
    FH_HIP(hipHostMalloc(&g.h, cap, hipHostMallocDefault));
    g.cap = cap;
  }
  std::memcpy(g.h, src, bytes);
  FH_HIP(hipMemcpyAsync(dst, g.h, bytes, hipMemcpyHostToDevice, st));
  FH_HIP(hipEventRecord(g.ev, st));
  g.used = true;
  return g.ev;
}

// ---- Profiler ----
hipEvent_t Profiler::take() {
  if (!spare.empty()) {
    hipEvent_t e = spare.back();
    spare.pop_back();
    return e;
  }
  hipEvent_t e;
  FH_HIP(hipEventCreate(&e));
  return e;
}

Profiler::~Profiler() {
  for (auto& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : spare) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(const Context& ctx, const char* n, hipStream_t s) : c(ctx), name(n), st(s) {
  if (!c.prof.on) return;
  std::lock_guard<std::mutex> lk(c.prof.mu);
  b = c.prof.take();
  FH_HIP(hipEventRecord(b, st));
}

ProfScope::~ProfScope() {
  if (!b) return;
  std::lock_guard<std::mutex> lk(c.prof.mu);
  hipEvent_t e = c.prof.take();
  if (hipEventRecord(e, st) == hipSuccess) c.prof.recs.push_back(Profiler::Rec{name, b, e});
}

// ---- Context ----
template <class T>
T* Context::upload(const std::vector<T>& v) {
  T* d = nullptr;
  size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  FH_HIP(hipMalloc(&d, bytes));
  owned.push_back(d);
  if (!v.empty()) FH_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

static uint32_t bitrev(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
  return r;
}

void Context::ensure_twists(int l1) {
  if (l1 < 0 || l1 > 8 || l1 >= logN) throw Error(-1, "radix split out of range");
  TwistSet& t = twists[l1];
  if (t.tw) return;
  // tau_r = psi^e_r, e_r = 1 + 2 brev_l1(r) - S1 (mod 2N)
  const uint64_t m = 2ull * N;
  const uint32_t S1 = 1u << l1, S2 = N >> l1;
  std::vector<uint64_t> ts((size_t)np * N), tssh(ts.size()), its(ts.size()), itssh(ts.size());
  std::vector<double> tsd(ts.size()), itsd(ts.size());
  for (int i = 0; i < np; ++i) {
    const uint64_t qi = q[i];
    for (uint32_t r = 0; r < S1; ++r) {
      const int64_t e = (1 + 2 * (int64_t)bitrev(r, l1) - (int64_t)S1) % (int64_t)m;
      const uint64_t eu = (uint64_t)(e < 0 ? e + (int64_t)m : e);
      const uint64_t step = h_powmod(psi[i], eu, qi), istep = h_powmod(psi[i], (m - eu) % m, qi);
      uint64_t x = 1, xi = 1;
      for (uint32_t c = 0; c < S2; ++c) {
        const size_t o = (size_t)i * N + (size_t)r * S2 + c;
        ts[o] = x;
        its[o] = xi;
        tssh[o] = h_shoup(x, qi);
        itssh[o] = h_shoup(xi, qi);
        tsd[o] = x > qi / 2 ? (double)x - (double)qi : (double)x;
        itsd[o] = xi > qi / 2 ? (double)xi - (double)qi : (double)xi;
        x = h_mulmod(x, step, qi);
        xi = h_mulmod(xi, istep, qi);
      }
    }
  }
  TwistSet n;
  n.tw = upload(ts);
  n.tw_sh = upload(tssh);
  n.itw = upload(its);
  n.itw_sh = upload(itssh);
  n.twd = upload(tsd);
  n.itwd = upload(itsd);
  t = n;
}

Context::Context(int logN_, int L1_, int K_, int dnum_, const std::vector<uint64_t>& primes)
    : logN(logN_), N(1u << logN_), L1(L1_), K(K_), dnum(dnum_) {
  if (logN < 10 || logN > 16) throw Error(-1, "logN must be in [10, 16]");
  if (L1 < 1 || K < 0 || L1 + K > kMaxPrimes) throw Error(-1, "prime counts out of range");
  if (dnum < 1 || dnum > L1) throw Error(-1, "dnum out of range");
  if ((int)primes.size() != L1 + K) throw Error(-1, "primes size must be L1 + K");
  alpha = (L1 + dnum - 1) / dnum;
  if (L1 - alpha * (dnum - 1) <= 0) throw Error(-1, "cannot distribute the Q towers into dnum digits");
  np = L1 + K;
  q = primes;
  const uint64_t m = 2ull * N;
  for (int i = 0; i < np; ++i) {
    if (q[i] >= (1ull << 62) || q[i] % m != 1 || !h_is_prime(q[i])) throw Error(-1, "modulus " + std::to_string(i) + " is not an NTT-friendly prime < 2^62");
    for (int j = 0; j < i; ++j) if (q[j] == q[i]) throw Error(-1, "duplicate modulus");
  }
  psi.resize(np);
  fp64.assign(np, 0);
  std::vector<PrimeConst> pc(np);
  std::vector<uint64_t> tw((size_t)np * N), twsh((size_t)np * N), itw((size_t)np * N), itwsh((size_t)np * N);
  std::vector<double> twd((size_t)np * N), itwd((size_t)np * N);
  for (int i = 0; i < np; ++i) {
    const uint64_t qi = q[i];
    psi[i] = h_min_root_of_unity(m, qi);
    PrimeConst& P = pc[i];
    P.q = qi;
    P.q2 = 2 * qi;
    P.k = 64 - __builtin_clzll(qi);
    P.mu = (uint64_t)(((u128)1 << (2 * P.k)) / qi);
    P.ninv = h_invmod(N % qi, qi);
    P.ninv_sh = h_shoup(P.ninv, qi);
    P.r64 = (uint64_t)(((u128)1 << 64) % qi);
    P.r64_sh = h_shoup(P.r64, qi);
    P.qd = (double)qi;
    P.qinv = 1.0 / (double)qi;
    // fp64 NTT: every intermediate stays below 2^53 when q < 2^50.2 (values are bounded by ~3.7q)
    P.fp64 = ((double)qi < 1.1487 * (double)(1ull << 50)) ? 1u : 0u;
    fp64[i] = (uint8_t)P.fp64;
    P.ninv_d = P.ninv > qi / 2 ? (double)P.ninv - (double)qi : (double)P.ninv;
    P.mont_ninv = (uint64_t)0 - h_inv_odd(qi, 0);  // modulus 2^64: mask of m-1 = all ones
    P.mont_r2 = (uint64_t)(((((u128)1 << 64) % qi) * (((u128)1 << 64) % qi)) % qi);
    // twiddles: tw[brev(j)] = psi^j (OpenFHE ChineseRemainderTransformFTTNat::PreCompute)
    const uint64_t pinv = h_invmod(psi[i], qi);
    uint64_t x = 1, xi = 1;
    for (uint32_t j = 0; j < N; ++j) {
      size_t r = (size_t)i * N + bitrev(j, logN);
      tw[r] = x;
      itw[r] = xi;
      twsh[r] = h_shoup(x, qi);
      itwsh[r] = h_shoup(xi, qi);
      // centred representatives: |w| <= q/2 halves the product magnitude in the fp64 path
      twd[r] = x > qi / 2 ? (double)x - (double)qi : (double)x;
      itwd[r] = xi > qi / 2 ? (double)xi - (double)qi : (double)xi;
      x = h_mulmod(x, psi[i], qi);
      xi = h_mulmod(xi, pinv, qi);
    }
  }
  d_prime = upload(pc);
  d_tw = upload(tw);
  d_tw_sh = upload(twsh);
  d_itw = upload(itw);
  d_itw_sh = upload(itwsh);
  d_twd = upload(twd);
  d_itwd = upload(itwd);
  ensure_twists(logN / 2);

  // ---- ModUp tables (OpenFHE PartQlHatInvModq / PartQlHatModp) ----
  const int E = L1 + K;
  std::vector<uint64_t> qhinv((size_t)L1 * dnum * alpha, 0), qhinv_sh(qhinv.size(), 0);
  std::vector<uint64_t> qhat((size_t)L1 * dnum * alpha * E, 0);
  for (int l = 0; l < L1; ++l) {
    for (int d = 0; d < num_digits(l); ++d) {
      int start = d * alpha, end = std::min(start + alpha, l + 1);
      for (int i = start; i < end; ++i) {
        uint64_t qi = q[i], prod = 1;
        for (int j = start; j < end; ++j) if (j != i) prod = h_mulmod(prod, q[j] % qi, qi);
        size_t o = ((size_t)l * dnum + d) * alpha + (i - start);
        qhinv[o] = h_invmod(prod, qi);
        qhinv_sh[o] = h_shoup(qhinv[o], qi);
        for (int pos = 0; pos < l + 1 + K; ++pos) {
          uint64_t pj = q[ext_prime(l, pos)], pr = 1;
          for (int j = start; j < end; ++j) if (j != i) pr = h_mulmod(pr, q[j] % pj, pj);
          qhat[o * E + pos] = pr;
        }
      }
    }
  }
  d_modup_qhinv = upload(qhinv);
  d_modup_qhinv_sh = upload(qhinv_sh);
  d_modup_qhat = upload(qhat);
  // Montgomery forms [Qhat_i 2^64]_{p_j}: the conversion reduces its 128-bit sum with one REDC
  std::vector<uint64_t> qhat_m(qhat.size(), 0);
  for (int l = 0; l < L1; ++l)
    for (int d = 0; d < num_digits(l); ++d)
      for (int i = 0; i < alpha; ++i)
        for (int pos = 0; pos < l + 1 + K; ++pos) {
          size_t o = (((size_t)l * dnum + d) * alpha + i) * E + pos;
          qhat_m[o] = (uint64_t)(((u128)qhat[o] << 64) % q[ext_prime(l, pos)]);
        }
  d_modup_qhat_m = upload(qhat_m);

  // ---- ModDown tables (OpenFHE PHatInvModp, PHatModq, PInvModq) ----
  std::vector<uint64_t> phinv(std::max(K, 1), 0), phinv_sh(std::max(K, 1), 0), phat((size_t)std::max(K, 1) * L1, 0);
  for (int k = 0; k < K; ++k) {
    uint64_t pk = q[L1 + k], prod = 1;
    for (int j = 0; j < K; ++j) if (j != k) prod = h_mulmod(prod, q[L1 + j] % pk, pk);
    phinv[k] = h_invmod(prod, pk);
    phinv_sh[k] = h_shoup(phinv[k], pk);
    for (int i = 0; i < L1; ++i) {
      uint64_t qi = q[i], pr = 1;
      for (int j = 0; j < K; ++j) if (j != k) pr = h_mulmod(pr, q[L1 + j] % qi, qi);
      phat[(size_t)k * L1 + i] = pr;
    }
  }
  h_pmod.assign(L1, 0);
  h_pinv.assign(L1, 0);
  std::vector<uint64_t> pinv_sh(L1), pmod_sh(L1);
  for (int i = 0; i < L1; ++i) {
    uint64_t qi = q[i], pr = 1;
    for (int j = 0; j < K; ++j) pr = h_mulmod(pr, q[L1 + j] % qi, qi);
    h_pmod[i] = pr;
    h_pinv[i] = h_invmod(pr, qi);
    pinv_sh[i] = h_shoup(h_pinv[i], qi);
    pmod_sh[i] = h_shoup(pr, qi);
  }
  d_md_phinv = upload(phinv);
  d_md_phinv_sh = upload(phinv_sh);
  d_md_phat = upload(phat);
  std::vector<uint64_t> phat_m(phat.size(), 0);
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < L1; ++i) phat_m[(size_t)k * L1 + i] = (uint64_t)(((u128)phat[(size_t)k * L1 + i] << 64) % q[i]);
  d_md_phat_m = upload(phat_m);
  // REDC needs (sum of products) < q 2^64: true when count * max prime <= 2^64
  uint64_t qmax = *std::max_element(q.begin(), q.end());
  int qbits = 64 - __builtin_clzll(qmax);
  const u128 lim = (u128)1 << 64;
  redc_conv_ok = ((u128)alpha << qbits) <= lim && ((u128)std::max(K, 1) << qbits) <= lim;
  split_mac_ok = qbits <= 60 && alpha <= 16 && K <= 16 && dnum <= 16;
  rn_conv_ok = ((u128)(alpha + 1) << qbits) <= lim && ((u128)(std::max(K, 1) + 1) << qbits) <= lim;
  {
    // ReducedNoise corrections in Montgomery form: [-Q_d 2^64]_{p_j} per (level, digit, target
    // position) and [-P 2^64]_{q_i}
    const int E = L1 + K;
    std::vector<uint64_t> negq((size_t)L1 * dnum * E, 0), negp(L1, 0);
    for (int l = 0; l < L1; ++l) {
      const int lq = l + 1, ext = lq + K;
      for (int d = 0; d < dnum; ++d) {
        const int start = d * alpha, end = std::min(start + alpha, lq);
        if (start >= end) continue;
        for (int pos = 0; pos < ext; ++pos) {
          if (pos >= start && pos < end) continue;
          const uint64_t p = q[ext_prime(l, pos)];
          uint64_t qd = 1;
          for (int j = start; j < end; ++j) qd = h_mulmod(qd, q[j] % p, p);
          const uint64_t neg = qd ? p - qd : 0;
          negq[((size_t)l * dnum + d) * E + pos] = (uint64_t)(((u128)neg << 64) % p);
        }
      }
    }
    for (int i = 0; i < L1; ++i) {
      const uint64_t neg = h_pmod[i] ? q[i] - h_pmod[i] : 0;
      negp[i] = (uint64_t)(((u128)neg << 64) % q[i]);
    }
    d_modup_negq_m = upload(negq);
    d_md_negp_m = upload(negp);
  }
  d_md_pinv = upload(h_pinv);
  d_md_pinv_sh = upload(pinv_sh);
  d_pmod = upload(h_pmod);
  d_pmod_sh = upload(pmod_sh);

  // ---- Rescale tables. OpenFHE QlQlInvModqlDivqlModq[l][i] = floor([Q_{l-1}^-1]_{q_l} Q_{l-1} / q_l) mod q_i.
  // With u = [Q_{l-1}^-1]_{q_l}, Q_{l-1} u = 1 + q_l R exactly, so R = -q_l^-1 (mod q_i) for i < l.
  std::vector<uint64_t> qlinv((size_t)L1 * L1, 0), qlinv_sh(qlinv.size(), 0), QlQl(qlinv.size(), 0), QlQl_sh(qlinv.size(), 0);
  for (int l = 1; l < L1; ++l) {
    for (int i = 0; i < l; ++i) {
      uint64_t qi = q[i];
      uint64_t v = h_invmod(q[l] % qi, qi);
      size_t o = (size_t)l * L1 + i;
      qlinv[o] = v;
      qlinv_sh[o] = h_shoup(v, qi);
      QlQl[o] = v ? qi - v : 0;
      QlQl_sh[o] = h_shoup(QlQl[o], qi);
    }
  }
  d_rs_qlinv = upload(qlinv);
  d_rs_qlinv_sh = upload(qlinv_sh);
  d_rs_QlQl = upload(QlQl);
  d_rs_QlQl_sh = upload(QlQl_sh);

  {
    const size_t words = (size_t)kMaxTeams * 19 + 64;
    FH_HIP(hipMalloc(&d_team, words * sizeof(uint32_t)));
    owned.push_back(d_team);
    FH_HIP(hipMemset(d_team, 0, words * sizeof(uint32_t)));
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    FH_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev));
  }
  FH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  ensure_aux_stream();
}

void Context::ensure_aux_stream() {
  if (aux_stream) return;
  FH_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
  FH_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  FH_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
}

Context::~Context() {
  if (aux_stream) {
    (void)hipStreamSynchronize(aux_stream);
    (void)hipStreamDestroy(aux_stream);
    (void)hipEventDestroy(ev_fork);
    (void)hipEventDestroy(ev_join);
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  pool.release_all();
  for (void* p : owned) (void)hipFree(p);
}

}  // namespace fh
