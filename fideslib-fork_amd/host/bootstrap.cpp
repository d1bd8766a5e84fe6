// CKKS bootstrapping and Chebyshev series evaluation over the facade.
//
// The bootstrap follows the reference's pipeline (src/CKKS/Bootstrap.cu:437-648):
//   ModRaise (fh_mod_raise: tower 0 lifted centred to the whole chain, Bootstrap.cu:708-930)
//   -> for sparse packing, the partial trace: the lazy rotation-fold sum at stride slots
//      (AccumulateBroadcast.cu:122-160), which keeps the coefficients at multiples of N/(2 slots)
//   -> CoeffsToSlots: the inverse special FFT as levelBudget[0] BSGS LinearTransforms, one per
//      group of butterfly stages (CoeffsToSlots.cu:75-151 applies the reference's factors the
//      same way); the bit reversal that opens the special FFT is left out of both transforms,
//      so the slots between them hold the coefficients in bit-reversed order, which EvalMod
//      (slot-wise) does not see
//   -> real and imaginary parts split with a conjugation and a multiplication by -i
//      (the monomial X^(3N/2), Bootstrap.cu:577-584)
//   -> EvalMod on both parts: a Chebyshev series of cos((2 pi K y - pi/2) / 2^r) on [-1, 1]
//      followed by r double-angle steps c <- 2c^2 - 1, which leaves sin(2 pi K y)
//      (ApproxModEval.cu:341-873); for a sparse secret the fewest double angles that bring
//      2 pi K / 2^r to at most 48, with a series of degree 63 or 127 (EvalBootstrapSetup)
//   -> recombine with i = X^(N/2) -> SlotsToCoeffs: the butterfly stages of the special FFT as
//      levelBudget[1] BSGS LinearTransforms.
// A group of m stages whose smallest butterfly half-length is h has its nonzero diagonals at the
// multiples of h within +-(2^m - 1) h: one LinearTransform with stride h, an offset, and at most
// 2^(m+1) - 1 diagonals.
// With a scaling factor close to q_0 (e.g. 59-bit scaling primes under a 60-bit q_0), the input is
// first multiplied by 2^-c and the output by the integer 2^c (the correction factor of
// Bootstrap.cu:478-497, 640-641), so that the message stays ~2^-10 of q_0 while EvalMod runs at
// the large scale.
// The scale factors are folded into the transforms: CoeffsToSlots multiplies by
// Delta / (2 gap q_0 K) so that each part enters EvalMod as (I + p / q_0) / K in [-1, 1] (the
// power of two 2^-e nearest below that factor is applied first as an exact scalar product, so that
// the diagonals keep entries near 1/slots and their encoding its full precision), and
// SlotsToCoeffs multiplies by q_0 / (2 pi Delta), undoing sin(2 pi p / q_0) ~ 2 pi p / q_0.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <set>

#include "fideslib.hpp"
#include "special_fft.hpp"

namespace fideslib {

using detail::Cx;
using detail::kPi;

struct BootstrapState {
  uint32_t slots = 0, gap = 1;
  double K = 0;          // EvalMod input range: |I| < K
  int r = 0;             // double-angle steps
  uint32_t degree = 0;   // Chebyshev degree of the series
  double pre = 1.0;      // exact power of two applied after the raise
  double scale_fix = 1.0;  // prod(q_r / scale) over the raise's rescales (in the first CoeffsToSlots group)
  uint32_t correction = 0;
  std::vector<double> cheb;
  std::vector<BootstrapGroup> cts, stc;         // transforms, in the order they are applied
  std::vector<CryptoContextImpl::PackedDiagonals> ctsDiags;  // per CoeffsToSlots group, at its level
  std::map<uint32_t, std::vector<CryptoContextImpl::PackedDiagonals>> stcDiags;  // by the towers entering SlotsToCoeffs
};

namespace {

using Diag = std::map<uint32_t, std::vector<std::complex<double>>>;  // offset -> row values

void put(Diag& D, uint32_t d, const std::vector<std::complex<double>>& v) {
  auto it = D.find(d);
  if (it == D.end()) D.emplace(d, v);
  else
    for (size_t i = 0; i < v.size(); ++i) it->second[i] += v[i];
}

// Butterfly stages of special_fft after its bit reversal (or their inverses), in diagonal form:
// row r of a stage reads column (r + offset) mod n.
std::vector<Diag> fft_stages(uint32_t n, uint64_t M, bool inverse) {
  std::vector<Diag> out;
  for (uint32_t len = 2; len <= n; len <<= 1) {
    const uint32_t lh = len >> 1;
    const uint64_t lq = (uint64_t)len << 2;
    std::vector<std::complex<double>> d0(n), dp(n), dm(n);
    for (uint32_t i = 0; i < n; i += len) {
      uint64_t g = 1;
      for (uint32_t j = 0; j < lh; ++j) {
        const uint64_t idx = (g % lq) * (M / lq);
        const double ang = 2.0 * kPi * (double)idx / (double)M;
        double wr, wi;
        ::sincos(ang, &wi, &wr);
        const std::complex<double> w(wr, wi);
        const uint32_t p = i + j, q = i + j + lh;
        if (inverse) {  // x_p = (y_p + y_q) / 2, x_q = (y_p - y_q) / (2 w)
          d0[p] = 0.5, dp[p] = 0.5;
          dm[q] = 0.5 / w, d0[q] = -0.5 / w;
        } else {        // y_p = x_p + w x_q, y_q = x_p - w x_q
          d0[p] = 1.0, dp[p] = w;
          dm[q] = 1.0, d0[q] = -w;
        }
        g = (g * 5) % M;
      }
    }
    Diag st;
    put(st, 0, d0);
    put(st, lh % n, dp);
    put(st, (n - lh) % n, dm);
    out.push_back(std::move(st));
  }
  return out;
}

// A B in diagonal form: offset d1 + d2, values A_d1[r] B_d2[r + d1]
Diag compose(const Diag& A, const Diag& B, uint32_t n) {
  Diag C;
  for (const auto& a : A)
    for (const auto& b : B) {
      std::vector<std::complex<double>> v(n);
      for (uint32_t r = 0; r < n; ++r) v[r] = a.second[r] * b.second[(r + a.first) % n];
      put(C, (a.first + b.first) % n, v);
    }
  return C;
}

// LinearTransform shape of a group whose smallest butterfly half-length is `low`
BootstrapGroup group_transform(const Diag& D, uint32_t n, uint32_t low) {
  uint32_t span = 0;  // largest |offset| / low
  for (const auto& d : D) {
    const uint32_t a = d.first <= n / 2 ? d.first : n - d.first;
    span = std::max(span, a / low);
  }
  BootstrapGroup g;
  g.stride = low;
  if (2 * span + 1 >= n / low) {
    g.rowSize = n / low;
    g.offset = 0;
  } else {
    g.rowSize = 2 * span + 1;
    g.offset = (n - span * low % n) % n;
  }
  g.diags.assign(g.rowSize, std::vector<std::complex<double>>(n));
  for (uint32_t k = 0; k < g.rowSize; ++k) {
    auto it = D.find((g.offset + k * low) % n);
    if (it != D.end()) g.diags[k] = it->second;
  }
  return g;
}

// groups of `budget` sizes (the first take the extra stages), in the order they are applied:
// CoeffsToSlots from S_L^-1 down, SlotsToCoeffs from S_1 up; the first group carries `factor`
std::vector<BootstrapGroup> factored(uint32_t n, uint64_t M, uint32_t budget, bool inverse, double factor) {
  const std::vector<Diag> S = fft_stages(n, M, inverse);
  const uint32_t L = (uint32_t)S.size();
  const uint32_t b = std::max(1u, std::min(budget, L));
  std::vector<uint32_t> order(L);
  for (uint32_t i = 0; i < L; ++i) order[i] = inverse ? L - 1 - i : i;
  std::vector<BootstrapGroup> groups;
  uint32_t pos = 0;
  for (uint32_t gi = 0; gi < b; ++gi) {
    const uint32_t m = L / b + (gi < L % b ? 1 : 0);
    Diag P;
    put(P, 0, std::vector<std::complex<double>>(n, 1.0));
    uint32_t low = n;
    for (uint32_t t = 0; t < m; ++t) {
      const uint32_t i = order[pos + t];
      P = compose(S[i], P, n);
      low = std::min(low, 1u << i);
    }
    pos += m;
    if (groups.empty())
      for (auto& d : P)
        for (auto& v : d.second) v *= factor;
    groups.push_back(group_transform(P, n, low));
  }
  return groups;
}

// baby steps of a bootstrap transform: the power of two up to 16 (the linear transform's bound)
// minimizing (b - 1) + 4 (ceil(rowSize / b) - 1). In the flat form a giant step costs a ModDown,
// a ModUp and a key inner product (batched with the other giant steps), a hoisted baby step only
// the key inner product: weighted 4 : 1. Ties go to the smaller b.
uint32_t bstep_for(uint32_t rowSize) {
  uint32_t best_b = 1, best_cost = UINT32_MAX;
  for (uint32_t b = 1; b <= 16; b *= 2) {
    const uint32_t cost = (b - 1) + 4 * ((rowSize + b - 1) / b - 1);
    if (cost < best_cost) {
      best_cost = cost;
      best_b = b;
    }
  }
  return best_b;
}

struct CosArg {
  double K;
  int r;
};
double shifted_cos(double y, const void* a) {
  const CosArg* c = static_cast<const CosArg*>(a);
  return std::cos((2.0 * kPi * c->K * y - 0.5 * kPi) / std::ldexp(1.0, c->r));
}

}  // namespace

// ---- Chebyshev series ----
std::vector<double> CryptoContextImpl::EvalChebyshevCoefficients(double (*f)(double, const void*), const void* arg, double a, double b,
                                                                 uint32_t degree) {
  const uint32_t n = degree + 1;
  std::vector<double> fx(n), c(n);
  for (uint32_t j = 0; j < n; ++j) {
    const double x = std::cos(kPi * (j + 0.5) / n);
    fx[j] = f(0.5 * (b - a) * x + 0.5 * (b + a), arg);
  }
  for (uint32_t k = 0; k < n; ++k) {
    double s = 0.0;
    for (uint32_t j = 0; j < n; ++j) s += fx[j] * std::cos(kPi * k * (j + 0.5) / n);
    c[k] = 2.0 * s / n;
  }
  return c;
}

// Two ciphertexts at one level as one of 4 components (back to back): the EvalMod of the real and
// imaginary parts runs as one sequence whose products go through fh_eval_mult_many (one launch
// sequence and one read of the relinearization key for both) and whose element-wise steps,
// rescales and weighted sums cover both parts in each launch.
struct BootstrapOps {
  static Ciphertext mult(CryptoContextImpl& cc, const Ciphertext& a0, const Ciphertext& b0) {
    if (a0->ncomp == 2 && b0->ncomp == 2) return cc.EvalMult(a0, b0);
    if (a0->ncomp != 4 || b0->ncomp != 4) throw std::runtime_error("fideslib: paired product of unpaired ciphertexts");
    if (!cc.relin_) throw std::runtime_error("fideslib: EvalMultKeyGen was not called");
    const uint32_t t = std::min(a0->towers, b0->towers);
    auto out = cc.newCt(t, 4);
    cc.check(fh_eval_mult_many(cc.ctx_, 2, a0->data->ptr, (int)a0->towers, b0->data->ptr, (int)b0->towers, (int)t - 1, cc.relin_,
                               out->data->ptr));
    out->noiseScaleDeg = a0->noiseScaleDeg + b0->noiseScaleDeg;
    out->scalingFactor = a0->scalingFactor * b0->scalingFactor;
    out->slots = a0->slots;
    return out;
  }
  static Ciphertext pack(CryptoContextImpl& cc, const Ciphertext& a, const Ciphertext& b) {
    if (a->ncomp != 2 || b->ncomp != 2 || a->towers != b->towers || a->noiseScaleDeg != b->noiseScaleDeg)
      throw std::runtime_error("fideslib: pairing needs two relinearized ciphertexts at one level");
    const size_t half = (size_t)2 * a->towers * cc.N_;
    auto p = cc.newCt(a->towers, 4);
    cc.check(fh_copy_d2d(cc.ctx_, p->data->ptr, a->data->ptr, half * 8));
    cc.check(fh_copy_d2d(cc.ctx_, p->data->ptr + half, b->data->ptr, half * 8));
    p->noiseScaleDeg = a->noiseScaleDeg;
    p->scalingFactor = a->scalingFactor;
    p->slots = a->slots;
    return p;
  }
  // ---- exact-scale forms (the bootstrap's EvalMod): scalingFactor is the ciphertext's true scale ----
  // a rescale that divides the tracked scale by the dropped prime itself
  static Ciphertext rescale_exact(CryptoContextImpl& cc, const Ciphertext& ct) {
    const double q = (double)cc.moduli_[ct->towers - 1];
    Ciphertext r = cc.modReduce(ct);
    r->scalingFactor = ct->scalingFactor / q;
    return r;
  }
  // every component times operand encoded at scale enc: degree + 1, scale times enc
  static Ciphertext mult_at(CryptoContextImpl& cc, const Ciphertext& a, double operand, double enc) {
    const auto f = cc.ElemForEvalMultAt(a->towers, operand, enc);
    auto out = cc.newCt(a->towers, a->ncomp);
    for (uint32_t c = 0; c < a->ncomp; ++c)
      cc.check(fh_eltwise_scalar(cc.ctx_, 1, a->data->ptr + (size_t)c * a->towers * cc.N_, f.data(),
                                 out->data->ptr + (size_t)c * a->towers * cc.N_, nullptr, (int)a->towers));
    out->noiseScaleDeg = a->noiseScaleDeg + 1;
    out->scalingFactor = a->scalingFactor * enc;
    out->slots = a->slots;
    return out;
  }
  // sum_k w_k cts[k] with weight k encoded at encs[k] (every term then at scale S): one fused
  // weighted sum at the level of the lowest input, degree + 1, scale S
  static Ciphertext wsum_at(CryptoContextImpl& cc, const std::vector<Ciphertext>& cts, const std::vector<double>& w,
                            const std::vector<double>& encs, double S) {
    if (cts.empty() || cts.size() > 16 || cts.size() != w.size() || w.size() != encs.size())
      throw std::runtime_error("fideslib: weighted sum of 1..16 terms");
    uint32_t t = ~0u;
    for (const auto& c : cts) t = std::min(t, c->towers);
    const size_t n = cts.size();
    std::vector<const uint64_t*> ptrs(n);
    std::vector<int32_t> towers(n);
    std::vector<uint64_t> res(n * t);
    for (size_t k = 0; k < n; ++k) {
      if (cts[k]->ncomp != cts[0]->ncomp || cts[k]->noiseScaleDeg != cts[0]->noiseScaleDeg)
        throw std::runtime_error("fideslib: weighted sum inputs differ in size or scale degree");
      ptrs[k] = cts[k]->data->ptr;
      towers[k] = (int32_t)cts[k]->towers;
      const auto f = cc.ElemForEvalMultAt(t, w[k], encs[k]);
      std::copy(f.begin(), f.end(), res.begin() + k * t);
    }
    auto out = cc.newCt(t, cts[0]->ncomp);
    cc.check(fh_linear_wsum(cc.ctx_, (int)n, ptrs.data(), towers.data(), res.data(), (int)t, (int)cts[0]->ncomp, nullptr, 0, out->data->ptr));
    out->noiseScaleDeg = cts[0]->noiseScaleDeg + 1;
    out->scalingFactor = S;
    out->slots = cts[0]->slots;
    return out;
  }
  // sum_k w_k cts[k] (integers) plus constant encoded at the inputs' common true scale
  static Ciphertext icomb_exact(CryptoContextImpl& cc, const std::vector<Ciphertext>& cts, const std::vector<int64_t>& w, double constant) {
    for (const auto& c : cts)
      if (c->scalingFactor != cts[0]->scalingFactor) throw std::runtime_error("fideslib: exact-scale combination of different scales");
    uint32_t t = ~0u;
    for (const auto& c : cts) t = std::min(t, c->towers);
    const size_t n = cts.size();
    std::vector<const uint64_t*> ptrs(n);
    std::vector<int32_t> towers(n);
    std::vector<uint64_t> res(n * t);
    for (size_t k = 0; k < n; ++k) {
      ptrs[k] = cts[k]->data->ptr;
      towers[k] = (int32_t)cts[k]->towers;
      for (uint32_t i = 0; i < t; ++i) {
        const int64_t r = w[k] % (int64_t)cc.moduli_[i];
        res[k * t + i] = (uint64_t)(r < 0 ? r + (int64_t)cc.moduli_[i] : r);
      }
    }
    std::vector<uint64_t> add;
    if (constant != 0.0) {
      add = cc.ElemForEvalAddOrSubAt(t, std::fabs(constant), 1, cts[0]->scalingFactor);
      if (constant < 0)
        for (uint32_t i = 0; i < t; ++i) add[i] = add[i] ? cc.moduli_[i] - add[i] : 0;
    }
    auto out = cc.newCt(t, cts[0]->ncomp);
    cc.check(fh_linear_wsum(cc.ctx_, (int)n, ptrs.data(), towers.data(), res.data(), (int)t, (int)cts[0]->ncomp,
                            add.empty() ? nullptr : add.data(), cts[0]->ncomp == 4 ? 5u : 1u, out->data->ptr));
    out->noiseScaleDeg = cts[0]->noiseScaleDeg;
    out->scalingFactor = cts[0]->scalingFactor;
    out->slots = cts[0]->slots;
    return out;
  }
  // rescale_exact(mult(a_i, b_i)) of several independent (paired or plain) pairs as one
  // mixed-level batch: one launch sequence for all of them, bit for bit the products and rescales
  // issued one at a time (operands may sit at different levels; the product's is the lower)
  static std::vector<Ciphertext> mult_rescale_many(CryptoContextImpl& cc, const std::vector<std::pair<Ciphertext, Ciphertext>>& ps) {
    if (!cc.relin_) throw std::runtime_error("fideslib: EvalMultKeyGen was not called");
    if (!cc.batch_) cc.check(fh_batch_create(cc.ctx_, &cc.batch_));
    cc.check(fh_batch_reset(cc.batch_));
    std::vector<Ciphertext> outs;
    for (const auto& pr : ps) {
      const Ciphertext &a = pr.first, &b = pr.second;
      if (a->ncomp != b->ncomp || (a->ncomp != 2 && a->ncomp != 4)) throw std::runtime_error("fideslib: batched product of mismatched ciphertexts");
      const uint32_t t = std::min(a->towers, b->towers);
      if (t < 2) throw std::runtime_error("fideslib: Rescale at the last tower");
      auto out = cc.newCt(t - 1, a->ncomp);
      for (uint32_t h = 0; h < a->ncomp / 2; ++h)
        cc.check(fh_batch_mult_strided(cc.batch_, a->data->ptr + (size_t)h * 2 * a->towers * cc.N_, (int)a->towers,
                                       b->data->ptr + (size_t)h * 2 * b->towers * cc.N_, (int)b->towers, (int)t - 1, cc.relin_, 1,
                                       out->data->ptr + (size_t)h * 2 * (t - 1) * cc.N_));
      const uint32_t deg = a->noiseScaleDeg + b->noiseScaleDeg;
      out->noiseScaleDeg = deg > 1 ? deg - 1 : 1;
      const double s = a->scalingFactor * b->scalingFactor;
      out->scalingFactor = s / (double)cc.moduli_[t - 1];
      out->slots = a->slots;
      outs.push_back(out);
    }
    cc.check(fh_batch_finalize(cc.batch_));
    cc.check(fh_batch_run(cc.batch_, 0));
    return outs;
  }
  static double prime(const CryptoContextImpl& cc, uint32_t i) { return (double)cc.moduli_[i]; }
  static double nominal(const CryptoContextImpl& cc) { return cc.scale_; }

  static Ciphertext half(CryptoContextImpl& cc, const Ciphertext& p, int which) {
    const size_t h = (size_t)2 * p->towers * cc.N_;
    auto x = cc.newCt(p->towers, 2);
    cc.check(fh_copy_d2d(cc.ctx_, x->data->ptr, p->data->ptr + which * h, h * 8));
    x->noiseScaleDeg = p->noiseScaleDeg;
    x->scalingFactor = p->scalingFactor;
    x->slots = p->slots;
    return x;
  }
};

namespace {
// Evaluation of sum_k c_k T_k(y) (full constant term) by recursive splitting at powers of two:
// p = q T_M + r with q_0 = c_M, q_j = 2 c_{M+j}, r_k = c_k - c_{2M-k} (T_{M+j} = 2 T_M T_j - T_{M-j});
// a piece of degree below 8, or below 16 where that costs no extra level, is a linear combination
// of T_1 .. T_d (one fused weighted sum); a constant q multiplies T_M as a scalar.
uint32_t clog2(uint32_t d) {
  uint32_t j = 0;
  while ((1u << j) < d) ++j;
  return j;
}
uint32_t cheb_depth(uint32_t d);
// levels a piece of degree d >= 1 consumes below T_1's, split at the largest power of two M <= d
uint32_t cheb_split_depth(uint32_t d) {
  uint32_t M = 1;
  while (2 * M <= d) M *= 2;
  const uint32_t lm = clog2(M);
  const uint32_t prod = (d == M ? lm : std::max(cheb_depth(d - M), lm)) + 1;
  return std::max(prod, cheb_depth(M - 1));
}
bool cheb_leaf(uint32_t d) { return d < 8 || (d < 16 && clog2(d) + 1 <= cheb_split_depth(d)); }
uint32_t cheb_depth(uint32_t d) { return cheb_leaf(d) ? clog2(d) + 1 : cheb_split_depth(d); }

// the T_k a series of degree d reads: its leaves' T_1..T_d and its split powers
void cheb_powers(uint32_t d, std::set<uint32_t>& need) {
  if (cheb_leaf(d)) {
    for (uint32_t k = 1; k <= d; ++k) need.insert(k);
    return;
  }
  uint32_t M = 1;
  while (2 * M <= d) M *= 2;
  need.insert(M);
  if (d > M) cheb_powers(d - M, need);
  cheb_powers(M - 1, need);
}
// the same with every T_k those need (T_2k from T_k, T_2k+1 from T_k and T_k+1)
std::set<uint32_t> cheb_powers_closed(uint32_t d) {
  std::set<uint32_t> need, all;
  cheb_powers(d, need);
  std::vector<uint32_t> stack(need.begin(), need.end());
  while (!stack.empty()) {
    const uint32_t k = stack.back();
    stack.pop_back();
    if (k <= 1 || !all.insert(k).second) continue;
    stack.push_back(k / 2);
    stack.push_back(k / 2 + k % 2);
  }
  return all;
}
// ciphertext products of the series' split pieces (a piece of degree d > M multiplies T_M by
// its upper part's value; exactly M a scalar product)
uint32_t cheb_tree_products(uint32_t d) {
  if (cheb_leaf(d)) return 0;
  uint32_t M = 1;
  while (2 * M <= d) M *= 2;
  return (d > M ? 1 + cheb_tree_products(d - M) : 0) + cheb_tree_products(M - 1);
}
// all ciphertext products of a series of degree d: its powers T_k (k >= 2) and its pieces'
uint32_t cheb_products(uint32_t d) { return (uint32_t)cheb_powers_closed(d).size() + cheb_tree_products(d); }

struct ChebEval {
  CryptoContextImpl& cc;
  std::map<uint32_t, Ciphertext> T;
  Ciphertext mul(const Ciphertext& a, const Ciphertext& b) { return BootstrapOps::mult(cc, a, b); }

  // T_k: T_2k = 2 T_k^2 - 1, T_2k+1 = 2 T_k T_k+1 - T_1, each with one rescale
  Ciphertext t(uint32_t k) {
    auto it = T.find(k);
    if (it != T.end()) return it->second;
    Ciphertext r;
    if (k % 2 == 0) {
      r = cc.IntegerCombination({cc.Rescale(mul(t(k / 2), t(k / 2)))}, {2}, -1.0);
    } else {
      r = cc.IntegerCombination({cc.Rescale(mul(t(k / 2), t(k / 2 + 1))), t(1)}, {2, -1}, 0.0);
    }
    T[k] = r;
    return r;
  }

  Ciphertext eval(const std::vector<double>& c) {
    const uint32_t d = (uint32_t)c.size() - 1;
    if (cheb_leaf(d)) {
      // at the level of the deepest T_k used: one fused weighted sum of the nonzero terms, one
      // rescale (the residues of the scalar products summed term by term)
      std::vector<Ciphertext> ts;
      std::vector<double> ws;
      for (uint32_t k = 1; k <= d; ++k) {
        Ciphertext tk = t(k);
        if (c[k] == 0.0) continue;
        ts.push_back(tk);
        ws.push_back(c[k]);
      }
      uint32_t towers = ~0u, used = ~0u;
      for (uint32_t k = 1; k <= d; ++k) towers = std::min(towers, t(k)->towers);
      for (const auto& x : ts) used = std::min(used, x->towers);
      if (used > towers) {
        // the sum's level is that of the deepest T_k, used or not
        ts.push_back(cc.LevelReduce(t(1), t(1)->towers - towers));
        ws.push_back(0.0);
      }
      return cc.IntegerCombination({cc.Rescale(cc.EvalLinearWSum(ts, ws))}, {1}, c[0]);
    }
    uint32_t M = 1;
    while (2 * M <= d) M *= 2;
    std::vector<double> q(d - M + 1), r(c.begin(), c.begin() + M);
    q[0] = c[M];
    for (uint32_t j = 1; j <= d - M; ++j) {
      q[j] = 2.0 * c[M + j];
      r[M - j] -= c[M + j];
    }
    Ciphertext P = q.size() == 1 ? cc.Rescale(cc.EvalMult(t(M), q[0])) : cc.Rescale(mul(eval(q), t(M)));
    Ciphertext R = eval(r);
    return cc.IntegerCombination({P, R}, {1, 1}, 0.0);
  }
};
// The same series with every ciphertext's true scale tracked (the rescales divide by the dropped
// primes, not the nominal scale): every constant is encoded at the true scale of what it is added
// to, every leaf term's weight at the scale that puts all terms of the leaf at one scale S, and
// each remainder piece at the scale of the product it is added to, so no sum combines different
// scales. With the nominal bookkeeping instead, the primes' distance from the scale (about 2^-31
// at N=2^16) enters every T_k and is doubled by each T_2k = 2 T_k^2 - 1, which bounded the
// bootstrap's precision. T_2k+1 = 2 T_k T_k+1 - T_1 takes T_1 through a scalar product that lands
// it on the product's scale.
struct ChebExact {
  CryptoContextImpl& cc;
  std::map<uint32_t, Ciphertext> T;
  std::map<uint32_t, uint32_t> lv;  // output towers of pieces, by degree (structure only)
  Ciphertext mul(const Ciphertext& a, const Ciphertext& b) { return BootstrapOps::mult(cc, a, b); }
  Ciphertext rs(const Ciphertext& a) { return BootstrapOps::rescale_exact(cc, a); }

  // T_k from the product p = T_k/2 T_k/2 or T_k/2 T_k/2+1 (already rescaled)
  Ciphertext finish(uint32_t k, const Ciphertext& p) {
    if (k % 2 == 0) return BootstrapOps::icomb_exact(cc, {p}, {2}, -1.0);
    const Ciphertext& t1 = T.at(1);
    // T_1 times 1 encoded at p's scale q / s_1: after the rescale by q, exactly p's scale
    Ciphertext a = rs(BootstrapOps::mult_at(cc, t1, 1.0, p->scalingFactor * BootstrapOps::prime(cc, t1->towers - 1) / t1->scalingFactor));
    a->scalingFactor = p->scalingFactor;
    return BootstrapOps::icomb_exact(cc, {p, a}, {2, -1}, 0.0);
  }
  Ciphertext t(uint32_t k) {
    auto it = T.find(k);
    if (it != T.end()) return it->second;
    const Ciphertext r = finish(k, rs(mul(t(k / 2), t(k / 2 + k % 2))));
    T[k] = r;
    return r;
  }
  // Every T_k the series reads, computed ahead in dependency waves: the products of one wave
  // (independent of each other) run as one batch. Values and scales are those of the lazy t(k).
  void precompute(uint32_t d) {
    const std::set<uint32_t> all = cheb_powers_closed(d);
    std::map<uint32_t, uint32_t> wave{{1, 0}};
    uint32_t waves = 0;
    for (uint32_t k : all) {  // ascending: a T_k's inputs have smaller indices
      const uint32_t w = 1 + std::max(wave.at(k / 2), wave.at(k / 2 + k % 2));
      wave[k] = w;
      waves = std::max(waves, w);
    }
    for (uint32_t w = 1; w <= waves; ++w) {
      std::vector<uint32_t> ks;
      std::vector<std::pair<Ciphertext, Ciphertext>> ops;
      for (uint32_t k : all)
        if (wave.at(k) == w && !T.count(k)) {
          ks.push_back(k);
          ops.push_back({T.at(k / 2), T.at(k / 2 + k % 2)});
        }
      if (ks.empty()) continue;
      const std::vector<Ciphertext> ps = BootstrapOps::mult_rescale_many(cc, ops);
      for (size_t i = 0; i < ks.size(); ++i) T[ks[i]] = finish(ks[i], ps[i]);
    }
  }
  // the towers the evaluation of a piece of degree d ends at
  uint32_t towers_of(uint32_t d) {
    if (cheb_leaf(d)) {
      uint32_t tw = ~0u;
      for (uint32_t k = 1; k <= d; ++k) tw = std::min(tw, t(k)->towers);
      return tw - 1;
    }
    uint32_t M = 1;
    while (2 * M <= d) M *= 2;
    const uint32_t tq = d - M == 0 ? t(M)->towers : towers_of(d - M);
    return std::min(tq, t(M)->towers) - 1;
  }
  // the piece at scale `target` (0: the nominal scale at its leaf)
  Ciphertext eval(const std::vector<double>& c, double target) {
    const uint32_t d = (uint32_t)c.size() - 1;
    if (cheb_leaf(d)) {
      uint32_t towers = ~0u;
      for (uint32_t k = 1; k <= d; ++k) towers = std::min(towers, t(k)->towers);
      const double S = (target != 0.0 ? target : BootstrapOps::nominal(cc)) * BootstrapOps::prime(cc, towers - 1);
      std::vector<Ciphertext> ts;
      std::vector<double> ws, es;
      for (uint32_t k = 1; k <= d; ++k) {
        if (c[k] == 0.0) continue;
        ts.push_back(t(k));
        ws.push_back(c[k]);
        es.push_back(S / t(k)->scalingFactor);
      }
      uint32_t used = ~0u;
      for (const auto& x : ts) used = std::min(used, x->towers);
      if (ts.empty() || used > towers) {
        // the sum's level is that of the deepest T_k, used or not
        ts.push_back(cc.LevelReduce(t(1), t(1)->towers - towers));
        ws.push_back(0.0);
        es.push_back(S / t(1)->scalingFactor);
      }
      return BootstrapOps::icomb_exact(cc, {rs(BootstrapOps::wsum_at(cc, ts, ws, es, S))}, {1}, c[0]);
    }
    uint32_t M = 1;
    while (2 * M <= d) M *= 2;
    std::vector<double> q(d - M + 1), r(c.begin(), c.begin() + M);
    q[0] = c[M];
    for (uint32_t j = 1; j <= d - M; ++j) {
      q[j] = 2.0 * c[M + j];
      r[M - j] -= c[M + j];
    }
    const Ciphertext tM = t(M);
    Ciphertext P;
    if (q.size() == 1) {
      const double enc = target != 0.0 ? target * BootstrapOps::prime(cc, tM->towers - 1) / tM->scalingFactor : BootstrapOps::nominal(cc);
      P = rs(BootstrapOps::mult_at(cc, tM, q[0], enc));
    } else {
      const uint32_t tp = std::min(towers_of(d - M), tM->towers);
      const double tq = target != 0.0 ? target * BootstrapOps::prime(cc, tp - 1) / tM->scalingFactor : 0.0;
      P = rs(mul(eval(q, tq), tM));
    }
    if (target != 0.0) P->scalingFactor = target;
    Ciphertext R = eval(r, P->scalingFactor);
    R->scalingFactor = P->scalingFactor;
    return BootstrapOps::icomb_exact(cc, {P, R}, {1, 1}, 0.0);
  }
};
}  // namespace

Ciphertext CryptoContextImpl::EvalChebyshevSeries(const Ciphertext& ct, const std::vector<double>& coeffs, double a, double b) {
  if (coeffs.size() < 2) throw std::runtime_error("fideslib: EvalChebyshevSeries needs degree >= 1");
  if (!(b > a)) throw std::runtime_error("fideslib: EvalChebyshevSeries needs a < b");
  Ciphertext y = ct->noiseScaleDeg == 2 ? modReduce(ct) : ct;
  // the series manages its rescales and tracked scales explicitly (FIXEDMANUAL semantics inside,
  // also under FLEXIBLEAUTO: the result carries the scale it tracked)
  ManualScope manual(*this);
  if (y->ncomp == 4 && (a != -1.0 || b != 1.0)) throw std::runtime_error("fideslib: EvalChebyshevSeries of a pair runs on [-1, 1] only");
  if (a != -1.0 || b != 1.0) {
    // y = (2x - (a + b)) / (b - a)
    y = Rescale(EvalMult(y, 2.0 / (b - a)));
    y = EvalAdd(y, -(a + b) / (b - a));
  }
  std::vector<double> c = coeffs;
  c[0] *= 0.5;
  ChebEval e{*this, {}};
  e.T[1] = y;
  return e.eval(c);
}

// ---- integer combinations (the chains of EvalAdd / EvalSub / MultByInteger they replace give the
// same residues: every step is exact modulo each q_i) ----
Ciphertext CryptoContextImpl::IntegerCombination(const std::vector<Ciphertext>& cts, const std::vector<int64_t>& w, double constant) {
  if (cts.empty() || cts.size() != w.size() || cts.size() > 16) throw std::runtime_error("fideslib: IntegerCombination: 1..16 terms");
  uint32_t t = ~0u;
  for (const auto& c : cts) {
    if (c->ncomp != cts[0]->ncomp || c->noiseScaleDeg != cts[0]->noiseScaleDeg)
      throw std::runtime_error("fideslib: IntegerCombination: inputs differ in size or scale degree");
    t = std::min(t, c->towers);
  }
  const size_t n = cts.size();
  std::vector<const uint64_t*> ptrs(n);
  std::vector<int32_t> towers(n);
  std::vector<uint64_t> res(n * t);
  for (size_t k = 0; k < n; ++k) {
    ptrs[k] = cts[k]->data->ptr;
    towers[k] = (int32_t)cts[k]->towers;
    for (uint32_t i = 0; i < t; ++i) {
      const int64_t r = w[k] % (int64_t)moduli_[i];
      res[k * t + i] = (uint64_t)(r < 0 ? r + (int64_t)moduli_[i] : r);
    }
  }
  std::vector<uint64_t> add;
  if (constant != 0.0) {
    add = ElemForEvalAddOrSub(t, std::fabs(constant), cts[0]->noiseScaleDeg);
    if (constant < 0)
      for (uint32_t i = 0; i < t; ++i) add[i] = add[i] ? moduli_[i] - add[i] : 0;
  }
  auto out = newCt(t, cts[0]->ncomp);
  // the constant goes to component 0 of each ciphertext (of both halves of a pair)
  check(fh_linear_wsum(ctx_, (int)n, ptrs.data(), towers.data(), res.data(), (int)t, (int)cts[0]->ncomp, add.empty() ? nullptr : add.data(),
                       cts[0]->ncomp == 4 ? 5u : 1u, out->data->ptr));
  out->noiseScaleDeg = cts[0]->noiseScaleDeg;
  out->scalingFactor = cts[0]->scalingFactor;
  out->slots = cts[0]->slots;
  return out;
}

// ---- multiplication by an integer (no scale change: multIntScalar, ApproxModEval.cu:169-175) ----
Ciphertext CryptoContextImpl::MultByInteger(const Ciphertext& ct, uint64_t v) {
  std::vector<uint64_t> f(ct->towers);
  for (uint32_t i = 0; i < ct->towers; ++i) f[i] = v % moduli_[i];
  auto out = newCt(ct->towers, ct->ncomp);
  for (uint32_t c = 0; c < ct->ncomp; ++c)
    check(fh_eltwise_scalar(ctx_, 1, ct->data->ptr + (size_t)c * ct->towers * N_, f.data(), out->data->ptr + (size_t)c * ct->towers * N_,
                            nullptr, (int)ct->towers));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

// ---- multiplication by a monomial X^power (exact: a plaintext of one +-1 coefficient) ----
Ciphertext CryptoContextImpl::MultByMonomial(const Ciphertext& ct, uint32_t power) {
  power %= 2 * N_;
  const auto key = std::make_pair(power, ct->towers);
  auto it = monomials_.find(key);
  if (it == monomials_.end()) {
    std::vector<int64_t> m(N_, 0);
    m[power % N_] = power < N_ ? 1 : -1;
    std::vector<int32_t> pids(ct->towers);
    for (uint32_t i = 0; i < ct->towers; ++i) pids[i] = (int32_t)i;
    it = monomials_.emplace(key, smallToEval(m, pids)).first;
  }
  auto out = newCt(ct->towers, ct->ncomp);
  for (uint32_t c = 0; c < ct->ncomp; ++c)
    check(fh_eltwise_binary(ctx_, 2, ct->data->ptr + (size_t)c * ct->towers * N_, it->second->ptr,
                            out->data->ptr + (size_t)c * ct->towers * N_, nullptr, (int)ct->towers));
  out->noiseScaleDeg = ct->noiseScaleDeg;
  out->scalingFactor = ct->scalingFactor;
  out->slots = ct->slots;
  return out;
}

// ---- bootstrapping ----
// a group's diagonals, pre-rotated by -(bStep floor(k / bStep) stride + offset) as LinearTransform
// expects (v'[j] = v[(j - s) mod n]), encoded extended at OpenFHE level `level`, packed once
CryptoContextImpl::PackedDiagonals CryptoContextImpl::groupDiagonals(const BootstrapGroup& g, uint32_t n, uint32_t level) {
  // all rowSize diagonals encoded in one device call straight into the packed layout
  const uint32_t b = bstep_for(g.rowSize);
  std::vector<double> vals((size_t)g.rowSize * n * 2);
  for (uint32_t k = 0; k < g.rowSize; ++k) {
    const uint32_t sft = (b * (k / b) * g.stride + g.offset) % n;
    for (uint32_t j = 0; j < n; ++j) {
      const std::complex<double> x = g.diags[k][(j + n - sft) % n];
      vals[((size_t)k * n + j) * 2] = x.real();
      vals[((size_t)k * n + j) * 2 + 1] = x.imag();
    }
  }
  PackedDiagonals p;
  p.towers = L1_ - level;
  p.rowSize = g.rowSize;
  p.noiseScaleDeg = 1;
  p.scalingFactor = scale_;
  std::vector<int32_t> pids;
  for (uint32_t i = 0; i < p.towers; ++i) pids.push_back((int32_t)i);
  for (uint32_t k = 0; k < K_; ++k) pids.push_back((int32_t)(L1_ + k));
  const size_t ext = pids.size() * N_;
  p.data = std::make_shared<DeviceBuffer>(native_, (size_t)g.rowSize * ext);
  check(fh_encode_slots(ctx_, vals.data(), (int)g.rowSize, (int)n, scale_, pids.data(), (int)pids.size(), p.data->ptr, ext));
  return p;
}

void CryptoContextImpl::EvalBootstrapSetup(const std::vector<uint32_t>& levelBudget, const std::vector<uint32_t>& dim1, uint32_t slots,
                                           uint32_t correctionFactor) {
  if (levelBudget.size() != 2 || levelBudget[0] < 1 || levelBudget[1] < 1)
    throw std::runtime_error("fideslib: EvalBootstrapSetup: the level budget is two counts >= 1");
  (void)dim1;
  if (correctionFactor > 40) throw std::runtime_error("fideslib: EvalBootstrapSetup: correction factor above 40");
  const uint32_t n = slots ? slots : N_ / 2;
  if (n < 2 || (n & (n - 1)) || n > N_ / 2) throw std::runtime_error("fideslib: bootstrap slots must be a power of two in [2, N/2]");
  auto st = std::make_shared<BootstrapState>();
  st->slots = n;
  st->gap = (N_ / 2) / n;
  // 0 = automatic: keep the message 2^-10 below q_0
  const int room = (int)params_.first_bits_ - (int)params_.scale_bits_;
  st->correction = correctionFactor ? correctionFactor : (uint32_t)std::max(0, 10 - room);
  // EvalMod range and shape: |I| <= (h + 1) / 2 for a sparse secret of weight h, and at most
  // OpenFHE's K_SPARSE = 28 (above that bound the overflow's spread, about sqrt(h / 12) per
  // coefficient, keeps |I| below 28 with overwhelming probability: OpenFHE's choice for its
  // SPARSE_TERNARY secrets of weight 192)
  const bool sparse = params_.dist_ == SecretKeyDist::SPARSE_TERNARY;
  st->K = sparse ? std::min(std::ceil(params_.hw_ / 2.0) + 4.0, 28.0) : 64.0;
  // sparse: the series approximates cos(a y - pi / 2^(r+1)) with a = 2 pi K / 2^r, which degree
  // 63 does to ~2^-45 up to a = 32, degree 127 up to a = 80 and degree 255 up to a = 160 (the
  // Chebyshev coefficients fall like the Bessel values J_k(a)). Among those (r, degree) the lowest
  // EvalMod depth, then the fewest products (the series' plus one per double angle). With the
  // series' scales exact a double angle costs well under a bit: at N=2^16, h=32 this picks r = 2
  // with degree 63, measured 22.2 bits in 25 ms against 23.4 bits in 35 ms for r = 0 with degree
  // 255 at the same depth
  st->r = 5;
  st->degree = 63;
  if (sparse) {
    static const uint32_t kDeg[3] = {63, 127, 255};
    static const double kArg[3] = {32.0, 80.0, 160.0};
    uint32_t best_depth = ~0u, best_prod = ~0u;
    for (int r = 0; r <= 10; ++r)
      for (int j = 0; j < 3; ++j) {
        if (2.0 * M_PI * st->K / std::ldexp(1.0, r) > kArg[j]) continue;
        const uint32_t depth = cheb_depth(kDeg[j]) + (uint32_t)r, prod = cheb_products(kDeg[j]) + (uint32_t)r;
        if (depth < best_depth || (depth == best_depth && prod < best_prod)) {
          best_depth = depth;
          best_prod = prod;
          st->r = r;
          st->degree = kDeg[j];
        }
      }
  }
  CosArg arg{st->K, st->r};
  st->cheb = EvalChebyshevCoefficients(shifted_cos, &arg, -1.0, 1.0, st->degree);
  const uint64_t M = 2ull * N_;
  const double q0 = (double)moduli_[0];
  const double fc = scale_ / (2.0 * st->gap * q0 * st->K), fs = q0 / (2.0 * kPi * scale_);
  const int e = (int)std::ceil(-std::log2(fc));
  st->pre = std::ldexp(1.0, -e);
  const double fd = fc * std::ldexp(1.0, e);
  // The raise's rescales (the prescale's, then one per CoeffsToSlots group) divide by the chain's
  // top primes, not by the nominal scale the constants are encoded at, so the EvalMod input would
  // arrive scaled by prod(scale / q_r). Its sine period must be exact (a relative error delta in
  // that scale moves the result by 2 pi I delta for an overflow I, about 2^-13 at N=2^16 where
  // the NTT-friendly primes sit ~2^-31 from the scale), so the first group carries the inverse.
  const uint32_t ncts = std::max(1u, std::min(levelBudget[0], (uint32_t)fft_stages(n, M, true).size()));
  double fix = 1.0;
  for (uint32_t j = 0; j <= ncts; ++j) fix *= (double)moduli_[L1_ - 1 - j] / scale_;
  st->scale_fix = fix;
  st->cts = factored(n, M, levelBudget[0], true, fd * fix);
  st->stc = factored(n, M, levelBudget[1], false, fs);
  // CoeffsToSlots group g runs at OpenFHE level 1 + g (after the prescale)
  for (size_t g = 0; g < st->cts.size(); ++g) st->ctsDiags.push_back(groupDiagonals(st->cts[g], n, (uint32_t)(1 + g)));
  boot_ = st;
}

void CryptoContextImpl::EvalBootstrapKeyGen(const PrivateKey& sk, uint32_t slots) {
  if (!boot_ || boot_->slots != (slots ? slots : N_ / 2)) throw std::runtime_error("fideslib: EvalBootstrapSetup for these slots first");
  const uint32_t n = boot_->slots;
  std::vector<int32_t> steps;
  if (n < N_ / 2)  // partial trace: the rotation-fold sum over the N/(2n) copies at stride n
    for (int32_t r : GetAccumulateSumRotationIndices((int)(N_ / 2 / n), (int)n)) steps.push_back(r);
  std::vector<BootstrapGroup> all = boot_->cts;
  all.insert(all.end(), boot_->stc.begin(), boot_->stc.end());
  for (const auto& g : all) {
    const uint32_t b = bstep_for(g.rowSize);
    for (uint32_t i = 1; i < b; ++i) steps.push_back((int32_t)((i * g.stride) % n));  // baby steps
    for (uint32_t j = 1; j * b < g.rowSize; ++j)  // giant steps: each multiple rotated by its own key
      if ((uint64_t)j * b * g.stride % n) steps.push_back((int32_t)((uint64_t)j * b * g.stride % n));
    if (g.offset) steps.push_back((int32_t)g.offset);
  }
  EvalRotateKeyGen(sk, steps);
  EvalConjugateKeyGen(sk);
}

BootstrapConstants CryptoContextImpl::GetBootstrapConstants() const {
  if (!boot_) throw std::runtime_error("fideslib: EvalBootstrapSetup first");
  return BootstrapConstants{boot_->slots, boot_->correction, boot_->K, boot_->r, boot_->pre, boot_->scale_fix, boot_->cheb, boot_->cts, boot_->stc};
}

Ciphertext CryptoContextImpl::EvalBootstrapToLevel(const Ciphertext& ct, uint32_t outputLevel, uint32_t numIterations, uint32_t precision) {
  Ciphertext out = EvalBootstrap(ct, numIterations, precision);
  if (outputLevel >= L1_) throw std::runtime_error("fideslib: EvalBootstrapToLevel: level out of range");
  const uint32_t towers = L1_ - outputLevel;
  if (towers > out->towers) throw std::runtime_error("fideslib: EvalBootstrapToLevel: the bootstrap ends below that level");
  return towers < out->towers ? LevelReduce(out, out->towers - towers) : out;
}

uint32_t CryptoContextImpl::GetBootstrapDepth() const {
  if (!boot_) throw std::runtime_error("fideslib: EvalBootstrapSetup first");
  // prescale 1 + one per CoeffsToSlots group + series + double angle r + one per SlotsToCoeffs
  // group (a series of degree 2^j - 1 >= 7 takes j + 1 levels)
  return 1 + (uint32_t)boot_->cts.size() + cheb_depth(boot_->degree) + (uint32_t)boot_->r + (uint32_t)boot_->stc.size();
}

Ciphertext CryptoContextImpl::EvalBootstrap(const Ciphertext& ct0, uint32_t numIterations, uint32_t precision) {
  if (!boot_) throw std::runtime_error("fideslib: EvalBootstrapSetup first");
  if (numIterations == 2) {
    // Meta-BTS (OpenFHE's EvalBootstrap with two iterations, as the reference composes it around
    // its device bootstrap, api/engine/cuda/CudaEngine.cpp:930-1040): bootstrap, scale the result
    // and the input up by 2^precision, bootstrap their difference (the first bootstrap's error,
    // now large enough to survive a second refresh), subtract it and scale back down.
    if (precision > 40) throw std::runtime_error("fideslib: EvalBootstrap: precision must be at most 40 bits");
    const uint64_t pw = 1ull << precision;
    Ciphertext b1 = EvalBootstrap(ct0, 1, 0);
    if (b1->noiseScaleDeg > 1) b1 = modReduce(b1);
    Ciphertext c1 = MultByInteger(b1, pw);
    if (c1->towers <= ct0->towers) return ct0;   // the bootstrap gains no towers: the input as it is
    Ciphertext up = ct0;
    if (!flex() && up->noiseScaleDeg > 1) up = modReduce(up);
    up = MultByInteger(up, pw);
    Ciphertext err = EvalSub(c1, up);
    Ciphertext b2 = EvalBootstrap(err, 1, 0);
    if (b2->noiseScaleDeg > 1) b2 = modReduce(b2);
    return EvalMult(EvalSub(c1, b2), 1.0 / (double)pw);
  }
  if (numIterations != 1) throw std::runtime_error("fideslib: EvalBootstrap: one or two iterations");
  (void)precision;
  if (ct0->ncomp != 2) throw std::runtime_error("fideslib: EvalBootstrap: relinearize first");
  BootstrapState& st = *boot_;
  const uint32_t n = st.slots;
  if (GetBootstrapDepth() + 1 > L1_) throw std::runtime_error("fideslib: EvalBootstrap: the chain is too short for the bootstrap depth");
  Ciphertext ct = ct0->noiseScaleDeg == 2 ? modReduce(ct0) : ct0;
  if (flex()) {
    // FLEXIBLEAUTO input (scale s at level factor sf): one scalar product and rescale bring the
    // message to the nominal scale 2^scaleModSize times 2^-correction, the state the steps below
    // assume (the reference's adjustment factor (target / source) (dropped prime / source)
    // 2^-correction, src/CKKS/Bootstrap.cu:680-690, with this bootstrap's nominal target)
    if (ct->towers < 2) throw std::runtime_error("fideslib: EvalBootstrap under FLEXIBLEAUTO needs two towers");
    const uint32_t l = ct->towers - 1;
    const double adj = scale_ / ct->scalingFactor * (mrf_[l] / sfReal_[l]) * std::ldexp(1.0, -(int)st.correction);
    ct = modReduce(multCore(ct, adj));
    ct->scalingFactor = scale_;
  }
  // the steps below manage their rescales and scales explicitly (FIXEDMANUAL semantics)
  ManualScope manual(*this);
  if (st.correction && GetScalingTechnique() == ScalingTechnique::FIXEDMANUAL) {
    if (ct->towers < 2) throw std::runtime_error("fideslib: EvalBootstrap with a correction factor needs two towers");
    ct = Rescale(EvalMult(ct, std::ldexp(1.0, -(int)st.correction)));
  }
  ct = LevelReduce(ct, ct->towers - 1);
  // 1. raise tower 0 to the whole chain: decrypts to p + q_0 I
  auto raised = newCt(L1_, 2);
  check(fh_mod_raise(ctx_, ct->data->ptr, 0, 2, (int)L1_ - 1, raised->data->ptr));
  raised->noiseScaleDeg = 1;
  raised->scalingFactor = ct->scalingFactor;
  raised->slots = n;
  raised = Rescale(EvalMult(raised, st.pre));
  // 2. sparse packing: trace onto the coefficients at multiples of gap (times gap)
  //    by the lazy rotation-fold sum (CKKS::Accumulate(ct, bStep, slots, N/2/slots),
  //    src/CKKS/Bootstrap.cu:548, src/CKKS/AccumulateBroadcast.cu:122-160)
  if (n < N_ / 2) raised = AccumulateSum(raised, (int)(N_ / 2 / n), (int)n);
  // 3. CoeffsToSlots: slots (I + p / q_0) / (2K), real and imaginary coefficient halves
  Ciphertext t = raised;
  for (size_t g = 0; g < st.cts.size(); ++g)
    t = Rescale(LinearTransformFlat(t, (int)bstep_for(st.cts[g].rowSize), st.ctsDiags[g], (int)st.cts[g].stride, (int)st.cts[g].offset));
  // 4. real and imaginary parts: t + conj(t), -i (t - conj(t))
  Ciphertext cj = EvalConjugate(t);
  Ciphertext re = EvalAdd(t, cj);
  Ciphertext im = MultByMonomial(EvalSub(t, cj), 3 * N_ / 2);
  // 5. EvalMod: sin(2 pi K y) on both parts, as one paired sequence with exact scales (the slots
  //    are exact at the nominal scale: the first CoeffsToSlots group carries the raise's drift)
  Ciphertext y = BootstrapOps::pack(*this, re, im);
  y->scalingFactor = scale_;
  std::vector<double> cheb = st.cheb;
  cheb[0] *= 0.5;
  ChebExact ex{*this, {}, {}};
  ex.T[1] = y;
  ex.precompute((uint32_t)cheb.size() - 1);
  Ciphertext c = ex.eval(cheb, 0.0);
  for (int i = 0; i < st.r; ++i) c = BootstrapOps::icomb_exact(*this, {BootstrapOps::rescale_exact(*this, BootstrapOps::mult(*this, c, c))}, {2}, -1.0);
  // the rest keeps the nominal bookkeeping (the output's relative scale error is the drift of the
  // EvalMod and SlotsToCoeffs levels, about 2^-27 at N=2^16)
  c->scalingFactor = scale_;
  Ciphertext fre = BootstrapOps::half(*this, c, 0), fim = BootstrapOps::half(*this, c, 1);
  // 6. recombine (i = X^(N/2)) and SlotsToCoeffs
  Ciphertext u = EvalAdd(fre, MultByMonomial(fim, N_ / 2));
  auto sit = st.stcDiags.find(u->towers);
  if (sit == st.stcDiags.end()) {
    std::vector<PackedDiagonals> pts;
    for (size_t g = 0; g < st.stc.size(); ++g) pts.push_back(groupDiagonals(st.stc[g], n, L1_ - u->towers + (uint32_t)g));
    sit = st.stcDiags.emplace(u->towers, std::move(pts)).first;
  }
  Ciphertext out = u;
  for (size_t g = 0; g < st.stc.size(); ++g)
    out = Rescale(LinearTransformFlat(out, (int)bstep_for(st.stc[g].rowSize), sit->second[g], (int)st.stc[g].stride, (int)st.stc[g].offset));
  if (st.correction) out = MultByInteger(out, 1ull << st.correction);
  out->slots = ct0->slots;
  return out;
}

}  // namespace fideslib
