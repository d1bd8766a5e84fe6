// Device modular arithmetic over 64-bit RNS words for gfx950.
//
// Every residue is a uint64 below q < 2^62. gfx950 has no 64x64->128 multiply,
// so products are assembled from 32-bit multiplies by the compiler
// (__umul64hi). Products by a fixed twiddle use Shoup's precomputed quotient;
// products of two variables use Barrett reduction of the 128-bit product.
// Reference: src/ModMult.cuh:42-95, 199-213 (Barrett/Shoup), src/AddSub.cuh:37-50.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

// Per-prime constants, one entry per prime index, resident in device memory.
struct PrimeConst {
  uint64_t q;
  uint64_t q2;        // 2q
  uint64_t mu;        // Barrett: floor(2^(2k) / q), k = bit length of q
  uint32_t k;         // bit length of q
  uint32_t fp64;      // 1 if the fp64 NTT path is exact for q (q < 2^50.2)
  uint64_t ninv;      // N^-1 mod q
  uint64_t ninv_sh;   // Shoup quotient of ninv
  uint64_t r64;       // 2^64 mod q
  uint64_t r64_sh;    // Shoup quotient of r64
  double qd;          // q as double (fp64 path)
  double qinv;        // 1/q rounded
  double ninv_d;      // N^-1 mod q, centred into (-q/2, q/2], as double
  uint64_t mont_ninv; // -q^-1 mod 2^64 (Montgomery reduction)
  uint64_t mont_r2;   // 2^128 mod q (to Montgomery form)
};

__device__ __forceinline__ uint64_t mulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// a*w mod q in [0, 2q) for any a < 2^64 and w < q, with wp = floor(w 2^64 / q).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t a, uint64_t w, uint64_t wp, uint64_t q) {
  return a * w - mulhi(a, wp) * q;
}

__device__ __forceinline__ uint64_t reduce2(uint64_t a, uint64_t q) { return a >= q ? a - q : a; }

__device__ __forceinline__ uint64_t shoup(uint64_t a, uint64_t w, uint64_t wp, uint64_t q) {
  return reduce2(shoup_lazy(a, w, wp, q), q);
}

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) {
  uint64_t s = a + b;
  return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) {
  return a >= b ? a - b : a + q - b;
}

// Barrett reduction of the 128-bit value (hi, lo) < q^2, result in [0, q).
// t = x >> (k-1) < 2^(k+1); qhat = (t * mu) >> (k+1); x - qhat q < 3q.
__device__ __forceinline__ uint64_t barrett128(uint64_t hi, uint64_t lo, const PrimeConst& p) {
  const uint32_t k = p.k;
  uint64_t t = (hi << (65 - k)) | (lo >> (k - 1));
  uint64_t ph = mulhi(t, p.mu), pl = t * p.mu;
  uint64_t qh = (ph << (63 - k)) | (pl >> (k + 1));
  uint64_t r = lo - qh * p.q;
  r = reduce2(r, p.q2);
  return reduce2(r, p.q);
}

__device__ __forceinline__ uint64_t mul_mod(uint64_t a, uint64_t b, const PrimeConst& p) {
  return barrett128(mulhi(a, b), a * b, p);
}

// 128-bit accumulator for sums of products (basis conversion, inner products).
struct U128 {
  uint64_t lo, hi;
  __device__ __forceinline__ void zero() { lo = 0; hi = 0; }
  __device__ __forceinline__ void mac(uint64_t a, uint64_t b) {
    uint64_t pl = a * b, ph = mulhi(a, b);
    uint64_t nl = lo + pl;
    hi += ph + (nl < lo ? 1 : 0);
    lo = nl;
  }
};

// Exact sum of up to 16 products y c with y, c < 2^60 from 30-bit halves (y = y1 2^30 + y0):
// every partial product y_a c_b is below 2^60, so four plain 64-bit sums of at most 16 terms
// cannot overflow, and each term costs four 32x32->64 multiply-adds with no carry handling.
// The 128-bit value s00 + (s01 + s10) 2^30 + s11 2^60 is assembled once.
struct Acc30 {
  uint64_t s00, s01, s10, s11;
  __device__ __forceinline__ void zero() { s00 = s01 = s10 = s11 = 0; }
  __device__ __forceinline__ void mac(uint32_t y0, uint32_t y1, uint32_t c0, uint32_t c1) {
    s00 += (uint64_t)y0 * c0;
    s01 += (uint64_t)y0 * c1;
    s10 += (uint64_t)y1 * c0;
    s11 += (uint64_t)y1 * c1;
  }
  __device__ __forceinline__ void get(uint64_t& hi, uint64_t& lo) const {
    const uint64_t m = s01 + s10;
    const uint64_t mc = m < s01 ? 1 : 0;  // carry of the middle sum, weight 2^94
    uint64_t h = (m >> 34) + (mc << 30) + (s11 >> 4);
    const uint64_t l1 = m << 30, l2 = s11 << 60;
    const uint64_t l = s00 + l1;
    h += l < l1 ? 1 : 0;
    const uint64_t l3 = l + l2;
    h += l3 < l2 ? 1 : 0;
    lo = l3;
    hi = h;
  }
};

// Multiply-accumulate policies of the basis conversions and key inner products. Operands are
// prepared once (prep) and reused across every product they enter.
struct MacWide {   // any residues below 2^64: 128-bit product per term
  using Op = uint64_t;
  U128 a;
  __device__ __forceinline__ static Op prep(uint64_t v) { return v; }
  __device__ __forceinline__ static Op zero_op() { return 0; }
  __device__ __forceinline__ void zero() { a.zero(); }
  __device__ __forceinline__ void mac(Op y, Op c) { a.mac(y, c); }
  __device__ __forceinline__ void get(uint64_t& hi, uint64_t& lo) const { hi = a.hi; lo = a.lo; }
};
struct Mac30 {     // residues below 2^60, at most 16 terms per sum
  struct Op {
    uint32_t lo, hi;
  };
  Acc30 a;
  __device__ __forceinline__ static Op prep(uint64_t v) { return Op{(uint32_t)v & 0x3FFFFFFFu, (uint32_t)(v >> 30)}; }
  __device__ __forceinline__ static Op zero_op() { return Op{0, 0}; }
  __device__ __forceinline__ void zero() { a.zero(); }
  __device__ __forceinline__ void mac(Op y, Op c) { a.mac(y.lo, y.hi, c.lo, c.hi); }
  __device__ __forceinline__ void get(uint64_t& hi, uint64_t& lo) const { a.get(hi, lo); }
};

// Montgomery reduction: for T = hi 2^64 + lo < q 2^64, returns T 2^-64 mod q in [0, q).
// m = lo (-q^-1) mod 2^64 makes T + m q divisible by 2^64; the low word of the sum is zero,
// so its carry is 1 exactly when lo != 0.
__device__ __forceinline__ uint64_t redc(uint64_t hi, uint64_t lo, uint64_t q, uint64_t ninv) {
  const uint64_t m = lo * ninv;
  uint64_t t = hi + mulhi(m, q) + (lo != 0 ? 1 : 0);
  return t >= q ? t - q : t;
}
__device__ __forceinline__ uint64_t redc(uint64_t hi, uint64_t lo, const PrimeConst& p) { return redc(hi, lo, p.q, p.mont_ninv); }

// a * b * 2^-64 mod q for a, b < q (Montgomery product)
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const PrimeConst& p) {
  return redc(mulhi(a, b), a * b, p);
}

// (hi, lo) mod q for any 128-bit value: x = hi [2^64]_q + lo (mod q).
__device__ __forceinline__ uint64_t mod128(uint64_t hi, uint64_t lo, const PrimeConst& p) {
  uint64_t h = shoup(hi, p.r64, p.r64_sh, p.q);  // [0, q)
  uint64_t l = barrett128(0, lo, p);              // lo < 2^64 < q^2
  return add_mod(h, l, p.q);
}

}  // namespace fh
