// Rescale, hybrid key switching (ModUp, key inner product, ModDown) and the
// automorphism, plus the ciphertext compositions built from them: relinearized
// EvalMult, EvalRotate in the hoisted formulation, the extended (Q_l ∪ P)
// rotation used by hoisted linear transforms.
//
// Reference: src/CKKS/LimbPartition.cu:711-797 (rescale), :851-968 (modup),
// :1345-1556 (dotKSK), :2145-2217 (moddown), :2224-2281 (automorph);
// src/CKKS/Conv.cu:27-430 (ModDown2 / DecompAndModUpConv);
// src/CKKS/ElemenwiseBatchKernels.cu:202-377 (fusedDotKSK_2_, hoistedRotateDotKSK_2_);
// src/Rotation.cuh:14-34 (automorph_slot); src/CKKS/Ciphertext.cpp:569-607,
// 618-622, 846-866, 1102-1168 (op compositions).
#include <algorithm>

#include "fh_keyswitch.hpp"

namespace fh {

namespace {

constexpr int kT = 256;
constexpr int kMaxAlpha = 32;
constexpr int kMaxK = 32;

// ----------------------------------------------------------------------------------- ModUp conversion
struct ModUpParams {
  uint32_t N, lq, ext, alpha, L1, K, beta, level, dnum;
  uint64_t in_stride;        // words between consecutive input polys (EVAL form)
  const PrimeConst* prime;
  const uint64_t* qhinv;     // [L1][dnum][alpha]
  const uint64_t* qhinv_sh;
  const uint64_t* qhat;      // [L1][dnum][alpha][L1+K]
  const uint64_t* qhat_m;    // Montgomery form of qhat
  const uint64_t* negq_m;    // ReducedNoise: [L1][dnum][L1+K] Montgomery form of [-Q_d]_{p_j}; null = off
};

// blockIdx.y = digit. in_coef: INTT of the input (lq limbs), in_eval: the input itself.
// Writes the digit's own limbs (copied in EVAL form) and the converted coefficient-form
// residues for every other extended position.
__global__ __launch_bounds__(kT) void k_modup_conv(const uint64_t* __restrict__ in_coef, const uint64_t* __restrict__ in_eval,
                                                    uint64_t* __restrict__ out, ModUpParams mp) {
  __shared__ uint64_t s_qhat[kMaxAlpha * kMaxPrimes];
  __shared__ uint64_t s_inv[kMaxAlpha], s_invsh[kMaxAlpha], s_q[kMaxAlpha];
  // blockIdx.y = poly * beta + digit; in_coef is packed (lq limbs per poly), out is beta*ext limbs per poly
  const uint32_t d = blockIdx.y % mp.beta, b = blockIdx.y / mp.beta;
  in_coef += (size_t)b * mp.lq * mp.N;
  if (in_eval) in_eval += (size_t)b * mp.in_stride;
  out += (size_t)b * mp.beta * mp.ext * mp.N;
  const uint32_t start = d * mp.alpha;
  const uint32_t sz = min(mp.alpha, mp.lq - start);
  const uint32_t E = mp.L1 + mp.K;
  const size_t tb = ((size_t)mp.level * mp.dnum + d) * mp.alpha;
  for (uint32_t i = threadIdx.x; i < sz * mp.ext; i += blockDim.x) {
    uint32_t si = i / mp.ext, pos = i % mp.ext;
    s_qhat[si * mp.ext + pos] = mp.qhat[(tb + si) * E + pos];
  }
  for (uint32_t i = threadIdx.x; i < sz; i += blockDim.x) {
    s_inv[i] = mp.qhinv[tb + i];
    s_invsh[i] = mp.qhinv_sh[tb + i];
    s_q[i] = mp.prime[start + i].q;
  }
  __syncthreads();
  uint64_t* od = out + (size_t)d * mp.ext * mp.N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < mp.N; n += gridDim.x * blockDim.x) {
    uint64_t y[kMaxAlpha];
#pragma unroll 4
    for (uint32_t i = 0; i < sz; ++i) {
      const size_t o = (size_t)(start + i) * mp.N + n;
      y[i] = shoup(in_coef[o], s_inv[i], s_invsh[i], s_q[i]);
      if (in_eval) od[o] = in_eval[o];
    }
    for (uint32_t pos = 0; pos < mp.ext; ++pos) {
      if (pos == start) { pos += sz - 1; continue; }
      const uint32_t p = pos < mp.lq ? pos : mp.L1 + (pos - mp.lq);
      U128 acc;
      acc.zero();
      for (uint32_t i = 0; i < sz; ++i) acc.mac(y[i], s_qhat[i * mp.ext + pos]);
      od[(size_t)pos * mp.N + n] = mod128(acc.hi, acc.lo, mp.prime[p]);
    }
  }
}

// Register-resident form for digit sizes up to SZ: two consecutive coefficients per lane (16-byte
// accesses), the sources prescaled once and prepared for the multiply-accumulate policy M, each
// target's 128-bit sum reduced with one Montgomery reduction against the Montgomery-form matrix
// [Qhat_i 2^64]_{p_j}. Sources beyond the digit's size are zero, so the product loop has no
// branches.
template <int SZ, class M, bool RN>
__global__ __launch_bounds__(kT) void k_modup_conv_r(const uint64_t* __restrict__ in_coef, const uint64_t* __restrict__ in_eval,
                                                      uint64_t* __restrict__ out, ModUpParams mp) {
  using Op = typename M::Op;
  __shared__ Op s_qhat[SZ * kMaxPrimes];
  __shared__ uint64_t s_pq[kMaxPrimes], s_pn[kMaxPrimes];
  const uint32_t d = blockIdx.y % mp.beta, b = blockIdx.y / mp.beta;
  in_coef += (size_t)b * mp.lq * mp.N;
  if (in_eval) in_eval += (size_t)b * mp.in_stride;
  out += (size_t)b * mp.beta * mp.ext * mp.N;
  const uint32_t start = d * mp.alpha;
  const uint32_t sz = min(mp.alpha, mp.lq - start);
  const uint32_t E = mp.L1 + mp.K;
  const size_t tb = ((size_t)mp.level * mp.dnum + d) * mp.alpha;
  for (uint32_t i = threadIdx.x; i < SZ * mp.ext; i += blockDim.x) {
    const uint32_t si = i / mp.ext, pos = i % mp.ext;
    s_qhat[si * mp.ext + pos] = si < sz ? M::prep(mp.qhat_m[(tb + si) * E + pos]) : M::zero_op();
  }
  __shared__ Op s_negq[RN ? kMaxPrimes : 1];
  for (uint32_t pos = threadIdx.x; pos < mp.ext; pos += blockDim.x) {
    const PrimeConst& P = mp.prime[pos < mp.lq ? pos : mp.L1 + (pos - mp.lq)];
    s_pq[pos] = P.q;
    s_pn[pos] = P.mont_ninv;
    if constexpr (RN) s_negq[pos] = M::prep(mp.negq_m[((size_t)mp.level * mp.dnum + d) * E + pos]);
  }
  // per-source constants live in LDS: held as uniform scalars they exhaust the SGPR file
  __shared__ uint64_t inv[SZ], invsh[SZ], qs[SZ];
  for (uint32_t i = threadIdx.x; i < sz; i += blockDim.x) {
    inv[i] = mp.qhinv[tb + i];
    invsh[i] = mp.qhinv_sh[tb + i];
    qs[i] = mp.prime[start + i].q;
  }
  __syncthreads();
  uint64_t* od = out + (size_t)d * mp.ext * mp.N;
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < mp.N; n += 2 * gridDim.x * blockDim.x) {
    Op y0[SZ], y1[SZ];
    uint32_t lift0 = 0, lift1 = 0;  // ReducedNoise: sources above q_i/2 (read as v - q_i)
#pragma unroll
    for (int i = 0; i < SZ; ++i) {
      if (i < (int)sz) {
        const size_t o = (size_t)(start + i) * mp.N + n;
        const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(in_coef + o);
        const uint64_t v0 = shoup(c.x, inv[i], invsh[i], qs[i]), v1 = shoup(c.y, inv[i], invsh[i], qs[i]);
        if constexpr (RN) {
          lift0 += v0 > (qs[i] >> 1);
          lift1 += v1 > (qs[i] >> 1);
        }
        y0[i] = M::prep(v0);
        y1[i] = M::prep(v1);
        if (in_eval) *reinterpret_cast<ulonglong2*>(od + o) = *reinterpret_cast<const ulonglong2*>(in_eval + o);
      } else {
        y0[i] = M::zero_op();
        y1[i] = M::zero_op();
      }
    }
    const Op l0 = M::prep(lift0), l1 = M::prep(lift1);
    for (uint32_t pos = 0; pos < mp.ext; ++pos) {
      if (pos == start) {
        pos += sz - 1;
        continue;
      }
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int i = 0; i < SZ; ++i) {
        const Op c = s_qhat[i * mp.ext + pos];
        a0.mac(y0[i], c);
        a1.mac(y1[i], c);
      }
      if constexpr (RN) {
        a0.mac(l0, s_negq[pos]);
        a1.mac(l1, s_negq[pos]);
      }
      const uint64_t pq = s_pq[pos], pn = s_pn[pos];
      uint64_t h0, l0, h1, l1;
      a0.get(h0, l0);
      a1.get(h1, l1);
      *reinterpret_cast<ulonglong2*>(od + (size_t)pos * mp.N + n) = make_ulonglong2(redc(h0, l0, pq, pn), redc(h1, l1, pq, pn));
    }
  }
}

// ----------------------------------------------------------------------------------- key inner product
struct KskParams {
  uint32_t N, lq, ext, L1, K, beta, logN;
  uint32_t galois;          // automorphism index (1 = none)
  const PrimeConst* prime;
  const uint64_t* key;      // [dnum][2][L1+K][N]
  const uint64_t* c0;       // optional: c0 over Q_l, added as P*c0 to out0's Q limbs
  int c0_ext;               // 1: c0 is already P-scaled over Q_l ∪ P, added on every limb
  const uint64_t* pmod;     // [L1] [P]_{q_i}
  const uint64_t* pmod_sh;
  const uint64_t* own;      // optional: the key-switched polynomial in EVAL form (lq limbs); digit d's
  uint32_t alpha;           // own limbs are read from it instead of from the digits
  int accumulate;           // 1: add the result to what out0 / out1 hold (after the automorphism)
};

// blockIdx.y = extended position. out0 = sum_d digit_d b_d (+ P c0), out1 = sum_d digit_d a_d,
// both written through the automorphism (scatter to the image slot).
template <class M>
__global__ __launch_bounds__(kT) void k_ksk_inner(const uint64_t* __restrict__ digits, uint64_t* __restrict__ out0,
                                                   uint64_t* __restrict__ out1, KskParams kp) {
  const uint32_t pos = blockIdx.y;
  const uint32_t p = pos < kp.lq ? pos : kp.L1 + (pos - kp.lq);
  const PrimeConst P = kp.prime[p];
  const uint32_t E = kp.L1 + kp.K;
  const size_t N = kp.N;
  const bool addc0 = kp.c0 != nullptr && !kp.c0_ext && pos < kp.lq;
  const bool addext = kp.c0 != nullptr && kp.c0_ext;
  const uint64_t pm = addc0 ? kp.pmod[pos] : 0, pmsh = addc0 ? kp.pmod_sh[pos] : 0;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < kp.N; n += gridDim.x * blockDim.x) {
    M a0, a1;
    a0.zero();
    a1.zero();
    const uint32_t own_d = kp.own && pos < kp.lq ? pos / kp.alpha : ~0u;
    for (uint32_t d = 0; d < kp.beta; ++d) {
      const auto x = M::prep(d == own_d ? kp.own[(size_t)pos * N + n] : digits[((size_t)d * kp.ext + pos) * N + n]);
      a0.mac(x, M::prep(kp.key[(((size_t)d * 2 + 0) * E + p) * N + n]));
      a1.mac(x, M::prep(kp.key[(((size_t)d * 2 + 1) * E + p) * N + n]));
    }
    uint64_t h0, l0, h1, l1;
    a0.get(h0, l0);
    a1.get(h1, l1);
    uint64_t r0 = redc(h0, l0, P), r1 = redc(h1, l1, P);  // keys are in Montgomery form
    if (addc0) r0 = add_mod(r0, shoup(kp.c0[(size_t)pos * N + n], pm, pmsh, P.q), P.q);
    if (addext) r0 = add_mod(r0, kp.c0[(size_t)pos * N + n], P.q);
    // gather form: out[j] = in[auto_src(j)]  <=>  scatter: out[auto_dst(n)] = in[n]
    uint32_t j = n;
    if (kp.galois != 1) {
      // invert the gather map: find j with auto_src(j) = n. auto_src with k^-1 is the inverse map.
      j = auto_src(n, kp.galois, kp.logN);
    }
    if (kp.accumulate) {
      r0 = add_mod(out0[(size_t)pos * N + j], r0, P.q);
      r1 = add_mod(out1[(size_t)pos * N + j], r1, P.q);
    }
    out0[(size_t)pos * N + j] = r0;
    out1[(size_t)pos * N + j] = r1;
  }
}

// ----------------------------------------------------------------------------------- batched key inner product
// Ciphertexts that share a rotation key form a group; a thread loads the key words of its
// (position, coefficient) once and applies them to every ciphertext of the group, so the key
// (the dominant HBM stream of a rotation) is read once per group instead of once per ciphertext.
constexpr int kMaxGroups = 64;
constexpr int kMaxBatch = 256;
constexpr int kMaxDnum = 8;

struct RotGroups {
  const uint64_t* key[kMaxGroups];
  uint32_t gal_inv[kMaxGroups];   // inverse automorphism index (scatter form), 1 = none
  uint16_t first[kMaxGroups], count[kMaxGroups];
  uint16_t order[kMaxBatch];      // batch index of the i-th ciphertext in group order
};

struct KskBatchParams {
  uint32_t N, lq, ext, L1, K, beta, logN;
  uint64_t ct_stride;       // words between the c0 sources of consecutive source ciphertexts
  uint64_t dig_stride;      // words between digit blocks of consecutive source ciphertexts
  uint64_t out_stride;      // words between output ext ciphertexts (2 * ext limbs)
  uint32_t per_src;         // items per source: item b reads the digits and c0 of source b / per_src
  int add_c0;               // fold c0: P-scaled over Q_l (c0_ext = 0) or as given over Q_l ∪ P (c0_ext = 1)
  int c0_ext;
  const PrimeConst* prime;
  const uint64_t* cts;
  const uint64_t* pmod;
  const uint64_t* pmod_sh;
  const uint64_t* own;      // optional: source b's key-switched polynomial (EVAL, lq limbs) at
  uint64_t own_stride;      // own + b * own_stride; digit d's own limbs are read from it
  uint32_t alpha;
  int accumulate;           // 1: add each result to what its output holds (after the automorphism)
};

template <class M>
__global__ __launch_bounds__(kT) void k_ksk_batch(const uint64_t* __restrict__ digits, uint64_t* __restrict__ out, KskBatchParams kp,
                                                   RotGroups g) {
  using Op = typename M::Op;
  const uint32_t pos = blockIdx.y, grp = blockIdx.z;
  const uint32_t p = pos < kp.lq ? pos : kp.L1 + (pos - kp.lq);
  const PrimeConst P = kp.prime[p];
  const uint32_t E = kp.L1 + kp.K;
  const size_t N = kp.N;
  const uint64_t* key = g.key[grp];
  const uint32_t kinv = g.gal_inv[grp];
  const bool addc0 = kp.add_c0 && !kp.c0_ext && pos < kp.lq;
  const bool addext = kp.add_c0 && kp.c0_ext;
  const uint64_t pm = addc0 ? kp.pmod[pos] : 0, pmsh = addc0 ? kp.pmod_sh[pos] : 0;
  // two consecutive coefficients per lane: 16-byte loads of keys, digits and c0
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < kp.N; n += 2 * gridDim.x * blockDim.x) {
    Op kb0[kMaxDnum], kb1[kMaxDnum], ka0[kMaxDnum], ka1[kMaxDnum];
#pragma unroll
    for (uint32_t d = 0; d < kMaxDnum; ++d) {
      if (d < kp.beta) {
        const ulonglong2 b = *reinterpret_cast<const ulonglong2*>(key + (((size_t)d * 2 + 0) * E + p) * N + n);
        const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(key + (((size_t)d * 2 + 1) * E + p) * N + n);
        kb0[d] = M::prep(b.x);
        kb1[d] = M::prep(b.y);
        ka0[d] = M::prep(a.x);
        ka1[d] = M::prep(a.y);
      }
    }
    const uint32_t j0 = kinv != 1 ? auto_src(n, kinv, kp.logN) : n;
    const uint32_t j1 = kinv != 1 ? auto_src(n + 1, kinv, kp.logN) : n + 1;
    const uint32_t cnt = g.count[grp], first = g.first[grp];
#pragma unroll 2
    for (uint32_t t = 0; t < cnt; ++t) {
      const uint32_t b = g.order[first + t];
      const uint32_t src = b / kp.per_src;
      const uint64_t* dg = digits + src * kp.dig_stride + (size_t)pos * N + n;
      const uint32_t own_d = kp.own && pos < kp.lq ? pos / kp.alpha : ~0u;
      const uint64_t* ow = kp.own ? kp.own + src * kp.own_stride + (size_t)pos * N + n : nullptr;
      M a0, a1, b0, b1;
      a0.zero();
      a1.zero();
      b0.zero();
      b1.zero();
#pragma unroll
      for (uint32_t d = 0; d < kMaxDnum; ++d) {
        if (d < kp.beta) {
          const ulonglong2 x = *reinterpret_cast<const ulonglong2*>(d == own_d ? ow : dg + (size_t)d * kp.ext * N);
          const Op x0 = M::prep(x.x), x1 = M::prep(x.y);
          a0.mac(x0, kb0[d]);
          a1.mac(x0, ka0[d]);
          b0.mac(x1, kb1[d]);
          b1.mac(x1, ka1[d]);
        }
      }
      // keys are in Montgomery form: one REDC gives sum digit * key mod p
      uint64_t h, l;
      a0.get(h, l);
      uint64_t r0 = redc(h, l, P);
      a1.get(h, l);
      uint64_t r1 = redc(h, l, P);
      b0.get(h, l);
      uint64_t s0 = redc(h, l, P);
      b1.get(h, l);
      uint64_t s1 = redc(h, l, P);
      if (addc0) {
        const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(kp.cts + src * kp.ct_stride + (size_t)pos * N + n);
        r0 = add_mod(r0, shoup(c.x, pm, pmsh, P.q), P.q);
        s0 = add_mod(s0, shoup(c.y, pm, pmsh, P.q), P.q);
      } else if (addext) {
        const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(kp.cts + src * kp.ct_stride + (size_t)pos * N + n);
        r0 = add_mod(r0, c.x, P.q);
        s0 = add_mod(s0, c.y, P.q);
      }
      uint64_t* o = out + b * kp.out_stride;
      if (kp.accumulate) {
        r0 = add_mod(o[(size_t)pos * N + j0], r0, P.q);
        r1 = add_mod(o[((size_t)kp.ext + pos) * N + j0], r1, P.q);
        s0 = add_mod(o[(size_t)pos * N + j1], s0, P.q);
        s1 = add_mod(o[((size_t)kp.ext + pos) * N + j1], s1, P.q);
      }
      o[(size_t)pos * N + j0] = r0;
      o[((size_t)kp.ext + pos) * N + j0] = r1;
      o[(size_t)pos * N + j1] = s0;
      o[((size_t)kp.ext + pos) * N + j1] = s1;
    }
  }
}

// ----------------------------------------------------------------------------------- ModDown
struct ModDownParams {
  uint32_t N, lq, ext, L1, K;
  const PrimeConst* prime;
  const uint64_t* phinv;     // [K]
  const uint64_t* phinv_sh;
  const uint64_t* phat;      // [K][L1]
  const uint64_t* phat_m;    // Montgomery form of phat
  const uint64_t* pinv;      // [L1]
  const uint64_t* pinv_sh;
  const uint64_t* negp_m;    // ReducedNoise: [L1] Montgomery form of [-P]_{q_i}; null = off
};

// blockIdx.y = poly. pcoef: INTT of the K special limbs of each poly (K limbs per poly);
// conv: lq limbs per poly, coefficient form.
__global__ __launch_bounds__(kT) void k_moddown_conv(const uint64_t* __restrict__ pcoef, uint64_t* __restrict__ conv, ModDownParams mp) {
  __shared__ uint64_t s_phat[kMaxK * kMaxPrimes];
  __shared__ uint64_t s_inv[kMaxK], s_invsh[kMaxK], s_p[kMaxK];
  for (uint32_t i = threadIdx.x; i < mp.K * mp.lq; i += blockDim.x) {
    uint32_t k = i / mp.lq, t = i % mp.lq;
    s_phat[k * mp.lq + t] = mp.phat[(size_t)k * mp.L1 + t];
  }
  for (uint32_t k = threadIdx.x; k < mp.K; k += blockDim.x) {
    s_inv[k] = mp.phinv[k];
    s_invsh[k] = mp.phinv_sh[k];
    s_p[k] = mp.prime[mp.L1 + k].q;
  }
  __syncthreads();
  const uint64_t* pc = pcoef + (size_t)blockIdx.y * mp.K * mp.N;
  uint64_t* cv = conv + (size_t)blockIdx.y * mp.lq * mp.N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < mp.N; n += gridDim.x * blockDim.x) {
    uint64_t y[kMaxK];
#pragma unroll 4
    for (uint32_t k = 0; k < mp.K; ++k) y[k] = shoup(pc[(size_t)k * mp.N + n], s_inv[k], s_invsh[k], s_p[k]);
    for (uint32_t i = 0; i < mp.lq; ++i) {
      U128 acc;
      acc.zero();
      for (uint32_t k = 0; k < mp.K; ++k) acc.mac(y[k], s_phat[k * mp.lq + i]);
      cv[(size_t)i * mp.N + n] = mod128(acc.hi, acc.lo, mp.prime[i]);
    }
  }
}

// Register-resident form for K <= KK special primes: two coefficients per lane, Montgomery-form
// matrix [P^_k 2^64]_{q_i}, one Montgomery reduction per output residue; sources beyond K are zero.
template <int KK, class M, bool RN>
__global__ __launch_bounds__(kT) void k_moddown_conv_r(const uint64_t* __restrict__ pcoef, uint64_t* __restrict__ conv, ModDownParams mp) {
  using Op = typename M::Op;
  __shared__ Op s_phat[KK * kMaxPrimes];
  __shared__ uint64_t s_q[kMaxPrimes], s_n[kMaxPrimes];
  for (uint32_t i = threadIdx.x; i < KK * mp.lq; i += blockDim.x) {
    uint32_t k = i / mp.lq, t = i % mp.lq;
    s_phat[k * mp.lq + t] = k < mp.K ? M::prep(mp.phat_m[(size_t)k * mp.L1 + t]) : M::zero_op();
  }
  __shared__ Op s_negp[RN ? kMaxPrimes : 1];
  for (uint32_t i = threadIdx.x; i < mp.lq; i += blockDim.x) {
    s_q[i] = mp.prime[i].q;
    s_n[i] = mp.prime[i].mont_ninv;
    if constexpr (RN) s_negp[i] = M::prep(mp.negp_m[i]);
  }
  __shared__ uint64_t inv[KK], invsh[KK], ps[KK];
  for (uint32_t k = threadIdx.x; k < mp.K; k += blockDim.x) {
    inv[k] = mp.phinv[k];
    invsh[k] = mp.phinv_sh[k];
    ps[k] = mp.prime[mp.L1 + k].q;
  }
  __syncthreads();
  const uint64_t* pc = pcoef + (size_t)blockIdx.y * mp.K * mp.N;
  uint64_t* cv = conv + (size_t)blockIdx.y * mp.lq * mp.N;
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < mp.N; n += 2 * gridDim.x * blockDim.x) {
    Op y0[KK], y1[KK];
    uint32_t lift0 = 0, lift1 = 0;  // ReducedNoise: sources above p_k/2 (read as v - p_k)
#pragma unroll
    for (int k = 0; k < KK; ++k) {
      if (k < (int)mp.K) {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(pc + (size_t)k * mp.N + n);
        const uint64_t v0 = shoup(v.x, inv[k], invsh[k], ps[k]), v1 = shoup(v.y, inv[k], invsh[k], ps[k]);
        if constexpr (RN) {
          lift0 += v0 > (ps[k] >> 1);
          lift1 += v1 > (ps[k] >> 1);
        }
        y0[k] = M::prep(v0);
        y1[k] = M::prep(v1);
      } else {
        y0[k] = M::zero_op();
        y1[k] = M::zero_op();
      }
    }
    const Op l0 = M::prep(lift0), l1 = M::prep(lift1);
    for (uint32_t i = 0; i < mp.lq; ++i) {
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int k = 0; k < KK; ++k) {
        const Op c = s_phat[k * mp.lq + i];
        a0.mac(y0[k], c);
        a1.mac(y1[k], c);
      }
      if constexpr (RN) {
        a0.mac(l0, s_negp[i]);
        a1.mac(l1, s_negp[i]);
      }
      const uint64_t q = s_q[i], nv = s_n[i];
      uint64_t h0, l0, h1, l1;
      a0.get(h0, l0);
      a1.get(h1, l1);
      *reinterpret_cast<ulonglong2*>(cv + (size_t)i * mp.N + n) = make_ulonglong2(redc(h0, l0, q, nv), redc(h1, l1, q, nv));
    }
  }
}

// ----------------------------------------------------------------------------------- Rescale
struct RescaleParams {
  uint32_t N, l, L1;
  const PrimeConst* prime;
  const uint64_t* qlinv;     // [L1][L1]
  const uint64_t* qlinv_sh;
  const uint64_t* QlQl;      // [L1][L1]
  const uint64_t* QlQl_sh;
};

// blockIdx.y = comp * l + i: e_i = SwitchModulus_{q_l -> q_i}(top) * QlQl[l][i] (coefficient form)
__global__ __launch_bounds__(kT) void k_rescale_expand(const uint64_t* __restrict__ top, uint64_t* __restrict__ e, RescaleParams rp) {
  const uint32_t comp = blockIdx.y / rp.l, i = blockIdx.y % rp.l;
  const uint64_t ql = rp.prime[rp.l].q, qi = rp.prime[i].q;
  const uint64_t half = ql >> 1;
  const size_t o = (size_t)rp.l * rp.L1 + i;
  const uint64_t f = rp.QlQl[o], fsh = rp.QlQl_sh[o];
  // OpenFHE NativeVector::SwitchModulus, old modulus q_l > or < q_i
  const uint64_t diff = ql > qi ? qi - (ql % qi) : qi - ql;
  const uint64_t* t = top + (size_t)comp * rp.N;
  uint64_t* ei = e + ((size_t)comp * rp.l + i) * rp.N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < rp.N; n += gridDim.x * blockDim.x) {
    uint64_t v = t[n];
    if (v > half) v += diff;
    if (v >= qi) v %= qi;
    ei[n] = shoup(v, f, fsh, qi);
  }
}

// ----------------------------------------------------------------------------------- ModRaise
// blockIdx.y = comp * (l+1) + i: out_i = SwitchModulus_{q_0 -> q_i}(x) of the coefficient-form
// tower 0 of each component (the reference's tower-0 broadcast, ElemenwiseBatchKernels.cu:140-148).
__global__ __launch_bounds__(kT) void k_mod_raise(const uint64_t* __restrict__ x0, uint64_t* __restrict__ out, uint32_t N, uint32_t l,
                                                   const PrimeConst* __restrict__ prime) {
  const uint32_t comp = blockIdx.y / (l + 1), i = blockIdx.y % (l + 1);
  const uint64_t q0 = prime[0].q, qi = prime[i].q;
  const uint64_t half = q0 >> 1;
  const uint64_t diff = q0 > qi ? qi - (q0 % qi) : qi - q0;
  const uint64_t* t = x0 + (size_t)comp * N;
  uint64_t* o = out + ((size_t)comp * (l + 1) + i) * N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    uint64_t v = t[n];
    if (i != 0) {
      if (v > half) v += diff;
      if (v >= qi) v %= qi;
    }
    o[n] = v;
  }
}

// ----------------------------------------------------------------------------------- automorphism
__global__ __launch_bounds__(kT) void k_automorph(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint32_t N, uint32_t logN,
                                                   uint32_t k) {
  const uint64_t* a = in + (size_t)blockIdx.y * N;
  uint64_t* o = out + (size_t)blockIdx.y * N;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) o[j] = a[auto_src(j, k, logN)];
}

inline dim3 grid_n(uint32_t N, uint32_t y, uint32_t per_thread = 1) {
  uint32_t gx = (N / per_thread + kT - 1) / kT;
  return dim3(std::max(1u, std::min(gx, 256u)), y);
}

}  // namespace

// ===================================================================================== host side

void modup(Context& c, const uint64_t* x, int l, uint64_t* digits, hipStream_t st) { modup_batch(c, x, l, 1, 0, digits, st); }

void modup_batch(Context& c, const uint64_t* x, int l, int npolys, uint64_t x_stride, uint64_t* digits, hipStream_t st, bool copy_own,
                 bool rows) {
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  if (c.K < 1) throw Error(-1, "key switching needs at least one special prime");
  if ((uint32_t)c.alpha > (uint32_t)kMaxAlpha) throw Error(-3, "digit size above kernel limit");
  if (npolys < 1) return;
  if (!x_stride) x_stride = (uint64_t)lq * c.N;
  std::vector<int> ids(lq);
  for (uint32_t i = 0; i < lq; ++i) ids[i] = i;
  DBuf tmp(c.pool, (size_t)npolys * lq * c.N);
  ntt_inv(c, make_set(const_cast<uint64_t*>(x), lq, ids.data(), npolys, x_stride),
          make_set(tmp.p, lq, ids.data(), npolys, (uint64_t)lq * c.N), st);
  ModUpParams mp{c.N, lq, ext, (uint32_t)c.alpha, (uint32_t)c.L1, (uint32_t)c.K, beta, (uint32_t)l, (uint32_t)c.dnum,
                 x_stride, c.d_prime, c.d_modup_qhinv, c.d_modup_qhinv_sh, c.d_modup_qhat, c.d_modup_qhat_m,
                 c.reduced_noise ? c.d_modup_negq_m : nullptr};
  if (c.reduced_noise && (!c.rn_conv_ok || c.alpha > 16)) throw Error(-3, "ReducedNoise conversions need (alpha + 1) * max prime <= 2^64");
  {
    ProfScope ps_(c, "modup_conv", st);
    const dim3 g2 = grid_n(c.N, beta * npolys, 2);
    const uint64_t* own = copy_own ? x : nullptr;
    if (!c.redc_conv_ok || c.alpha > 16) k_modup_conv<<<grid_n(c.N, beta * npolys), kT, 0, st>>>(tmp.p, own, digits, mp);
    else FH_CONV_DISPATCH(c, c.alpha, k_modup_conv_r, (g2, kT, 0, st), (tmp.p, own, digits, mp))
  }
  FH_HIP(hipGetLastError());
  ModUpAddr a{digits, ext, lq, (uint32_t)c.alpha, (uint32_t)c.L1, beta, (uint32_t)npolys, (uint64_t)beta * ext * c.N};
  if (rows) ntt_fwd_modup(c, a, a.count(), st);
  else ntt_fwd_modup_cols(c, a, a.count(), st);
}

void ksk_inner(Context& c, const uint64_t* digits, int l, const uint64_t* key, uint64_t* out0, uint64_t* out1,
               const uint64_t* c0, uint32_t galois, hipStream_t st, bool c0_ext, const uint64_t* own, bool accumulate) {
  const uint32_t lq = l + 1, ext = lq + c.K;
  KskParams kp{c.N, lq, ext, (uint32_t)c.L1, (uint32_t)c.K, (uint32_t)c.num_digits(l), (uint32_t)c.logN, galois,
               c.d_prime, key, c0, c0_ext ? 1 : 0, c.d_pmod, c.d_pmod_sh, own, (uint32_t)c.alpha, accumulate ? 1 : 0};
  if (galois != 1) {
    // the kernel scatters through the inverse map: pass k^-1 so that out[j] = in[src_k(j)]
    kp.galois = (uint32_t)h_inv_odd(galois, 2ull * c.N);
  }
  {
    ProfScope ps_(c, "ksk_inner", st);
    if (c.split_mac_ok) k_ksk_inner<Mac30><<<grid_n(c.N, ext), kT, 0, st>>>(digits, out0, out1, kp);
    else k_ksk_inner<MacWide><<<grid_n(c.N, ext), kT, 0, st>>>(digits, out0, out1, kp);
  }
  FH_HIP(hipGetLastError());
}

namespace {
// [a, a + an) and [b, b + bn) words overlap
inline bool overlaps(const uint64_t* a, size_t an, const uint64_t* b, size_t bn) { return a < b + bn && b < a + an; }
}  // namespace

void moddown(Context& c, const uint64_t* x, int l, uint64_t* out, int npolys, uint64_t x_stride, uint64_t out_stride, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K, K = c.K;
  if (K > (uint32_t)kMaxK) throw Error(-3, "special prime count above kernel limit");
  if (npolys < 1) return;
  if (!x_stride) x_stride = (uint64_t)ext * c.N;
  if (!out_stride) out_stride = (uint64_t)lq * c.N;
  const size_t xn = (size_t)(npolys - 1) * x_stride + (size_t)ext * c.N, on = (size_t)(npolys - 1) * out_stride + (size_t)lq * c.N;
  if (overlaps(x, xn, out, on)) {
    // the fused epilogue reads x while other blocks write out: go through a separate buffer
    DBuf tmp(c.pool, (size_t)npolys * lq * c.N);
    moddown(c, x, l, tmp.p, npolys, x_stride, (uint64_t)lq * c.N, st);
    for (int b = 0; b < npolys; ++b)
      FH_HIP(hipMemcpyAsync(out + b * out_stride, tmp.p + (size_t)b * lq * c.N, (size_t)lq * c.N * 8, hipMemcpyDeviceToDevice, st));
    return;
  }
  DBuf pc(c.pool, (size_t)npolys * K * c.N), conv(c.pool, (size_t)npolys * lq * c.N);
  std::vector<int> pids(K), qids(lq);
  for (uint32_t k = 0; k < K; ++k) pids[k] = c.L1 + k;
  for (uint32_t i = 0; i < lq; ++i) qids[i] = i;
  // INTT of the special limbs straight out of the extended polys into a packed buffer
  ntt_inv(c, make_set(const_cast<uint64_t*>(x) + (size_t)lq * c.N, K, pids.data(), npolys, x_stride),
          make_set(pc.p, K, pids.data(), npolys, (uint64_t)K * c.N), st);
  ModDownParams mp{c.N, lq, ext, (uint32_t)c.L1, K, c.d_prime, c.d_md_phinv, c.d_md_phinv_sh, c.d_md_phat, c.d_md_phat_m,
                   c.d_md_pinv, c.d_md_pinv_sh, c.reduced_noise ? c.d_md_negp_m : nullptr};
  if (c.reduced_noise && (!c.rn_conv_ok || K > 16)) throw Error(-3, "ReducedNoise conversions need (K + 1) * max prime <= 2^64");
  {
    ProfScope ps_(c, "moddown_conv", st);
    const dim3 g2 = grid_n(c.N, npolys, 2);
    if (!c.redc_conv_ok || K > 16) k_moddown_conv<<<grid_n(c.N, npolys), kT, 0, st>>>(pc.p, conv.p, mp);
    else FH_CONV_DISPATCH(c, (int)K, k_moddown_conv_r, (g2, kT, 0, st), (pc.p, conv.p, mp))
  }
  FH_HIP(hipGetLastError());
  // forward NTT of the conversion with (x_i - conv_i) * P^-1 fused into its last launch
  ntt_fwd_moddown(c, make_set(conv.p, lq, qids.data(), npolys, (uint64_t)lq * c.N), x, x_stride, out, out_stride, st);
}

void rescale(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, hipStream_t st) {
  if (l < 1) throw Error(-1, "cannot rescale a level-0 ciphertext");
  if (ncomp < 1) return;
  if (overlaps(x, (size_t)ncomp * (l + 1) * c.N, out, (size_t)ncomp * l * c.N)) {
    DBuf tmp(c.pool, (size_t)ncomp * l * c.N);
    rescale(c, x, l, ncomp, tmp.p, st);
    FH_HIP(hipMemcpyAsync(out, tmp.p, (size_t)ncomp * l * c.N * 8, hipMemcpyDeviceToDevice, st));
    return;
  }
  DBuf work(c.pool, rescale_work_words(c, l, ncomp));
  rescale_core(c, x, l, ncomp, out, work.p, st);
}

void rescale_core(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, uint64_t* work, hipStream_t st) {
  if (l < 1) throw Error(-1, "cannot rescale a level-0 ciphertext");
  if (ncomp < 1) return;
  if (overlaps(x, (size_t)ncomp * (l + 1) * c.N, out, (size_t)ncomp * l * c.N)) throw Error(-1, "rescale_core cannot run in place");
  uint64_t* top = work;                          // ncomp limbs
  uint64_t* e = work + (size_t)ncomp * c.N;      // ncomp x l limbs
  int pid_top = l;
  // INTT of the top limb of every component straight into a packed buffer
  ntt_inv(c, make_set(const_cast<uint64_t*>(x) + (size_t)l * c.N, 1, &pid_top, ncomp, (uint64_t)(l + 1) * c.N),
          make_set(top, 1, &pid_top, ncomp, c.N), st);
  RescaleParams rp{c.N, (uint32_t)l, (uint32_t)c.L1, c.d_prime, c.d_rs_qlinv, c.d_rs_qlinv_sh, c.d_rs_QlQl, c.d_rs_QlQl_sh};
  { ProfScope ps_(c, "rescale_expand", st); k_rescale_expand<<<grid_n(c.N, ncomp * l), kT, 0, st>>>(top, e, rp); }
  FH_HIP(hipGetLastError());
  int ids[kMaxLimbs];   // on the stack: this path makes no allocation of any kind
  for (int i = 0; i < l; ++i) ids[i] = i;
  // forward NTT of the correction with x_i * q_l^-1 + e_i fused into its last launch
  ntt_fwd_rescale(c, make_set(e, l, ids, ncomp, (uint64_t)l * c.N), x, (uint32_t)l, out, st);
}

void mod_raise(Context& c, const uint64_t* x, int l_in, int ncomp, int l, uint64_t* out, hipStream_t st) {
  if (l_in < 0 || l < 0 || l >= c.L1) throw Error(-1, "ModRaise level out of range");
  if (ncomp < 1) return;
  if (overlaps(x, (size_t)ncomp * (l_in + 1) * c.N, out, (size_t)ncomp * (l + 1) * c.N)) throw Error(-1, "ModRaise cannot run in place");
  DBuf t0(c.pool, (size_t)ncomp * c.N);
  int pid0 = 0;
  // tower 0 of every component to coefficient form, packed
  ntt_inv(c, make_set(const_cast<uint64_t*>(x), 1, &pid0, ncomp, (uint64_t)(l_in + 1) * c.N), make_set(t0.p, 1, &pid0, ncomp, c.N), st);
  { ProfScope ps_(c, "mod_raise", st); k_mod_raise<<<grid_n(c.N, ncomp * (l + 1)), kT, 0, st>>>(t0.p, out, c.N, (uint32_t)l, c.d_prime); }
  FH_HIP(hipGetLastError());
  std::vector<int> ids(l + 1);
  for (int i = 0; i <= l; ++i) ids[i] = i;
  ntt_fwd(c, make_set(out, l + 1, ids.data(), ncomp, (uint64_t)(l + 1) * c.N), st);
}

void automorph(Context& c, const uint64_t* x, int nlimbs, uint32_t k, uint64_t* out, hipStream_t st) {
  if (x == out) throw Error(-1, "automorphism cannot run in place");
  if ((k & 1) == 0 || k >= 2 * c.N) throw Error(-1, "automorphism index must be odd and < 2N");
  { ProfScope ps_(c, "automorph", st); k_automorph<<<grid_n(c.N, nlimbs), kT, 0, st>>>(x, out, c.N, c.logN, k); }
  FH_HIP(hipGetLastError());
}

// ks = ModDown(<ModUp(c2), key>); out = (c0 + ks0, c1 + ks1)
void relinearize(Context& c, const uint64_t* ct3, int l, const uint64_t* key, uint64_t* out2, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t P = (size_t)lq * c.N;
  DBuf dig(c.pool, (size_t)beta * ext * c.N), t(c.pool, (size_t)2 * ext * c.N), ks(c.pool, 2 * P);
  modup_batch(c, ct3 + 2 * P, l, 1, 0, dig.p, st, false);
  ksk_inner(c, dig.p, l, key, t.p, t.p + (size_t)ext * c.N, nullptr, 1, st, false, ct3 + 2 * P);
  moddown(c, t.p, l, ks.p, 2, 0, 0, st);
  std::vector<int> ids(lq);
  for (uint32_t i = 0; i < lq; ++i) ids[i] = i;
  ew_binary(c, 0, make_set(const_cast<uint64_t*>(ct3), lq, ids.data(), 2, P), make_set(ks.p, lq, ids.data(), 2, P),
            make_set(out2, lq, ids.data(), 2, P), st);
}

void eval_mult(Context& c, const uint64_t* ct1, const uint64_t* ct2, int l, const uint64_t* key, uint64_t* out2, hipStream_t st) {
  const size_t P = (size_t)(l + 1) * c.N;
  DBuf t(c.pool, 3 * P);
  ew_tensor(c, ct1, ct2, t.p, l + 1, st);
  relinearize(c, t.p, l, key, out2, st);
}

void eval_mult_many(Context& c, int n, const uint64_t* ct1, int limbs1, const uint64_t* ct2, int limbs2, int l, const uint64_t* key, uint64_t* out2,
                    hipStream_t st) {
  if (n < 1) return;
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t P = (size_t)lq * c.N, E = (size_t)ext * c.N;
  if (n == 1) {
    DBuf t(c.pool, 3 * P);
    ew_tensor(c, ct1, ct2, t.p, (int)lq, st, 1, limbs1, limbs2);
    return relinearize(c, t.p, l, key, out2, st);
  }
  DBuf t(c.pool, (size_t)n * 3 * P), dig(c.pool, (size_t)n * beta * E), eo(c.pool, (size_t)n * 2 * E), ks(c.pool, (size_t)n * 2 * P);
  ew_tensor(c, ct1, ct2, t.p, (int)lq, st, n, limbs1, limbs2);
  // every d2 through one ModUp, one grouped key inner product (the key read once) and one ModDown
  modup_batch(c, t.p + 2 * P, l, n, 3 * P, dig.p, st, false);
  std::vector<const uint64_t*> keys(n, key);
  std::vector<uint32_t> gal(n, 1);
  ksk_items(c, n, 1, dig.p, (uint64_t)beta * E, nullptr, 0, false, false, l, keys.data(), gal.data(), eo.p, 2 * E, st, t.p + 2 * P, 3 * P);
  moddown(c, eo.p, l, ks.p, 2 * n, E, P, st);
  std::vector<int> ids(lq);
  for (uint32_t i = 0; i < lq; ++i) ids[i] = i;
  for (int b = 0; b < n; ++b)
    ew_binary(c, 0, make_set(t.p + (size_t)b * 3 * P, lq, ids.data(), 2, P), make_set(ks.p + (size_t)b * 2 * P, lq, ids.data(), 2, P),
              make_set(out2 + (size_t)b * 2 * P, lq, ids.data(), 2, P), st);
}

// hoisted formulation: ext key switch of c1 (+ P c0), automorph, result over Q_l ∪ P
void rotate_ext(Context& c, const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, bool add_c0, uint64_t* out_ext, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t P = (size_t)lq * c.N;
  DBuf dig(c.pool, (size_t)beta * ext * c.N);
  modup_batch(c, ct + P, l, 1, 0, dig.p, st, false);
  ksk_inner(c, dig.p, l, key, out_ext, out_ext + (size_t)ext * c.N, add_c0 ? ct : nullptr, galois, st, false, ct + P);
}

void eval_rotate(Context& c, const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, uint64_t* out2, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K;
  DBuf e(c.pool, (size_t)2 * ext * c.N);
  rotate_ext(c, ct, l, key, galois, true, e.p, st);
  moddown(c, e.p, l, out2, 2, 0, 0, st);
}

// Key inner products of nb items at level l, grouped by shared (key, galois) so each key limb is
// read once per group. Item b reads the digits (and c0) of source b / per_src and writes its
// extended result (2 x (l+1+K) limbs) at out + b * out_stride, through the automorphism.
// add_c0 folds c0: P c0 over Q_l (c0_ext false) or an extended c0 added on every limb.
void ksk_items(Context& c, int nb, int per_src, const uint64_t* digits, uint64_t dig_stride, const uint64_t* c0s, uint64_t c0_stride,
               bool add_c0, bool c0_ext, int l, const uint64_t* const* keys, const uint32_t* galois, uint64_t* eo, uint64_t out_stride,
               hipStream_t st, const uint64_t* own, uint64_t own_stride, bool accumulate) {
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  if (beta > (uint32_t)kMaxDnum) throw Error(-3, "dnum above kernel limit");
  if (per_src < 1) throw Error(-1, "per_src must be >= 1");
  if (nb > 65535) throw Error(-3, "too many key-switch items in one call");
  std::vector<int> idx(nb);
  for (int i = 0; i < nb; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (keys[a] != keys[b]) return keys[a] < keys[b];
    return galois[a] < galois[b];
  });
  KskBatchParams kp{c.N, lq, ext, (uint32_t)c.L1, (uint32_t)c.K, beta, (uint32_t)c.logN, c0_stride, dig_stride, out_stride,
                    (uint32_t)per_src, add_c0 ? 1 : 0, c0_ext ? 1 : 0, c.d_prime, c0s, c.d_pmod, c.d_pmod_sh, own, own_stride,
                    (uint32_t)c.alpha, accumulate ? 1 : 0};
  int i = 0;
  while (i < nb) {
    RotGroups g{};
    int ng = 0, base = i;
    // one launch takes at most kMaxGroups groups and kMaxBatch items (a group may continue in the next)
    while (i < nb && ng < kMaxGroups && i - base < kMaxBatch) {
      const uint64_t* k = keys[idx[i]];
      uint32_t gl = galois[idx[i]];
      if (!k) throw Error(-1, "null key");
      if ((gl & 1) == 0 || gl >= 2 * c.N) throw Error(-1, "galois index must be odd and < 2N");
      g.key[ng] = k;
      g.gal_inv[ng] = gl == 1 ? 1 : (uint32_t)h_inv_odd(gl, 2ull * c.N);
      g.first[ng] = (uint16_t)(i - base);
      int cnt = 0;
      while (i < nb && i - base < kMaxBatch && keys[idx[i]] == k && galois[idx[i]] == gl) {
        g.order[i - base] = (uint16_t)idx[i];
        ++i;
        ++cnt;
      }
      g.count[ng] = (uint16_t)cnt;
      ++ng;
    }
    ProfScope ps_(c, "ksk_batch", st);
    const dim3 gk(grid_n(c.N, 1, 2).x, ext, ng);
    if (c.split_mac_ok) k_ksk_batch<Mac30><<<gk, kT, 0, st>>>(digits, eo, kp, g);
    else k_ksk_batch<MacWide><<<gk, kT, 0, st>>>(digits, eo, kp, g);
    FH_HIP(hipGetLastError());
  }
}

namespace {

// Key inner products (+ P*c0, automorphism) for B ciphertexts whose digits are already ModUp'ed,
// grouped by shared (key, galois); then optional ModDown of the 2B extended results.
void rotate_core(Context& c, int nb, const uint64_t* digits, uint64_t dig_stride, const uint64_t* xs, uint64_t ct_stride, int l,
                 const uint64_t* const* keys, const uint32_t* galois, bool add_c0, bool do_moddown, uint64_t* out, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K;
  DBuf ext_buf;
  uint64_t* eo;
  if (do_moddown) {
    ext_buf = DBuf(c.pool, (size_t)nb * 2 * ext * c.N);
    eo = ext_buf.p;
  } else {
    eo = out;
  }
  ksk_items(c, nb, 1, digits, dig_stride, xs, ct_stride, add_c0, false, l, keys, galois, eo, 2ull * ext * c.N, st, xs + (size_t)lq * c.N,
            ct_stride);
  if (do_moddown) moddown(c, eo, l, out, 2 * nb, 0, 0, st);
}

}  // namespace

// B ciphertexts at level l (ct_stride words apart, 0 = packed), each with its own (key, galois);
// ciphertexts that share a key are grouped so the key is streamed once per group. out: B packed
// results, over Q_l (2 x lq limbs each) when do_moddown, else over Q_l ∪ P (2 x ext limbs each).
void rotate_batch(Context& c, int B, const uint64_t* cts, uint64_t ct_stride, int l, const uint64_t* const* keys,
                  const uint32_t* galois, bool add_c0, bool do_moddown, uint64_t* out, hipStream_t st) {
  if (B <= 0) return;
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  if (beta > (uint32_t)kMaxDnum) throw Error(-3, "dnum above kernel limit");
  if (!ct_stride) ct_stride = 2ull * lq * c.N;
  const uint64_t out_ct = do_moddown ? 2ull * lq * c.N : 2ull * ext * c.N;
  const bool fused = rows_ksk_ok(c);
  auto part = [&](int b0, int nb, hipStream_t s) {
    const uint64_t* xs = cts + (size_t)b0 * ct_stride;
    DBuf dig(c.pool, (size_t)nb * beta * ext * c.N);
    modup_batch(c, xs + (size_t)lq * c.N, l, nb, ct_stride, dig.p, s, false, !fused);
    if (!fused) {
      rotate_core(c, nb, dig.p, (uint64_t)beta * ext * c.N, xs, ct_stride, l, keys + b0, galois + b0, add_c0, do_moddown,
                  out + (size_t)b0 * out_ct, s);
      return;
    }
    // the targets' row stages fused with the key inner product (the digits stay after the column pass)
    DBuf ext_buf;
    uint64_t* eo = out + (size_t)b0 * out_ct;
    if (do_moddown) {
      ext_buf = DBuf(c.pool, (size_t)nb * 2 * ext * c.N);
      eo = ext_buf.p;
    }
    rows_ksk(c, nb, dig.p, (uint64_t)beta * ext * c.N, xs, ct_stride, add_c0, xs + (size_t)lq * c.N, ct_stride, l, keys + b0,
             galois + b0, eo, 2ull * ext * c.N, s);
    if (do_moddown) moddown(c, eo, l, out + (size_t)b0 * out_ct, 2 * nb, 0, 0, s);
  };
  if (c.tune.rot_streams > 1 && B >= 8 && B <= kMaxBatch) {
    // two halves on two streams: the second waits for everything issued before, the caller's
    // stream waits for the second; pool blocks freed meanwhile are parked until both are issued
    c.ensure_aux_stream();
    const int h = B / 2;
    c.pool.hold();
    try {
      FH_HIP(hipEventRecord(c.ev_fork, st));
      FH_HIP(hipStreamWaitEvent(c.aux_stream, c.ev_fork, 0));
      part(0, h, st);
      part(h, B - h, c.aux_stream);
      FH_HIP(hipEventRecord(c.ev_join, c.aux_stream));
      FH_HIP(hipStreamWaitEvent(st, c.ev_join, 0));
    } catch (...) {
      (void)hipStreamSynchronize(c.aux_stream);
      (void)hipStreamSynchronize(st);
      c.pool.release_held();
      throw;
    }
    c.pool.release_held();
    return;
  }
  for (int b0 = 0; b0 < B; b0 += kMaxBatch) part(b0, std::min(B - b0, kMaxBatch), st);
}

// n rotations of ONE ciphertext from a single shared ModUp (hoisted rotation).
void rotate_hoisted(Context& c, const uint64_t* ct, int l, int n, const uint64_t* const* keys, const uint32_t* galois, bool add_c0,
                    bool do_moddown, uint64_t* out, hipStream_t st) {
  if (n <= 0) return;
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  if (beta > (uint32_t)kMaxDnum) throw Error(-3, "dnum above kernel limit");
  const uint64_t out_ct = do_moddown ? 2ull * lq * c.N : 2ull * ext * c.N;
  DBuf dig(c.pool, (size_t)beta * ext * c.N);
  modup_batch(c, ct + (size_t)lq * c.N, l, 1, 0, dig.p, st, false);
  for (int b0 = 0; b0 < n; b0 += kMaxBatch) {
    const int nb = std::min(n - b0, kMaxBatch);
    rotate_core(c, nb, dig.p, 0, ct, 0, l, keys + b0, galois + b0, add_c0, do_moddown, out + (size_t)b0 * out_ct, st);
  }
}

}  // namespace fh
