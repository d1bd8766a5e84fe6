// Hoisted BSGS linear transform (plaintext-diagonal matrix-vector product) in the
// extended-basis formulation of the reference: baby-step rotations share one ModUp and
// stay in Q_l ∪ P, one multiply-accumulate kernel forms every giant-step partial sum, and
// the giant steps are folded by Horner's rule with only c1 brought back to Q_l before each
// giant rotation (the extended c0 enters the key-switch product directly).
// Reference: src/CKKS/LinearTransform.cu:65-254 (FUSED branch, ONLY_C1), MAC kernel
// dotProductLtBatchedPt2___ (src/CKKS/ElemenwiseBatchKernels.cu:785-835) via
// LTdotProductPtBatch (src/CKKS/LimbPartitionBatch.cu:266-524), Ciphertext::extend
// (src/CKKS/Ciphertext.cpp:1087-1100), contract api/CryptoContext.hpp:460-517.
#include <algorithm>

#include "fh_keyswitch.hpp"

namespace fh {

namespace {

constexpr int kT = 256;
constexpr int kMaxBaby = 16;

// out[comp][pos] = [P]_{q_pos} x[comp][pos] for Q limbs, 0 for the special limbs (the P-scaled
// extended form of a ciphertext over Q_l).
__global__ __launch_bounds__(kT) void k_extend_p(const uint64_t* __restrict__ x, uint64_t* __restrict__ out, uint32_t N, uint32_t lq,
                                                  uint32_t ext, const PrimeConst* prime, const uint64_t* pmod, const uint64_t* pmod_sh) {
  const uint32_t comp = blockIdx.y / ext, pos = blockIdx.y % ext;
  uint64_t* o = out + ((size_t)comp * ext + pos) * N;
  if (pos >= lq) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) o[n] = 0;
    return;
  }
  const uint64_t q = prime[pos].q, w = pmod[pos], wsh = pmod_sh[pos];
  const uint64_t* xi = x + ((size_t)comp * lq + pos) * N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) o[n] = shoup(xi[n], w, wsh, q);
}

struct MacParams {
  uint32_t N, lq, ext, L1, rowSize, bStep, gStep;
  const PrimeConst* prime;
  uint32_t i0 = 0;   // first baby step of this launch (a launch covers baby steps i0 .. i0 + BB - 1)
  uint32_t acc = 0;  // 1: add into R (a later launch over the next baby steps)
};

// R[j][comp][pos] = sum_{i: bStep j + i < rowSize} fast[i][comp][pos] (.) diag[bStep j + i][pos]
// One lane holds the bStep baby-step words of both components for its (position, coefficient)
// and streams the diagonals once; each sum (at most 16 products) is accumulated exactly under the
// multiply policy M and reduced once.
template <int BB, class M>
__global__ __launch_bounds__(kT) void k_lt_mac(const uint64_t* __restrict__ fast, const uint64_t* __restrict__ diags,
                                                uint64_t* __restrict__ R, MacParams mp) {
  using Op = typename M::Op;
  const uint32_t pos = blockIdx.y;
  const PrimeConst P = mp.prime[pos < mp.lq ? pos : mp.L1 + (pos - mp.lq)];
  const size_t N = mp.N, E = (size_t)mp.ext * N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < mp.N; n += gridDim.x * blockDim.x) {
    Op f0[BB], f1[BB];
#pragma unroll
    for (int i = 0; i < BB; ++i) {
      if (mp.i0 + i < mp.bStep) {
        f0[i] = M::prep(fast[((size_t)(mp.i0 + i) * 2 + 0) * E + (size_t)pos * N + n]);
        f1[i] = M::prep(fast[((size_t)(mp.i0 + i) * 2 + 1) * E + (size_t)pos * N + n]);
      }
    }
    for (uint32_t j = 0; j < mp.gStep; ++j) {
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int i = 0; i < BB; ++i) {
        const uint32_t k = j * mp.bStep + mp.i0 + i;
        if (mp.i0 + i < mp.bStep && k < mp.rowSize) {
          const Op d = M::prep(diags[(size_t)k * E + (size_t)pos * N + n]);
          a0.mac(f0[i], d);
          a1.mac(f1[i], d);
        }
      }
      uint64_t hi, lo;
      uint64_t* r0 = R + ((size_t)j * 2 + 0) * E + (size_t)pos * N + n;
      uint64_t* r1 = R + ((size_t)j * 2 + 1) * E + (size_t)pos * N + n;
      a0.get(hi, lo);
      const uint64_t v0 = mod128(hi, lo, P);
      a1.get(hi, lo);
      const uint64_t v1 = mod128(hi, lo, P);
      *r0 = mp.acc ? add_mod(*r0, v0, P.q) : v0;
      *r1 = mp.acc ? add_mod(*r1, v1, P.q) : v1;
    }
  }
}

inline dim3 grid_n(uint32_t N, uint32_t y) { return dim3(std::max(1u, std::min((N + kT - 1) / kT, 256u)), y); }

// acc += sum_{t < nt} T[t] over the extended basis of both components (acc: the LimbSet's two
// polynomials; T[t] at T + t * t_stride in the same layout): one read of each summand and one
// read-modify-write of acc instead of nt separate additions. blockIdx.y = (comp, pos).
__global__ __launch_bounds__(kT) void k_lt_sum(LimbSet acc, const uint64_t* __restrict__ T, uint64_t t_stride, int nt, uint32_t N,
                                              const PrimeConst* prime) {
  const uint32_t comp = blockIdx.y / acc.nlimbs, pos = blockIdx.y % acc.nlimbs;
  const uint64_t q = prime[acc.pid[pos]].q;
  const size_t off = (size_t)comp * acc.poly_stride + (size_t)pos * N;
  uint64_t* a = acc.base + off;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    uint64_t s = a[n];
    for (int t = 0; t < nt; ++t) s = add_mod(s, T[t * t_stride + off + n], q);
    a[n] = s;
  }
}

// k_extend_p for B ciphertexts: blockIdx.y = (b, comp, pos)
__global__ __launch_bounds__(kT) void k_extend_p_b(const uint64_t* __restrict__ x, uint64_t x_stride, uint64_t* __restrict__ out,
                                                    uint64_t out_stride, uint32_t N, uint32_t lq, uint32_t ext, const PrimeConst* prime,
                                                    const uint64_t* pmod, const uint64_t* pmod_sh) {
  const uint32_t b = blockIdx.y / (2 * ext), r = blockIdx.y % (2 * ext);
  const uint32_t comp = r / ext, pos = r % ext;
  uint64_t* o = out + b * out_stride + ((size_t)comp * ext + pos) * N;
  if (pos >= lq) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) o[n] = 0;
    return;
  }
  const uint64_t q = prime[pos].q, w = pmod[pos], wsh = pmod_sh[pos];
  const uint64_t* xi = x + b * x_stride + ((size_t)comp * lq + pos) * N;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) o[n] = shoup(xi[n], w, wsh, q);
}

constexpr int kMacTile = 64;   // coefficients per workgroup of the batched MAC
constexpr int kMacLanes = kT / kMacTile;

struct MacBatchParams {
  uint32_t N, lq, ext, L1, rowSize, bStep, gStep, B, nrot;
  const PrimeConst* prime;
  int8_t slot[kMaxBaby];   // baby step i: rotation slot (>= 0) or the P-scaled identity (-1)
};

// Giant-step partial sums for B ciphertexts sharing one diagonal set. blockIdx.y = extended
// position, blockIdx.x = a 64-coefficient tile: the tile of every diagonal is staged in LDS once
// (as multiply operands, split into 30-bit halves under Mac30) and applied to all B ciphertexts
// (four ciphertexts in flight per workgroup), so the diagonals are read once per batch.
// R[b][j][comp][pos] = sum_i fast_b[i][comp][pos] (.) diag[bStep j + i][pos], each sum of at most
// 16 products accumulated exactly and reduced once.
template <int BB, class M>
__global__ __launch_bounds__(kT) void k_lt_mac_b(const uint64_t* __restrict__ rot, const uint64_t* __restrict__ ident,
                                                  const uint64_t* __restrict__ diags, uint64_t* __restrict__ R, MacBatchParams mp) {
  using Op = typename M::Op;
  static_assert(sizeof(Op) == 8, "one LDS word per diagonal coefficient");
  extern __shared__ uint64_t sd_raw[];   // rowSize x kMacTile
  Op* sd = reinterpret_cast<Op*>(sd_raw);
  const uint32_t pos = blockIdx.y;
  const PrimeConst P = mp.prime[pos < mp.lq ? pos : mp.L1 + (pos - mp.lq)];
  const size_t N = mp.N, E = (size_t)mp.ext * N;
  const uint32_t lane = threadIdx.x % kMacTile, cl = threadIdx.x / kMacTile;
  const uint32_t n = blockIdx.x * kMacTile + lane;
  for (uint32_t k = cl; k < mp.rowSize; k += kMacLanes) sd[k * kMacTile + lane] = M::prep(diags[k * E + (size_t)pos * N + n]);
  __syncthreads();
  for (uint32_t b = cl; b < mp.B; b += kMacLanes) {
    Op f0[BB], f1[BB];
#pragma unroll
    for (int i = 0; i < BB; ++i) {
      if (i < (int)mp.bStep) {
        const uint64_t* src = mp.slot[i] >= 0 ? rot + ((size_t)b * mp.nrot + mp.slot[i]) * 2 * E : ident + (size_t)b * 2 * E;
        f0[i] = M::prep(src[(size_t)pos * N + n]);
        f1[i] = M::prep(src[E + (size_t)pos * N + n]);
      }
    }
    uint64_t* Rb = R + (size_t)b * mp.gStep * 2 * E;
    for (uint32_t j = 0; j < mp.gStep; ++j) {
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int i = 0; i < BB; ++i) {
        const uint32_t k = j * mp.bStep + i;
        if (i < (int)mp.bStep && k < mp.rowSize) {
          const Op d = sd[k * kMacTile + lane];
          a0.mac(f0[i], d);
          a1.mac(f1[i], d);
        }
      }
      uint64_t hi, lo;
      a0.get(hi, lo);
      Rb[((size_t)j * 2 + 0) * E + (size_t)pos * N + n] = mod128(hi, lo, P);
      a1.get(hi, lo);
      Rb[((size_t)j * 2 + 1) * E + (size_t)pos * N + n] = mod128(hi, lo, P);
    }
  }
}

}  // namespace

namespace {

// Baby steps (the identity in P-scaled extended form, the others from one shared ModUp) and the
// multiply-accumulate of every giant-step partial sum: R_j, 2 x (l+1+K) limbs each, j < gStep.
DBuf lt_partials(Context& c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int slots, const uint64_t* diags,
                 const uint64_t* const* baby_keys, const uint32_t* baby_gal, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K;
  const size_t E = (size_t)ext * c.N;
  const int gStep = (rowSize + bStep - 1) / bStep;
  DBuf fast(c.pool, (size_t)bStep * 2 * E);
  std::vector<int> rot_idx;
  std::vector<const uint64_t*> rk;
  std::vector<uint32_t> rg;
  for (int i = 0; i < bStep; ++i) {
    if (((long long)i * stride) % slots == 0) {
      ProfScope ps_(c, "lt_extend", st);
      k_extend_p<<<grid_n(c.N, 2 * ext), kT, 0, st>>>(ct, fast.p + (size_t)i * 2 * E, c.N, lq, ext, c.d_prime, c.d_pmod, c.d_pmod_sh);
      FH_HIP(hipGetLastError());
    } else {
      if (!baby_keys[i]) throw Error(-1, "baby-step key missing");
      rot_idx.push_back(i);
      rk.push_back(baby_keys[i]);
      rg.push_back(baby_gal[i]);
    }
  }
  if (!rot_idx.empty()) {
    bool contiguous = true;
    for (size_t t = 1; t < rot_idx.size(); ++t) contiguous &= rot_idx[t] == rot_idx[t - 1] + 1;
    if (contiguous) {
      rotate_hoisted(c, ct, l, (int)rot_idx.size(), rk.data(), rg.data(), true, false, fast.p + (size_t)rot_idx[0] * 2 * E, st);
    } else {
      DBuf tmp(c.pool, rot_idx.size() * 2 * E);
      rotate_hoisted(c, ct, l, (int)rot_idx.size(), rk.data(), rg.data(), true, false, tmp.p, st);
      for (size_t t = 0; t < rot_idx.size(); ++t)
        FH_HIP(hipMemcpyAsync(fast.p + (size_t)rot_idx[t] * 2 * E, tmp.p + t * 2 * E, 2 * E * 8, hipMemcpyDeviceToDevice, st));
    }
  }
  DBuf R(c.pool, (size_t)gStep * 2 * E);
  MacParams mp{c.N, lq, ext, (uint32_t)c.L1, (uint32_t)rowSize, (uint32_t)bStep, (uint32_t)gStep, c.d_prime};
  {
    ProfScope ps_(c, "lt_mac", st);
    const dim3 g = grid_n(c.N, ext);
    // more than 8 baby steps: launches of 8 each, the later ones adding into R (the sums are
    // exact residues, so the split result is the same word); 16 per lane would hold 3 waves per
    // SIMD instead of 5
    if (c.split_mac_ok) {
      if (bStep <= 4) k_lt_mac<4, Mac30><<<g, kT, 0, st>>>(fast.p, diags, R.p, mp);
      else
        for (mp.i0 = 0; mp.i0 < (uint32_t)bStep; mp.i0 += 8, mp.acc = 1) k_lt_mac<8, Mac30><<<g, kT, 0, st>>>(fast.p, diags, R.p, mp);
    } else {
      if (bStep <= 4) k_lt_mac<4, MacWide><<<g, kT, 0, st>>>(fast.p, diags, R.p, mp);
      else
        for (mp.i0 = 0; mp.i0 < (uint32_t)bStep; mp.i0 += 8, mp.acc = 1) k_lt_mac<8, MacWide><<<g, kT, 0, st>>>(fast.p, diags, R.p, mp);
    }
    FH_HIP(hipGetLastError());
  }
  return R;
}

// The last step of both forms on the combined extended sum: the offset rotation (ModDown of c1,
// the extended c0 folded directly) and ModDown of both, or the ModDown of both.
void lt_finish(Context& c, uint64_t* acc, int l, int offset, const uint64_t* off_key, uint32_t off_gal, uint64_t* out, hipStream_t st) {
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t N = c.N, E = (size_t)ext * N;
  if (offset != 0) {
    DBuf c1q(c.pool, (size_t)lq * N), dig(c.pool, (size_t)beta * E), t(c.pool, 2 * E);
    moddown(c, acc + E, l, c1q.p, 1, 0, 0, st);
    modup(c, c1q.p, l, dig.p, st);
    ksk_inner(c, dig.p, l, off_key, t.p, t.p + E, acc, off_gal, st, true);
    moddown(c, t.p, l, out, 2, 0, 0, st);
  } else {
    moddown(c, acc, l, out, 2, 0, 0, st);
  }
}

}  // namespace

void linear_transform(Context& c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                      const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal, const uint64_t* giant_key,
                      uint32_t giant_gal, const uint64_t* off_key, uint32_t off_gal, uint64_t* out, hipStream_t st) {
  if (bStep < 1 || bStep > kMaxBaby) throw Error(-1, "bStep must be in [1, 16]");
  if (rowSize < 1) throw Error(-1, "rowSize must be >= 1");
  if (slots < 1 || slots > (int)c.N / 2) throw Error(-1, "slots out of range");
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t N = c.N, E = (size_t)ext * N;
  const int gStep = (rowSize + bStep - 1) / bStep;
  const bool giant = ((long long)bStep * stride) % slots != 0;
  if (gStep > 1 && giant && !giant_key) throw Error(-1, "giant-step key required");
  if (offset != 0 && !off_key) throw Error(-1, "offset key required");
  DBuf R = lt_partials(c, ct, l, rowSize, bStep, stride, slots, diags, baby_keys, baby_gal, st);

  // ---- Horner over the giant steps ----
  std::vector<int> epids(ext);
  for (uint32_t p = 0; p < ext; ++p) epids[p] = c.ext_prime(l, p);
  // R_{j-1} += rot_giant(R_j) for j = gStep-1 .. 1: the key switch adds its result into R_{j-1}.
  DBuf c1q(c.pool, (size_t)lq * N), dig(c.pool, (size_t)beta * E);
  for (int j = gStep - 1; j > 0; --j) {
    uint64_t* Rj = R.p + (size_t)j * 2 * E;
    uint64_t* Rp = Rj - 2 * E;
    if (giant) {
      moddown(c, Rj + E, l, c1q.p, 1, 0, 0, st);
      modup(c, c1q.p, l, dig.p, st);
      ksk_inner(c, dig.p, l, giant_key, Rp, Rp + E, Rj, giant_gal, st, true, nullptr, true);
    } else {
      LimbSet a = make_set(Rp, ext, epids.data(), 2, E), b = make_set(Rj, ext, epids.data(), 2, E);
      ew_binary(c, 0, a, b, a, st);
    }
  }
  lt_finish(c, R.p, l, offset, off_key, off_gal, out, st);
}

void linear_transform_flat(Context& c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                           const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal,
                           const uint64_t* const* giant_keys, const uint32_t* giant_gal, const uint64_t* off_key, uint32_t off_gal,
                           uint64_t* out, hipStream_t st) {
  if (bStep < 1 || bStep > kMaxBaby) throw Error(-1, "bStep must be in [1, 16]");
  if (rowSize < 1) throw Error(-1, "rowSize must be >= 1");
  if (slots < 1 || slots > (int)c.N / 2) throw Error(-1, "slots out of range");
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t N = c.N, E = (size_t)ext * N;
  const int gStep = (rowSize + bStep - 1) / bStep;
  if (offset != 0 && !off_key) throw Error(-1, "offset key required");
  // the partial sums rotated by a nonzero multiple of the giant step, and their keys
  std::vector<int> rot;
  std::vector<const uint64_t*> keys;
  std::vector<uint32_t> gals;
  for (int j = 1; j < gStep; ++j) {
    if (((long long)j * bStep * stride) % slots == 0) continue;
    if (!giant_keys || !giant_keys[j - 1]) throw Error(-1, "giant-step key required");
    rot.push_back(j);
    keys.push_back(giant_keys[j - 1]);
    gals.push_back(giant_gal[j - 1]);
  }
  DBuf R = lt_partials(c, ct, l, rowSize, bStep, stride, slots, diags, baby_keys, baby_gal, st);
  std::vector<int> epids(ext);
  for (uint32_t p = 0; p < ext; ++p) epids[p] = c.ext_prime(l, p);
  const LimbSet acc = make_set(R.p, ext, epids.data(), 2, E);
  if (!rot.empty()) {
    // the rotated partial sums side by side: one ModDown of their c1, one ModUp, one key inner
    // product with each item's own key (the extended c0 folded directly, as a Horner step does)
    const int nr = (int)rot.size();
    bool contiguous = true;
    for (int t = 1; t < nr; ++t) contiguous &= rot[t] == rot[0] + t;
    DBuf gathered, c1q(c.pool, (size_t)nr * lq * N), dig(c.pool, (size_t)nr * beta * E), T(c.pool, (size_t)nr * 2 * E);
    const uint64_t* src = R.p + (size_t)rot[0] * 2 * E;
    if (!contiguous) {
      gathered = DBuf(c.pool, (size_t)nr * 2 * E);
      for (int t = 0; t < nr; ++t)
        FH_HIP(hipMemcpyAsync(gathered.p + (size_t)t * 2 * E, R.p + (size_t)rot[t] * 2 * E, 2 * E * 8, hipMemcpyDeviceToDevice, st));
      src = gathered.p;
    }
    moddown(c, src + E, l, c1q.p, nr, 2 * E, (uint64_t)lq * N, st);
    modup_batch(c, c1q.p, l, nr, (uint64_t)lq * N, dig.p, st, true);
    ksk_items(c, nr, 1, dig.p, (uint64_t)beta * E, src, 2 * E, true, true, l, keys.data(), gals.data(), T.p, 2 * E, st);
    {
      ProfScope ps_(c, "lt_sum", st);
      k_lt_sum<<<grid_n(c.N, 2 * ext), kT, 0, st>>>(acc, T.p, 2 * E, nr, c.N, c.d_prime);
      FH_HIP(hipGetLastError());
    }
  }
  for (int j = 1; j < gStep; ++j)
    if (((long long)j * bStep * stride) % slots == 0) ew_binary(c, 0, acc, make_set(R.p + (size_t)j * 2 * E, ext, epids.data(), 2, E), acc, st);
  lt_finish(c, R.p, l, offset, off_key, off_gal, out, st);
}

// The same transform for count ciphertexts sharing the diagonal set and keys: the baby-step
// rotations of all ciphertexts form one key-switch batch (each baby key read once for all of
// them), one multiply-accumulate launch reads the diagonals once for the whole batch, and every
// giant step key-switches all ciphertexts with one read of the giant key. Results are bit-identical
// to count calls of linear_transform.
void linear_transform_batch(Context& c, int count, const uint64_t* cts, uint64_t ct_stride, int l, int rowSize, int bStep, int stride,
                            int offset, int slots, const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal,
                            const uint64_t* giant_key, uint32_t giant_gal, const uint64_t* off_key, uint32_t off_gal, uint64_t* out,
                            hipStream_t st) {
  if (count < 1) return;
  if (bStep < 1 || bStep > kMaxBaby) throw Error(-1, "bStep must be in [1, 16]");
  if (rowSize < 1) throw Error(-1, "rowSize must be >= 1");
  if ((size_t)rowSize * kMacTile * 8 > 64 * 1024) throw Error(-3, "batched linear transform takes at most 128 diagonals");
  if (slots < 1 || slots > (int)c.N / 2) throw Error(-1, "slots out of range");
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t N = c.N, E = (size_t)ext * N;
  if (!ct_stride) ct_stride = 2ull * lq * N;
  const int gStep = (rowSize + bStep - 1) / bStep;
  const bool giant = ((long long)bStep * stride) % slots != 0;
  if (gStep > 1 && giant && !giant_key) throw Error(-1, "giant-step key required");
  if (offset != 0 && !off_key) throw Error(-1, "offset key required");
  const uint32_t B = (uint32_t)count;

  // ---- baby steps ----
  MacBatchParams mp{c.N, lq, ext, (uint32_t)c.L1, (uint32_t)rowSize, (uint32_t)bStep, (uint32_t)gStep, B, 0, c.d_prime, {}};
  std::vector<int> rot_idx;
  bool has_ident = false;
  for (int i = 0; i < bStep; ++i) {
    if (((long long)i * stride) % slots == 0) {
      mp.slot[i] = -1;
      has_ident = true;
    } else {
      if (!baby_keys[i]) throw Error(-1, "baby-step key missing");
      mp.slot[i] = (int8_t)rot_idx.size();
      rot_idx.push_back(i);
    }
  }
  const uint32_t nrot = (uint32_t)rot_idx.size();
  mp.nrot = nrot;
  DBuf ident, rot, dig(c.pool, (size_t)B * beta * E);
  if (has_ident) {
    ident = DBuf(c.pool, (size_t)B * 2 * E);
    ProfScope ps_(c, "lt_extend", st);
    k_extend_p_b<<<grid_n(c.N, B * 2 * ext), kT, 0, st>>>(cts, ct_stride, ident.p, 2 * E, c.N, lq, ext, c.d_prime, c.d_pmod, c.d_pmod_sh);
    FH_HIP(hipGetLastError());
  }
  if (nrot) {
    rot = DBuf(c.pool, (size_t)B * nrot * 2 * E);
    modup_batch(c, cts + (size_t)lq * N, l, (int)B, ct_stride, dig.p, st);
    std::vector<const uint64_t*> ks((size_t)B * nrot);
    std::vector<uint32_t> gs((size_t)B * nrot);
    for (uint32_t b = 0; b < B; ++b)
      for (uint32_t r = 0; r < nrot; ++r) {
        ks[(size_t)b * nrot + r] = baby_keys[rot_idx[r]];
        gs[(size_t)b * nrot + r] = baby_gal[rot_idx[r]];
      }
    ksk_items(c, (int)(B * nrot), (int)nrot, dig.p, (uint64_t)beta * E, cts, ct_stride, true, false, l, ks.data(), gs.data(), rot.p, 2 * E,
              st);
  }

  // ---- giant-step partial sums: diagonals read once for the batch ----
  DBuf R(c.pool, (size_t)B * gStep * 2 * E);
  {
    ProfScope ps_(c, "lt_mac", st);
    const dim3 g(c.N / kMacTile, ext);
    const size_t lds = (size_t)rowSize * kMacTile * 8;
    // 30-bit split products when every prime is below 2^60 (sums of at most bStep <= 16 terms)
    if (c.split_mac_ok) {
      if (bStep <= 4) k_lt_mac_b<4, Mac30><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
      else if (bStep <= 8) k_lt_mac_b<8, Mac30><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
      else k_lt_mac_b<16, Mac30><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
    } else {
      if (bStep <= 4) k_lt_mac_b<4, MacWide><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
      else if (bStep <= 8) k_lt_mac_b<8, MacWide><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
      else k_lt_mac_b<16, MacWide><<<g, kT, lds, st>>>(rot.p, ident.p, diags, R.p, mp);
    }
    FH_HIP(hipGetLastError());
  }

  // ---- Horner over the giant steps, all ciphertexts together ----
  std::vector<int> epids(ext);
  for (uint32_t p = 0; p < ext; ++p) epids[p] = c.ext_prime(l, p);
  const uint64_t rstride = (uint64_t)gStep * 2 * E;  // between ciphertexts in R
  DBuf c1q(c.pool, (size_t)B * lq * N);
  std::vector<const uint64_t*> gk(B, giant_key), okv(B, off_key);
  std::vector<uint32_t> gg(B, giant_gal), og(B, off_gal);
  for (int j = gStep - 1; j >= 0; --j) {
    uint64_t* Rj = R.p + (size_t)j * 2 * E;
    if (j > 0) {
      uint64_t* Rp = Rj - 2 * E;  // every ciphertext's R_{j-1} += rot_giant(R_j)
      if (giant) {
        moddown(c, Rj + E, l, c1q.p, (int)B, rstride, (uint64_t)lq * N, st);
        modup_batch(c, c1q.p, l, (int)B, (uint64_t)lq * N, dig.p, st);
        ksk_items(c, (int)B, 1, dig.p, (uint64_t)beta * E, Rj, rstride, true, true, l, gk.data(), gg.data(), Rp, rstride, st,
                  nullptr, 0, true);
      } else {
        for (int comp = 0; comp < 2; ++comp) {
          LimbSet a = make_set(Rp + comp * E, ext, epids.data(), B, rstride), b = make_set(Rj + comp * E, ext, epids.data(), B, rstride);
          ew_binary(c, 0, a, b, a, st);
        }
      }
    } else if (offset != 0) {
      moddown(c, Rj + E, l, c1q.p, (int)B, rstride, (uint64_t)lq * N, st);
      modup_batch(c, c1q.p, l, (int)B, (uint64_t)lq * N, dig.p, st);
      DBuf t(c.pool, (size_t)B * 2 * E);
      ksk_items(c, (int)B, 1, dig.p, (uint64_t)beta * E, Rj, rstride, true, true, l, okv.data(), og.data(), t.p, 2 * E, st);
      moddown(c, t.p, l, out, (int)(2 * B), E, (uint64_t)lq * N, st);
    } else {
      moddown(c, Rj, l, out, (int)B, rstride, 2ull * lq * N, st);
      moddown(c, Rj + E, l, out + (size_t)lq * N, (int)B, rstride, 2ull * lq * N, st);
    }
  }
}

// Rotation-fold sum from a start factor (the cascade: CKKS::Accumulate(ct, bStep, stride, size,
// startFactor), src/CKKS/AccumulateBroadcast.cu:42-84): per fold level s = start, start*radix, ...
// < size, the rotations by i*stride*s (0 < i < radix, i*s < size) share one ModUp and stay in the
// extended basis with P*c0 folded in, are added to P*ct, and both components take a ModDown at
// every level. keys / galois: one per rotation in fold order.
void accumulate_cascade(Context& c, const uint64_t* ct, int l, int radix, int size, int start, const uint64_t* const* keys,
                    const uint32_t* galois, uint64_t* out, hipStream_t st) {
  const uint32_t N = c.N, lq = l + 1, ext = lq + c.K;
  const size_t E = (size_t)ext * N, Q = (size_t)lq * N;
  if (out != ct) FH_HIP(hipMemcpyAsync(out, ct, 2 * Q * 8, hipMemcpyDeviceToDevice, st));
  if (start <= 0 || size <= 0 || start >= size) return;
  std::vector<int> epids(ext);
  for (uint32_t p = 0; p < ext; ++p) epids[p] = c.ext_prime(l, p);
  DBuf fast(c.pool, (size_t)radix * 2 * E);
  int used = 0;
  for (long long s = start; s < size; s *= radix) {
    int n = 0;
    while (n + 1 < radix && (long long)(n + 1) * s < size) ++n;
    {
      ProfScope ps_(c, "lt_extend", st);
      k_extend_p<<<grid_n(N, 2 * ext), kT, 0, st>>>(out, fast.p, N, lq, ext, c.d_prime, c.d_pmod, c.d_pmod_sh);
      FH_HIP(hipGetLastError());
    }
    rotate_hoisted(c, out, l, n, keys + used, galois + used, true, false, fast.p + 2 * E, st);
    LimbSet a = make_set(fast.p, ext, epids.data(), 2, E);
    for (int i = 1; i <= n; ++i) {
      LimbSet b = make_set(fast.p + (size_t)i * 2 * E, ext, epids.data(), 2, E);
      ew_binary(c, 0, a, b, a, st);
    }
    moddown(c, fast.p, l, out, 2, 0, 0, st);
    used += n;
  }
}


// The lazy rotation-fold sum (CKKS::Accumulate(ct, bStep, stride, size),
// src/CKKS/AccumulateBroadcast.cu:122-160). Per fold level s = 1, radix, ... < size: one ModUp of c1
// shared by the level's rotations (by i*stride*s, 0 < i < radix, i*s < size), each kept in Q_l ∪ P;
// c0 is folded as P*c0 on the first level and afterwards stays in Q_l ∪ P, where it is added on
// every limb before the automorphism (the c0_modup branch of hoistedRotateDotKSK_2_,
// src/CKKS/ElemenwiseBatchKernels.cu:302-307). The rotations are summed into (c0, P*c1); c1 takes a
// ModDown at the end of each level and c0 one ModDown after the last. Buffer fast: slot 0 holds the
// accumulators (extended c0, extended c1), slots 1..radix-1 the level's rotations.
void accumulate_sum(Context& c, const uint64_t* ct, int l, int radix, int size, const uint64_t* const* keys, const uint32_t* galois,
                    uint64_t* out, hipStream_t st) {
  const uint32_t N = c.N, lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  const size_t E = (size_t)ext * N, Q = (size_t)lq * N;
  if (out != ct) FH_HIP(hipMemcpyAsync(out, ct, 2 * Q * 8, hipMemcpyDeviceToDevice, st));
  if (size <= 1 || radix < 2) return;
  std::vector<int> epids(ext);
  for (uint32_t p = 0; p < ext; ++p) epids[p] = c.ext_prime(l, p);
  DBuf fast(c.pool, (size_t)radix * 2 * E);
  DBuf dig(c.pool, (size_t)beta * E);
  bool c0_ext = false;
  int used = 0;
  for (long long s = 1; s < size; s *= radix) {
    int n = 0;
    while (n + 1 < radix && (long long)(n + 1) * s < size) ++n;
    // the level's rotations, all from the level's input (c0 or its extension, and c1)
    modup_batch(c, out + Q, l, 1, 0, dig.p, st, false);
    ksk_items(c, n, n, dig.p, 0, c0_ext ? fast.p : out, 0, true, c0_ext, l, keys + used, galois + used, fast.p + 2 * E, 2 * E, st,
              out + Q, 0);
    {
      // accumulators: P*c1 (and P*c0 on the first level; afterwards slot 0 already holds c0)
      ProfScope ps_(c, "lt_extend", st);
      const uint32_t comps = c0_ext ? 1 : 2;
      k_extend_p<<<grid_n(N, comps * ext), kT, 0, st>>>(c0_ext ? out + Q : out, c0_ext ? fast.p + E : fast.p, N, lq, ext, c.d_prime,
                                                       c.d_pmod, c.d_pmod_sh);
      FH_HIP(hipGetLastError());
    }
    LimbSet a = make_set(fast.p, ext, epids.data(), 2, E);
    for (int i = 1; i <= n; ++i) {
      LimbSet b = make_set(fast.p + (size_t)i * 2 * E, ext, epids.data(), 2, E);
      ew_binary(c, 0, a, b, a, st);
    }
    moddown(c, fast.p + E, l, out + Q, 1, 0, 0, st);
    c0_ext = true;
    used += n;
  }
  moddown(c, fast.p, l, out, 1, 0, 0, st);
}

}  // namespace fh
