// Element-wise RNS kernels: add, sub, negate, multiply, scalar ops and the
// CKKS tensor product. Each block walks one limb (blockIdx.y) with 16-byte
// vector loads; these kernels are HBM-bound by construction.
// Reference: src/AddSub.cuh:12-50, src/CKKS/ElemenwiseBatchKernels.cuh:13-99,
// src/CKKS/ElemenwiseBatchKernels.cu:1363-1510 (binomialMult_ / binomialSquare_).
#include "fh_ops.hpp"

namespace fh {

namespace {

constexpr int kEwThreads = 256;

struct V2 {
  uint64_t x, y;
};

__device__ __forceinline__ V2 ld2(const uint64_t* p) {
  ulonglong2 v = *reinterpret_cast<const ulonglong2*>(p);
  return V2{v.x, v.y};
}
__device__ __forceinline__ void st2(uint64_t* p, uint64_t a, uint64_t b) {
  *reinterpret_cast<ulonglong2*>(p) = make_ulonglong2(a, b);
}

// out = op(a, b) limb-wise; a, b, out share the prime map of `a`.
template <int OP>
__global__ __launch_bounds__(kEwThreads) void k_binary(LimbSet a, LimbSet b, LimbSet o, const PrimeConst* pr, uint32_t N) {
  uint64_t *pa, *pb, *po;
  int pid, pid2, pid3;
  SetAddr{a}.at(blockIdx.y, N, pa, pid);
  SetAddr{b}.at(blockIdx.y, N, pb, pid2);
  SetAddr{o}.at(blockIdx.y, N, po, pid3);
  const PrimeConst P = pr[pid];
  for (uint32_t i = 2 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 2 * gridDim.x * blockDim.x) {
    V2 x = ld2(pa + i), y = ld2(pb + i);
    uint64_t r0, r1;
    if constexpr (OP == 0) { r0 = add_mod(x.x, y.x, P.q); r1 = add_mod(x.y, y.y, P.q); }
    else if constexpr (OP == 1) { r0 = sub_mod(x.x, y.x, P.q); r1 = sub_mod(x.y, y.y, P.q); }
    else { r0 = mul_mod(x.x, y.x, P); r1 = mul_mod(x.y, y.y, P); }
    st2(po + i, r0, r1);
  }
}

struct Scalars {
  uint64_t v[kMaxLimbs];
  uint64_t sh[kMaxLimbs];
};

// OP 0: a + s_limb ; OP 1: a * s_limb (Shoup) ; OP 2: -a
template <int OP>
__global__ __launch_bounds__(kEwThreads) void k_unary(LimbSet a, LimbSet o, Scalars sc, const PrimeConst* pr, uint32_t N) {
  uint64_t *pa, *po;
  int pid, pid2;
  SetAddr{a}.at(blockIdx.y, N, pa, pid);
  SetAddr{o}.at(blockIdx.y, N, po, pid2);
  const uint64_t q = pr[pid].q;
  const uint32_t limb = blockIdx.y / a.npolys;
  const uint64_t s = sc.v[limb], ssh = sc.sh[limb];
  for (uint32_t i = 2 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 2 * gridDim.x * blockDim.x) {
    V2 x = ld2(pa + i);
    uint64_t r0, r1;
    if constexpr (OP == 0) { r0 = add_mod(x.x, s, q); r1 = add_mod(x.y, s, q); }
    else if constexpr (OP == 1) { r0 = shoup(x.x, s, ssh, q); r1 = shoup(x.y, s, ssh, q); }
    else { r0 = x.x ? q - x.x : 0; r1 = x.y ? q - x.y : 0; }
    st2(po + i, r0, r1);
  }
}

// (c0, c1) x (d0, d1) -> (c0 d0, c0 d1 + c1 d0, c1 d1) over the first nl limbs; square when SQ.
// The inputs hold lc / ld limbs per component; blockIdx.z: item (ciphertexts back to back, outputs
// 3 nl limbs apart).
template <bool SQ>
__global__ __launch_bounds__(kEwThreads) void k_tensor(const uint64_t* c, const uint64_t* d, uint64_t* o, uint32_t nl, const PrimeConst* pr, uint32_t N,
                                                       uint32_t lc, uint32_t ld) {
  const uint32_t limb = blockIdx.y;
  const PrimeConst P = pr[limb];
  const size_t cc = (size_t)lc * N, dc = (size_t)ld * N, comp = (size_t)nl * N, off = (size_t)limb * N;
  c += blockIdx.z * 2 * cc;
  d += blockIdx.z * 2 * dc;
  o += blockIdx.z * 3 * comp;
  const uint64_t *c0 = c + off, *c1 = c + cc + off, *d0 = d + off, *d1 = d + dc + off;
  uint64_t *o0 = o + off, *o1 = o + comp + off, *o2 = o + 2 * comp + off;
  for (uint32_t i = 2 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 2 * gridDim.x * blockDim.x) {
    V2 a0 = ld2(c0 + i), a1 = ld2(c1 + i);
    V2 b0 = SQ ? a0 : ld2(d0 + i), b1 = SQ ? a1 : ld2(d1 + i);
    uint64_t x0 = mul_mod(a0.x, b0.x, P), y0 = mul_mod(a0.y, b0.y, P);
    uint64_t x2 = mul_mod(a1.x, b1.x, P), y2 = mul_mod(a1.y, b1.y, P);
    uint64_t x1, y1;
    if constexpr (SQ) {
      x1 = mul_mod(a0.x, a1.x, P); x1 = add_mod(x1, x1, P.q);
      y1 = mul_mod(a0.y, a1.y, P); y1 = add_mod(y1, y1, P.q);
    } else {
      // c0 d1 + c1 d0 accumulated in 128 bits, one reduction
      U128 s; s.zero(); s.mac(a0.x, b1.x); s.mac(a1.x, b0.x); x1 = mod128(s.hi, s.lo, P);
      U128 t; t.zero(); t.mac(a0.y, b1.y); t.mac(a1.y, b0.y); y1 = mod128(t.hi, t.lo, P);
    }
    st2(o0 + i, x0, y0);
    st2(o1 + i, x1, y1);
    st2(o2 + i, x2, y2);
  }
}

// Weighted sum of n ciphertexts over their first `towers` limbs: out_c = sum_k w_k * in_k,c (+ add
// on the components whose bit is set in add_mask). in_k holds ncomp components of limbs_k >= towers
// limbs; w is [n][towers] residues, add (optional) [towers] residues. One block row per
// (component, limb).
struct WSumIn {
  const uint64_t* p[kMaxWSum];
  uint32_t limbs[kMaxWSum];
};
// weights and the added constant in the kernel arguments when n x towers <= kWSumInline (no upload)
constexpr int kWSumInline = 320;
struct WSumConst {
  uint64_t w[kWSumInline];
  uint64_t add[kMaxLimbs];
};

__device__ __forceinline__ void wsum_body(const WSumIn& in, int n, const uint64_t* w, uint64_t a0, uint32_t comp, uint32_t limb, uint32_t towers,
                                          uint64_t* out, const PrimeConst& P, uint32_t N) {
  uint64_t* po = out + ((size_t)comp * towers + limb) * N;
  for (uint32_t i = 2 * (blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 2 * gridDim.x * blockDim.x) {
    uint64_t r0 = a0, r1 = a0;
    for (int k = 0; k < n; ++k) {
      const uint64_t wk = w[(size_t)k * towers + limb];
      const V2 x = ld2(in.p[k] + ((size_t)comp * in.limbs[k] + limb) * N + i);
      r0 = add_mod(r0, mul_mod(x.x, wk, P), P.q);
      r1 = add_mod(r1, mul_mod(x.y, wk, P), P.q);
    }
    st2(po + i, r0, r1);
  }
}

__global__ __launch_bounds__(kEwThreads) void k_wsum_inline(WSumIn in, WSumConst kc, int n, int has_add, uint32_t add_mask, uint32_t towers,
                                                            uint64_t* out, const PrimeConst* pr, uint32_t N) {
  const uint32_t comp = blockIdx.y / towers, limb = blockIdx.y % towers;
  const PrimeConst P = pr[limb];
  const uint64_t a0 = has_add && ((add_mask >> comp) & 1u) ? kc.add[limb] : 0;
  wsum_body(in, n, kc.w, a0, comp, limb, towers, out, P, N);
}

__global__ __launch_bounds__(kEwThreads) void k_wsum(WSumIn in, int n, const uint64_t* __restrict__ w, const uint64_t* __restrict__ add,
                                                     uint32_t add_mask, uint32_t towers, uint64_t* out, const PrimeConst* pr, uint32_t N) {
  const uint32_t comp = blockIdx.y / towers, limb = blockIdx.y % towers;
  const PrimeConst P = pr[limb];
  const uint64_t a0 = add && ((add_mask >> comp) & 1u) ? add[limb] : 0;
  wsum_body(in, n, w, a0, comp, limb, towers, out, P, N);
}

// splitmix64 finalizer: a counter-based stream, one 64-bit word per (seed, limb job, index)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kEwThreads) void k_fill_uniform(LimbSet s, const PrimeConst* pr, uint32_t N, uint64_t seed) {
  uint64_t* p;
  int pid;
  SetAddr{s}.at(blockIdx.y, N, p, pid);
  const PrimeConst P = pr[pid];
  const uint64_t key = mix64(seed ^ ((uint64_t)blockIdx.y << 32));
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const uint64_t r = mix64(key + i);
    p[i] = barrett128(0, r, P);  // r mod q (uniform up to a 2^-13 bias, irrelevant for synthetic data)
  }
}

// STREAM copy (b[i] = a[i]) in 16-byte words: each workgroup moves contiguous chunks of 4 x 256
// words (16 KiB), its lanes holding 4 loads in flight, chunks dealt out grid-stride. Contiguous
// per-workgroup chunks keep DRAM pages open; on MI355X this reaches about 5.7 TB/s at 1 GiB against
// 4.6 TB/s for a lane-interleaved grid-stride loop and 5.2 TB/s for a DMA-engine hipMemcpy. It is
// the achievable HBM bandwidth the kernels are compared against.
constexpr int kCopyU = 4;
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n16) {
  constexpr size_t per = (size_t)kCopyU * 256;
  for (size_t base = (size_t)blockIdx.x * per; base < n16; base += (size_t)gridDim.x * per) {
    uint4 x[kCopyU];
#pragma unroll
    for (int u = 0; u < kCopyU; ++u) {
      const size_t i = base + (size_t)u * 256 + threadIdx.x;
      if (i < n16) x[u] = a[i];
    }
#pragma unroll
    for (int u = 0; u < kCopyU; ++u) {
      const size_t i = base + (size_t)u * 256 + threadIdx.x;
      if (i < n16) b[i] = x[u];
    }
  }
}

inline dim3 ew_grid(uint32_t N, uint32_t nlimbs) {
  uint32_t per_block = 2 * kEwThreads;
  uint32_t gx = (N + per_block - 1) / per_block;
  if (gx > 64) gx = 64;
  return dim3(gx, nlimbs);
}

}  // namespace

void ew_binary(const Context& c, int op, const LimbSet& a, const LimbSet& b, const LimbSet& o, hipStream_t st) {
  uint32_t n = a.nlimbs * a.npolys;
  if (!n) return;
  dim3 g = ew_grid(c.N, n);
  ProfScope ps_(c, "eltwise", st);
  switch (op) {
    case 0: k_binary<0><<<g, kEwThreads, 0, st>>>(a, b, o, c.d_prime, c.N); break;
    case 1: k_binary<1><<<g, kEwThreads, 0, st>>>(a, b, o, c.d_prime, c.N); break;
    default: k_binary<2><<<g, kEwThreads, 0, st>>>(a, b, o, c.d_prime, c.N); break;
  }
  FH_HIP(hipGetLastError());
}

void ew_wsum(Context& c, int n, const uint64_t* const* in, const uint32_t* in_limbs, const uint64_t* weights, int towers, int ncomp,
             const uint64_t* add, uint32_t add_mask, uint64_t* out, hipStream_t st) {
  if (n < 1 || n > kMaxWSum) throw Error(-1, "weighted sum of 1.." + std::to_string(kMaxWSum) + " ciphertexts");
  if (towers < 1 || towers > c.L1 || ncomp < 1 || ncomp > 8) throw Error(-1, "weighted sum: shape out of range");
  WSumIn wi{};
  for (int k = 0; k < n; ++k) {
    if (!in[k]) throw Error(-1, "weighted sum: null input");
    if (in_limbs[k] < (uint32_t)towers || in_limbs[k] > (uint32_t)c.L1) throw Error(-1, "weighted sum: an input's limb count is out of range");
    wi.p[k] = in[k];
    wi.limbs[k] = in_limbs[k];
  }
  dim3 g = ew_grid(c.N, (uint32_t)(towers * ncomp));
  if ((size_t)n * towers <= (size_t)kWSumInline) {
    // weights (reduced) and the constant travel in the kernel arguments
    WSumConst kc;
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < towers; ++i) kc.w[(size_t)k * towers + i] = weights[(size_t)k * towers + i] % c.q[i];
    for (int i = 0; i < towers; ++i) kc.add[i] = add ? add[i] % c.q[i] : 0;
    ProfScope ps_(c, "eltwise", st);
    k_wsum_inline<<<g, kEwThreads, 0, st>>>(wi, kc, n, add ? 1 : 0, add_mask, (uint32_t)towers, out, c.d_prime, c.N);
    FH_HIP(hipGetLastError());
    return;
  }
  // weights (reduced) and the added constant, uploaded in stream order into one pool buffer
  std::vector<uint64_t> h((size_t)(n + 1) * towers, 0);
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < towers; ++i) h[(size_t)k * towers + i] = weights[(size_t)k * towers + i] % c.q[i];
  if (add)
    for (int i = 0; i < towers; ++i) h[(size_t)n * towers + i] = add[i] % c.q[i];
  DBuf d(c.pool, h.size());
  FH_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, st));
  ProfScope ps_(c, "eltwise", st);
  k_wsum<<<g, kEwThreads, 0, st>>>(wi, n, d.p, add ? d.p + (size_t)n * towers : nullptr, add_mask, (uint32_t)towers, out, c.d_prime, c.N);
  FH_HIP(hipGetLastError());
}

void ew_unary(const Context& c, int op, const LimbSet& a, const LimbSet& o, const uint64_t* scalars, hipStream_t st) {
  uint32_t n = a.nlimbs * a.npolys;
  if (!n) return;
  Scalars sc{};
  for (uint32_t i = 0; i < a.nlimbs; ++i) {
    uint64_t q = c.q[a.pid[i]];
    uint64_t v = scalars ? scalars[i] % q : 0;
    sc.v[i] = v;
    sc.sh[i] = h_shoup(v, q);
  }
  dim3 g = ew_grid(c.N, n);
  ProfScope ps_(c, "eltwise", st);
  switch (op) {
    case 0: k_unary<0><<<g, kEwThreads, 0, st>>>(a, o, sc, c.d_prime, c.N); break;
    case 1: k_unary<1><<<g, kEwThreads, 0, st>>>(a, o, sc, c.d_prime, c.N); break;
    default: k_unary<2><<<g, kEwThreads, 0, st>>>(a, o, sc, c.d_prime, c.N); break;
  }
  FH_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kEwThreads) void k_to_mont(LimbSet s, const PrimeConst* pr, uint32_t N) {
  uint64_t* p;
  int pid;
  SetAddr{s}.at(blockIdx.y, N, p, pid);
  const PrimeConst P = pr[pid];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) p[i] = mont_mul(p[i], P.mont_r2, P);
}

void to_montgomery(const Context& c, const LimbSet& s, hipStream_t st) {
  uint32_t n = s.nlimbs * s.npolys;
  if (!n) return;
  k_to_mont<<<ew_grid(c.N, n), kEwThreads, 0, st>>>(s, c.d_prime, c.N);
  FH_HIP(hipGetLastError());
}

void stream_copy(const Context& c, void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) throw Error(-1, "stream copy: 16-byte aligned sizes and pointers only");
  const size_t n16 = bytes / 16;
  // one 16 KiB chunk per workgroup up to 32768 workgroups (128 per CU), then grid-stride
  const size_t per = (size_t)kCopyU * 256;
  const size_t g = std::min<size_t>(32768, (n16 + per - 1) / per);
  ProfScope ps_(c, "stream_copy", st);
  k_stream_copy<<<(unsigned)g, 256, 0, st>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
  FH_HIP(hipGetLastError());
}

void fill_uniform(const Context& c, const LimbSet& s, uint64_t seed, hipStream_t st) {
  uint32_t n = s.nlimbs * s.npolys;
  if (!n) return;
  if (n > 65535u) throw Error(-1, "too many limbs in one fill");
  k_fill_uniform<<<ew_grid(c.N, n), kEwThreads, 0, st>>>(s, c.d_prime, c.N, seed);
  FH_HIP(hipGetLastError());
}

void ew_tensor(const Context& c, const uint64_t* ct1, const uint64_t* ct2, uint64_t* out3, int nlimbs, hipStream_t st, int n, int limbs1,
               int limbs2) {
  if (nlimbs <= 0 || n <= 0) return;
  if (n > 65535) throw Error(-1, "too many tensor products in one launch");
  if (limbs1 == 0) limbs1 = nlimbs;
  if (limbs2 == 0) limbs2 = nlimbs;
  if (limbs1 < nlimbs || limbs2 < nlimbs) throw Error(-1, "tensor inputs have fewer limbs than the product");
  dim3 g = ew_grid(c.N, nlimbs);
  g.z = (uint32_t)n;
  ProfScope ps_(c, "tensor", st);
  if (ct1 == ct2 && limbs1 == limbs2) k_tensor<true><<<g, kEwThreads, 0, st>>>(ct1, ct1, out3, nlimbs, c.d_prime, c.N, limbs1, limbs1);
  else k_tensor<false><<<g, kEwThreads, 0, st>>>(ct1, ct2, out3, nlimbs, c.d_prime, c.N, limbs1, limbs2);
  FH_HIP(hipGetLastError());
}

}  // namespace fh
