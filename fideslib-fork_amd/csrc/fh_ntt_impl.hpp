// Negacyclic NTT / INTT over 64-bit RNS limbs for gfx950.
//
// The transform is OpenFHE's: forward Cooley-Tukey with the bit-reversed table
// tw[k] = psi^brev(k) (stage m = 1..N/2 uses tw[m + i] for group i), output in
// bit-reversed EVALUATION order; inverse Gentleman-Sande with psi^-brev(k) and
// a final N^-1. Reference: src/NTT.cu:42-310 (INTT), 331-756 (NTT),
// src/CKKS/LimbPartition.cu:433-550 (launch), src/NTThelper.cuh:21-72 (butterflies).
//
// MI355X design
//  * N = S1 * S2 (S1 = 2^l1 "column" points, S2 = 2^l2 "row" points). The first
//    l1 stages only couple elements of one column (stride S2), the last l2
//    stages only elements of one row (contiguous), so a limb is transformed by
//    two launches. Row rho's stages equal a standard S2-point negacyclic
//    transform of x_c tau_rho^c (tau_rho = psi^(1 + 2 brev(rho) - S1)), so the row
//    launch multiplies by a per-element twist table once and then uses the
//    same S2 twiddles (kept in LDS) for every row, like the column launch. Each lane owns 16 elements of one sub-transform: four radix-2
//    stages run in registers, one LDS exchange re-partitions the sub-transform,
//    the remaining stages run in registers again. Global loads and stores are
//    coalesced in both passes (rows are staged through LDS where the register
//    layout is not).
//  * Two exact arithmetics, chosen per limb by its prime:
//      - fp64 (q < 2^50.2, the 50-bit scaling primes): residues are doubles in a
//        signed lazy range; a product is split exactly with FMA (h = yw,
//        l = fma(y, w, -h)), the quotient is rint(h / q) and the remainder
//        fma(-c, q, h) + l is exact. gfx950 issues fp64 FMA at full rate, about
//        three times the rate of a 64-bit Shoup product built from 32-bit
//        multiplies. Values are re-centred into (-q/2, q/2] every few stages so
//        every intermediate stays an exact integer below 2^53.
//      - integer Harvey-lazy Shoup for the 60-bit primes: forward values in
//        [0, 4q), inverse values in [0, 2q).
//    Both are exact, so the canonical outputs are identical.
//  * Between the two launches the limb holds an internal form (lazy integers or
//    doubles). A batch may be split into chunks of limbs (tunable "ntt_chunk_mb")
//    so that the intermediate is still in the MALL when the second launch reads
//    it back; by default one launch pair covers the whole batch.
//  * The last forward launch hands its outputs to an epilogue, which fuses the
//    element-wise step that consumes them (ModDown, Rescale).
#pragma once
#include <algorithm>
#include <cstring>

#include "fh_kernels.hpp"

namespace fh {

namespace {

// workgroup sizes the passes are built for (the default and the wide tile)
constexpr int kNttThreads = 256, kNttWide = 512;

// ------------------------------------------------------------------------------- integer arithmetic
struct IntA {
  using T = uint64_t;
  uint64_t q, q2;
  struct W {
    uint64_t w, sh;
  };
  __device__ __forceinline__ W tw(const uint64_t* t, const uint64_t* tsh, uint32_t i) const { return W{t[i], tsh[i]}; }
  __device__ __forceinline__ T from_canon(uint64_t v) const { return v; }
  __device__ __forceinline__ T from_mid(uint64_t v) const { return v; }
  __device__ __forceinline__ uint64_t to_mid(T v) const { return v; }
  __device__ __forceinline__ void fwd(T& X, T& Y, W w) const {
    uint64_t x = X >= q2 ? X - q2 : X;         // [0, 2q)
    uint64_t t = shoup_lazy(Y, w.w, w.sh, q);  // [0, 2q)
    X = x + t;                                 // [0, 4q)
    Y = x - t + q2;                            // [0, 4q)
  }
  __device__ __forceinline__ void inv(T& X, T& Y, W w) const {
    uint64_t u = X + Y;                        // X, Y in [0, 2q)
    u = u >= q2 ? u - q2 : u;
    uint64_t v = X - Y + q2;                   // [0, 4q)
    X = u;
    Y = shoup_lazy(v, w.w, w.sh, q);           // [0, 2q)
  }
  __device__ __forceinline__ void recenter(T&) const {}
  // x * w for any x < 2^64, result in [0, 2q)
  __device__ __forceinline__ T mul(T x, W w) const { return shoup_lazy(x, w.w, w.sh, q); }
  __device__ __forceinline__ uint64_t canon(T x) const {
    x = x >= q2 ? x - q2 : x;
    return x >= q ? x - q : x;
  }
};

// ------------------------------------------------------------------------------- fp64 arithmetic
struct FpA {
  using T = double;
  using W = double;
  double q, qinv, half;
  __device__ __forceinline__ W tw(const double* t, const double*, uint32_t i) const { return t[i]; }
  // y * w mod q for |y| < 4q, |w| <= q/2: exact, result in about (-1.3q, 1.3q)
  __device__ __forceinline__ double mulmod(double y, double w) const {
    const double h = y * w;
    const double l = fma(y, w, -h);
    const double c = rint(h * qinv);
    const double d = fma(-c, q, h);
    return d + l;
  }
  __device__ __forceinline__ T from_canon(uint64_t v) const {
    const double x = (double)v;
    return x > half ? x - q : x;
  }
  __device__ __forceinline__ T from_mid(uint64_t v) const { return __longlong_as_double((long long)v); }
  __device__ __forceinline__ uint64_t to_mid(T v) const { return (uint64_t)__double_as_longlong(v); }
  __device__ __forceinline__ void fwd(T& X, T& Y, W w) const {
    const double t = mulmod(Y, w);
    Y = X - t;
    X = X + t;
  }
  __device__ __forceinline__ void inv(T& X, T& Y, W w) const {
    const double u = X + Y, v = X - Y;
    X = u;
    Y = mulmod(v, w);
  }
  __device__ __forceinline__ void recenter(T& x) const { x = fma(-rint(x * qinv), q, x); }
  // x * w re-centred into [-q/2, q/2]
  __device__ __forceinline__ T mul(T x, W w) const {
    T r = mulmod(x, w);
    recenter(r);
    return r;
  }
  __device__ __forceinline__ uint64_t canon(T x) const {
    recenter(x);
    x = x < 0 ? x + q : x;
    x = x >= q ? x - q : x;
    return (uint64_t)x;
  }
};

// Loads of the second pass. PlainLd: ordinary loads (the two-launch form: the first launch's
// stores are visible to the next launch). L2Ld: loads that skip the CU's L1 (sc0) and are served by
// the L2 -- the single-launch team form, where the data was written by other workgroups of the same
// launch after this CU may have cached the addresses.
struct PlainLd {
  __device__ __forceinline__ static uint64_t ld(const uint64_t* p) { return *p; }
};
struct L2Ld {
  __device__ __forceinline__ static uint64_t ld(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

// per-limb arithmetic selection
struct Sel {
  bool fp;
  IntA ia;
  FpA fa;
  const uint64_t *tw, *twsh;
  const double* twd;
  const uint64_t *ts, *tssh;   // row twists of the limb's prime (forward: tau^c, inverse: tau^-c)
  const double* tsd;
};

__device__ __forceinline__ Sel select(const NttTables& tb, int pid, bool inverse) {
  const PrimeConst& P = tb.prime[pid];
  Sel s;
  s.fp = P.fp64 != 0;
  s.ia = IntA{P.q, 2 * P.q};
  s.fa = FpA{P.qd, P.qinv, 0.5 * P.qd};
  const size_t o = (size_t)pid * tb.N;
  s.tw = (inverse ? tb.itw : tb.tw) + o;
  s.twsh = (inverse ? tb.itw_sh : tb.tw_sh) + o;
  s.twd = (inverse ? tb.itwd : tb.twd) + o;
  s.ts = (inverse ? tb.itwist : tb.twist) + o;
  s.tssh = (inverse ? tb.itwist_sh : tb.twist_sh) + o;
  s.tsd = (inverse ? tb.itwistd : tb.twistd) + o;
  return s;
}

// Thread index policies of the pass bodies. Inside a loop (the team kernels run their limbs in
// one), the compiler hoists the column passes' 16 strided per-lane addresses out of the loop as
// loop invariants, which doubled those kernels' registers; an index the compiler cannot see
// through (OpaqueTid) keeps them in the body. Single-pass launches use the plain index.
struct PlainTid {
  __device__ __forceinline__ static int get() { return threadIdx.x; }
};
struct OpaqueTid {
  __device__ __forceinline__ static int get() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
  }
};

// ------------------------------------------------------------------------------- forward, columns
// First l1 stages. Input: canonical (in place or out of place), output: internal form.
template <int LS, class Tid = PlainTid, class A, class TwT>
__device__ __forceinline__ void fwd_cols_body(const A& ar, const TwT* sw, const TwT* swp, const uint64_t* p, uint64_t* po,
                                              uint64_t* tile, uint32_t N2, int C, uint32_t bx) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1;
  const int tid = Tid::get();
  const int cc = tid % C, r0 = tid / C;
  const uint32_t col = bx * C + cc;
  typename A::T x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = ar.from_canon(p[(size_t)(r0 + T * k) * N2 + col]);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      ar.fwd(x[k], x[k + h], ar.tw(sw, swp, idx));
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) ar.recenter(x[k]);
  if constexpr (LS > 4) {
    uint64_t* t = tile + cc * PADS;
#pragma unroll
    for (int k = 0; k < 16; ++k) t[r0 + T * k] = ar.to_mid(x[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(t[16 * r0 + k]);
#pragma unroll
    for (int s = 4; s < LS; ++s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * r0 + k) >> (LS - s));
        ar.fwd(x[k], x[k + h], ar.tw(sw, swp, idx));
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      ar.recenter(x[k]);
      po[(size_t)(16 * r0 + k) * N2 + col] = ar.to_mid(x[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) po[(size_t)(r0 + T * k) * N2 + col] = ar.to_mid(x[k]);
  }
}

template <int LS, int TH, class Addr, class AddrOut>
__global__ __launch_bounds__(TH) void k_ntt_fwd_cols(Addr ad, AddrOut ao, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, CMAX = TH / T;
  __shared__ uint64_t tile[CMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t *p, *po;
  int pid, pido;
  ad.at(blockIdx.y, tb.N, p, pid);
  ao.at(blockIdx.y, tb.N, po, pido);
  const Sel sel = select(tb, pid, false);
  const int C = blockDim.x / T;
  const uint32_t N2 = tb.N >> LS;
  if (sel.fp) {
    double* swd = reinterpret_cast<double*>(sw);
    for (int i = threadIdx.x; i < S; i += blockDim.x) swd[i] = sel.twd[i];
    fwd_cols_body<LS>(sel.fa, (const double*)swd, (const double*)swd, p, po, tile, N2, C, blockIdx.x);
  } else {
    for (int i = threadIdx.x; i < S; i += blockDim.x) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
    fwd_cols_body<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, p, po, tile, N2, C, blockIdx.x);
  }
}

// ------------------------------------------------------------------------------- row tiles in LDS
// A row pass writes its tile row-major (word c0 + T k) and reads it transposed (word 16 c0 + k),
// i.e. at a stride of 16 words across lanes, which maps pairs of lanes onto the same LDS banks.
// One pad word every 16 words makes that stride 17: the T lanes of a row then hit distinct bank
// pairs in both patterns, and a row length of 17 T words (= T mod 32) places the rows sharing a
// half-wave on disjoint banks.
__device__ __forceinline__ int rsw(int e) { return e + (e >> 4); }
template <int LS>
struct RowTile {
  static constexpr int S = 1 << LS;
  static constexpr int PAD = S + S / 16;   // words per row in LDS; rsw(S - 1) = PAD - 2
};

// ------------------------------------------------------------------------------- forward, rows
// Last l2 stages. Row rho of the intermediate is finished by a standard S2-point negacyclic
// transform of x_c tau_rho^c (c = position in the row): one coalesced twist multiply per element,
// then stages whose twiddles are the first S2 entries of the table, the same for every row, read
// from LDS. Input: internal form, output: canonical residues handed to the epilogue.
//
// fwd_rows_core: the stages of one row (at `row`, its first word at limb position rbase) by the T
// threads of row slot rr of the tile. With LS > 4 the canonical outputs end in the LDS tile (word
// rsw(c0 + T k) of the slot holds position c0 + T k) after a barrier, where the row pass's store
// phase and the fused key inner product read them; with LS <= 4 (no LDS exchange) they go to the
// epilogue directly.
template <int LS, class Ld = PlainLd, class A, class TwT, class Epi>
__device__ __forceinline__ void fwd_rows_core(const A& ar, const TwT* sw, const TwT* swp, const TwT* ts, const TwT* tsp, const uint64_t* row,
                                              uint32_t rbase, uint64_t* tile, int rr, int c0, const Epi& epi) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD;
  typename A::T x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(Ld::ld(row + c0 + T * k));
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = ar.mul(x[k], ar.tw(ts + rbase, tsp + rbase, c0 + T * k));
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      ar.fwd(x[k], x[k + h], ar.tw(sw, swp, idx));
    }
  }
  if constexpr (LS > 4) {
    uint64_t* t = tile + rr * PADS;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      ar.recenter(x[k]);
      t[rsw(c0 + T * k)] = ar.to_mid(x[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(t[rsw(16 * c0 + k)]);
#pragma unroll
    for (int s = 4; s < LS; ++s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * c0 + k) >> (LS - s));
        ar.fwd(x[k], x[k + h], ar.tw(sw, swp, idx));
      }
    }
    // each thread rewrites exactly the LDS words it read: no barrier needed before the store
#pragma unroll
    for (int k = 0; k < 16; ++k) t[rsw(16 * c0 + k)] = ar.canon(x[k]);
    __syncthreads();
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) epi.store(const_cast<uint64_t*>(row), rbase, c0 + T * k, ar.canon(x[k]));
  }
}

template <int LS, class Ld = PlainLd, class A, class TwT, class Epi>
__device__ __forceinline__ void fwd_rows_body(const A& ar, const TwT* sw, const TwT* swp, const TwT* ts, const TwT* tsp, uint64_t* p,
                                              uint64_t* tile, int R, const Epi& epi, uint32_t bx) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD;
  const int tid = threadIdx.x;
  const int rr = tid / T, c0 = tid % T;
  const uint32_t rho = bx * R + rr;
  uint64_t* row = p + (size_t)rho * S;
  const uint32_t rbase = rho * S;
  fwd_rows_core<LS, Ld>(ar, sw, swp, ts, tsp, row, rbase, tile, rr, c0, epi);
  if constexpr (LS > 4) {
    const uint64_t* t = tile + rr * PADS;
#pragma unroll
    for (int k = 0; k < 16; ++k) epi.store(row, rbase, c0 + T * k, t[rsw(c0 + T * k)]);
  }
}

// ------------------------------------------------------------------------------- forward epilogues
// The last launch of a forward NTT hands every canonical output word to an epilogue, which
// either stores it in place (EpiStore) or fuses the element-wise step that consumes it: the
// ModDown combination (x_i - v) * [P^-1]_{q_i} (src/NTTfusions.cuh:237-252) or the Rescale
// combination x_i * [q_l^-1]_{q_i} + v (src/NTTfusions.cuh:214-235). Job y addresses
// (poly, limb) = (y % npolys, y / npolys), the order of SetAddr.
struct EpiStore {
  __device__ __forceinline__ void init(uint32_t, uint32_t, const PrimeConst*) {}
  __device__ __forceinline__ void store(uint64_t* row, uint32_t, uint32_t j, uint64_t v) const { row[j] = v; }
};

struct EpiModDown {
  const uint64_t* x;       // ext polys, limb i of poly b at x + b * x_stride + i * N
  uint64_t x_stride;
  uint64_t* out;           // Q_l polys, limb i of poly b at out + b * out_stride + i * N
  uint64_t out_stride;
  const uint64_t* w;       // [P^-1]_{q_i}, Shoup companions
  const uint64_t* wsh;
  uint32_t npolys;
  // per block
  const uint64_t* xi;
  uint64_t* oi;
  uint64_t q, wi, wish;
  __device__ __forceinline__ void init(uint32_t y, uint32_t N, const PrimeConst* prime) {
    const uint32_t i = y / npolys, b = y % npolys;
    xi = x + b * x_stride + (size_t)i * N;
    oi = out + b * out_stride + (size_t)i * N;
    q = prime[i].q;
    wi = w[i];
    wish = wsh[i];
  }
  __device__ __forceinline__ void store(uint64_t*, uint32_t rbase, uint32_t j, uint64_t v) const {
    oi[rbase + j] = shoup(sub_mod(xi[rbase + j], v, q), wi, wish, q);
  }
};

struct EpiRescale {
  const uint64_t* x;       // ncomp components of l+1 limbs each
  uint64_t* out;           // ncomp components of l limbs each
  const uint64_t* w;       // [q_l^-1]_{q_i} row of level l, Shoup companions
  const uint64_t* wsh;
  uint32_t l, npolys;
  const uint64_t* xi;
  uint64_t* oi;
  uint64_t q, wi, wish;
  __device__ __forceinline__ void init(uint32_t y, uint32_t N, const PrimeConst* prime) {
    const uint32_t i = y / npolys, b = y % npolys;
    xi = x + ((size_t)b * (l + 1) + i) * N;
    oi = out + ((size_t)b * l + i) * N;
    q = prime[i].q;
    wi = w[i];
    wish = wsh[i];
  }
  __device__ __forceinline__ void store(uint64_t*, uint32_t rbase, uint32_t j, uint64_t v) const {
    oi[rbase + j] = add_mod(shoup(xi[rbase + j], wi, wish, q), v, q);
  }
};

template <int LS, int TH, class Addr, class Epi>
__global__ __launch_bounds__(TH) void k_ntt_fwd_rows(Addr ad, NttTables tb, Epi epi) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD, RMAX = TH / T;
  __shared__ uint64_t tile[RMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t* p;
  int pid;
  ad.at(blockIdx.y, tb.N, p, pid);
  epi.init(blockIdx.y, tb.N, tb.prime);
  const Sel sel = select(tb, pid, false);
  const int R = blockDim.x / T;
  if (sel.fp) {
    double* swd = reinterpret_cast<double*>(sw);
    for (int i = threadIdx.x; i < S; i += blockDim.x) swd[i] = sel.twd[i];
    __syncthreads();
    fwd_rows_body<LS>(sel.fa, (const double*)swd, (const double*)swd, sel.tsd, sel.tsd, p, tile, R, epi, blockIdx.x);
  } else {
    for (int i = threadIdx.x; i < S; i += blockDim.x) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
    __syncthreads();
    fwd_rows_body<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, sel.ts, sel.tssh, p, tile, R, epi, blockIdx.x);
  }
}

// ------------------------------------------------------------------------------- inverse, rows
// First launch of the INTT: the inverse of the forward row transform, i.e. the standard S2-point
// Gentleman-Sande stages (first S2 entries of the inverse table, from LDS) followed by one
// coalesced multiply by tau_rho^-c. Input: canonical, output: internal form. fp64 values are
// re-centred after every second stage (Gentleman-Sande sums double per stage).
template <int LS, class A, class TwT>
__device__ __forceinline__ void inv_rows_body(const A& ar, const TwT* sw, const TwT* swp, const TwT* ts, const TwT* tsp, const uint64_t* p,
                                              uint64_t* po, uint64_t* tile, int R, uint32_t bx) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD;
  const int tid = threadIdx.x;
  const int rr = tid / T, c0 = tid % T;
  const uint32_t rho = bx * R + rr;
  const uint64_t* row = p + (size_t)rho * S;
  uint64_t* rowo = po + (size_t)rho * S;
  const uint32_t rbase = rho * S;
  uint64_t* t = tile + rr * PADS;
  typename A::T x[16];
  if constexpr (LS > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) t[rsw(c0 + T * k)] = row[c0 + T * k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_canon(t[rsw(16 * c0 + k)]);
#pragma unroll
    for (int s = LS - 1; s >= 4; --s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * c0 + k) >> (LS - s));
        ar.inv(x[k], x[k + h], ar.tw(sw, swp, idx));
      }
      if (((LS - s) & 1) == 0) {
#pragma unroll
        for (int k = 0; k < 16; ++k) ar.recenter(x[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      ar.recenter(x[k]);
      t[rsw(16 * c0 + k)] = ar.to_mid(x[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(t[rsw(c0 + T * k)]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_canon(row[c0 + T * k]);
  }
#pragma unroll
  for (int s = 3; s >= 0; --s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      ar.inv(x[k], x[k + h], ar.tw(sw, swp, idx));
    }
    if ((s & 1) == 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) ar.recenter(x[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) rowo[c0 + T * k] = ar.to_mid(ar.mul(x[k], ar.tw(ts + rbase, tsp + rbase, c0 + T * k)));
}

template <int LS, int TH, class Addr, class AddrOut>
__global__ __launch_bounds__(TH) void k_ntt_inv_rows(Addr ad, AddrOut ao, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD, RMAX = TH / T;
  __shared__ uint64_t tile[RMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t *p, *po;
  int pid, pido;
  ad.at(blockIdx.y, tb.N, p, pid);
  ao.at(blockIdx.y, tb.N, po, pido);
  const Sel sel = select(tb, pid, true);
  const int R = blockDim.x / T;
  if (sel.fp) {
    double* swd = reinterpret_cast<double*>(sw);
    for (int i = threadIdx.x; i < S; i += blockDim.x) swd[i] = sel.twd[i];
    __syncthreads();
    inv_rows_body<LS>(sel.fa, (const double*)swd, (const double*)swd, sel.tsd, sel.tsd, p, po, tile, R, blockIdx.x);
  } else {
    for (int i = threadIdx.x; i < S; i += blockDim.x) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
    __syncthreads();
    inv_rows_body<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, sel.ts, sel.tssh, p, po, tile, R, blockIdx.x);
  }
}

// ------------------------------------------------------------------------------- inverse, columns
// Second launch of the INTT (first l1 stages, reversed) with the N^-1 scaling. Input: internal
// form, output: canonical residues.
template <int LS, class Ld = PlainLd, class Tid = PlainTid, class A, class TwT, class Fin>
__device__ __forceinline__ void inv_cols_body(const A& ar, const TwT* sw, const TwT* swp, uint64_t* p, uint64_t* tile, uint32_t N2,
                                              int C, const Fin& fin, uint32_t bx) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1;
  const int tid = Tid::get();
  const int cc = tid % C, r0 = tid / C;
  const uint32_t col = bx * C + cc;
  typename A::T x[16];
  uint64_t* t = tile + cc * PADS;
  if constexpr (LS > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(Ld::ld(p + (size_t)(16 * r0 + k) * N2 + col));
    __syncthreads();
#pragma unroll
    for (int s = LS - 1; s >= 4; --s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * r0 + k) >> (LS - s));
        ar.inv(x[k], x[k + h], ar.tw(sw, swp, idx));
      }
      if (((LS - s) & 1) == 0) {
#pragma unroll
        for (int k = 0; k < 16; ++k) ar.recenter(x[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      ar.recenter(x[k]);
      t[16 * r0 + k] = ar.to_mid(x[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(t[r0 + T * k]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = ar.from_mid(Ld::ld(p + (size_t)(r0 + T * k) * N2 + col));
    __syncthreads();
  }
#pragma unroll
  for (int s = 3; s >= 0; --s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      ar.inv(x[k], x[k + h], ar.tw(sw, swp, idx));
    }
    if ((s & 1) == 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) ar.recenter(x[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) p[(size_t)(r0 + T * k) * N2 + col] = fin(x[k]);
}

struct IntFin {
  IntA a;
  uint64_t s, ssh;
  __device__ __forceinline__ uint64_t operator()(uint64_t v) const {
    uint64_t r = shoup_lazy(v, s, ssh, a.q);
    return r >= a.q ? r - a.q : r;
  }
};
struct FpFin {
  FpA a;
  double s;
  __device__ __forceinline__ uint64_t operator()(double v) const { return a.canon(a.mulmod(v, s)); }
};

template <int LS, int TH, class Addr>
__global__ __launch_bounds__(TH) void k_ntt_inv_cols(Addr ad, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, CMAX = TH / T;
  __shared__ uint64_t tile[CMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t* p;
  int pid;
  ad.at(blockIdx.y, tb.N, p, pid);
  const Sel sel = select(tb, pid, true);
  const PrimeConst& P = tb.prime[pid];
  const int C = blockDim.x / T;
  const uint32_t N2 = tb.N >> LS;
  if (sel.fp) {
    double* swd = reinterpret_cast<double*>(sw);
    for (int i = threadIdx.x; i < S; i += blockDim.x) swd[i] = sel.twd[i];
    inv_cols_body<LS>(sel.fa, (const double*)swd, (const double*)swd, p, tile, N2, C, FpFin{sel.fa, P.ninv_d}, blockIdx.x);
  } else {
    for (int i = threadIdx.x; i < S; i += blockDim.x) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
    inv_cols_body<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, p, tile, N2, C, IntFin{sel.ia, P.ninv, P.ninv_sh}, blockIdx.x);
  }
}

// ------------------------------------------------------------------------------- team form
// One launch transforms every limb with a single HBM round trip. For N = S x S, a limb is split
// into W = S / C column tiles and W row tiles; a team of W workgroups owns one limb at a time:
// member m runs the first pass on tile m and stores the intermediate in the output limb, the team
// meets at a barrier, and member m finishes row tile m (forward; the inverse runs rows first).
// The intermediate (8N bytes per limb, a few teams per XCD) stays in the L2 of the team's XCD
// between the passes, so HBM sees only the input read and the output write.
//
// Members of a team are workgroups b = x + 8 (t W + m) (the same b mod 8). Workgroups are dispatched
// to the XCDs round-robin, so a team normally shares one XCD and its L2, the point of coherence of
// all of that XCD's CUs: a store acknowledged to the wave (s_waitcnt vmcnt(0)) is visible to every
// CU of the XCD, and the second pass reads with loads that skip the CU's L1 (L2Ld). Whether a team
// does share an XCD is checked, not assumed: every member records its XCC id (hardware register
// XCC_ID) at launch, and a team whose members differ synchronizes with agent-scope release /
// acquire fences instead (L2 write-back and invalidate), which is correct on any placement.
// The launch is cooperative (all workgroups resident), every barrier wait is bounded, and a wait
// that runs past its bound sets the error word and makes every team leave the launch.
constexpr uint32_t kTeamSpinLimit = 1u << 20;

struct TeamArgs {
  uint32_t* ctr;      // [teams][3] arrival counters (placement, two limb slots), zero at launch
  uint32_t* xcc;      // [teams][W] XCC id of each member
  uint32_t* err;      // error word; err[1] counts teams whose members spanned XCDs
  uint32_t njobs, nteams, W;
};

__device__ __forceinline__ uint32_t xcc_id() { return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xF; }  // HW_REG_XCC_ID[3:0]

__device__ __forceinline__ uint32_t team_poll(uint32_t* p) { return __hip_atomic_fetch_add(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Arrival of this workgroup at a team counter (after its stores are visible to the team: the
// wave's stores acknowledged by the L2, or an agent-scope release when the team spans XCDs).
__device__ __forceinline__ void team_arrive(uint32_t* ctr, bool formal) {
  if (formal) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  else __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until the counter reaches target; false when the launch is being abandoned (a wait ran past
// its bound here or elsewhere).
__device__ __forceinline__ bool team_wait(const TeamArgs& ta, uint32_t* ctr, uint32_t target, bool formal, int* sflag) {
  if (threadIdx.x == 0) {
    int ok = 1;
    for (uint32_t spins = 0; team_poll(ctr) < target; ++spins) {
      if ((spins & 63) == 63 && (spins > kTeamSpinLimit || team_poll(ta.err))) {
        __hip_atomic_fetch_or(ta.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    *sflag = ok;
  }
  __syncthreads();
  const bool ok = *sflag != 0;
  __syncthreads();   // every thread has read the flag before it is written again
  if (formal) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok;
}

// A copy of v the compiler cannot compute before this point.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm volatile("" : "+s"(v));
  return v;
}

// Team identity of this workgroup and the launch-time placement check; returns false to leave.
// Counters: [team][3]: placement barrier, then the two alternating limb slots.
__device__ __forceinline__ bool team_setup(const TeamArgs& ta, uint32_t& team, uint32_t& m, bool& formal, int* sflag) {
  const uint32_t b = blockIdx.x, j = b / 8;
  team = (b % 8) + 8 * (j / ta.W);
  m = j % ta.W;
  if (threadIdx.x == 0) ta.xcc[team * ta.W + m] = xcc_id();
  team_arrive(ta.ctr + 3 * team, true);
  if (!team_wait(ta, ta.ctr + 3 * team, ta.W, true, sflag)) return false;
  if (threadIdx.x == 0) {
    int same = 1;
    const uint32_t x0 = ta.xcc[team * ta.W];
    for (uint32_t i = 1; i < ta.W; ++i) same &= ta.xcc[team * ta.W + i] == x0;
    *sflag = same;
    if (!same && m == 0) __hip_atomic_fetch_add(ta.err + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // diagnostic count
  }
  __syncthreads();
  formal = *sflag == 0;
  __syncthreads();
  return true;
}

// Each workgroup runs its limbs software-pipelined: the first pass of the team's next limb is done
// before waiting for the current limb's team barrier, so the barrier's latency hides behind a pass
// of work; two limbs per team are in flight (two alternating counter slots, targets W per use).
// One launch per arithmetic (FP: the fp64 path of the 50-bit primes, else integer Shoup): a team
// skips the limbs of the other kind, and each kernel holds one path's registers only.
template <int LS, int TH, bool FP, class Addr, class AddrOut, class Epi>
__global__ __launch_bounds__(TH) __attribute__((amdgpu_waves_per_eu(3))) void k_ntt_fwd_team(Addr ad, AddrOut ao, NttTables tb, Epi epi, TeamArgs ta) {
  constexpr int S = 1 << LS, T = S / 16, CMAX = TH / T;
  constexpr int TW = CMAX * RowTile<LS>::PAD > CMAX * (S + 1) ? CMAX * RowTile<LS>::PAD : CMAX * (S + 1);
  __shared__ uint64_t tile[TW];
  __shared__ uint64_t sw[S], swp[S];
  __shared__ int sflag;
  uint32_t team, m;
  bool formal;
  if (!team_setup(ta, team, m, formal, &sflag)) return;
  const uint32_t N2 = tb.N >> LS;
  const int C = (int)(blockDim.x / T);
  auto mine = [&](uint32_t y) {
    uint64_t* p;
    int pid;
    ad.at(y, tb.N, p, pid);
    return (tb.prime[pid].fp64 != 0) == FP;
  };
  auto next = [&](uint32_t y) {   // the team's next limb of this kernel's arithmetic after y
    for (y += ta.nteams; y < ta.njobs && !mine(y); y += ta.nteams) {}
    return y;
  };
  uint32_t* slot = ta.ctr + 3 * team + 1;
  uint32_t uses[2] = {0, 0};
  auto pass1 = [&](uint32_t y, uint32_t k) {
    uint64_t *p, *po;
    int pid, pido;
    ad.at(y, tb.N, p, pid);
    ao.at(y, tb.N, po, pido);
    const Sel sel = select(tb, pid, false);
    __syncthreads();   // earlier readers of sw / tile are done
    if constexpr (FP) {
      double* swd = reinterpret_cast<double*>(sw);
      for (int i = threadIdx.x; i < S; i += TH) swd[i] = sel.twd[i];
      __syncthreads();
      fwd_cols_body<LS, OpaqueTid>(sel.fa, (const double*)swd, (const double*)swd, p, po, tile, N2, C, opaque(m));
    } else {
      for (int i = threadIdx.x; i < S; i += TH) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
      __syncthreads();
      fwd_cols_body<LS, OpaqueTid>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, p, po, tile, N2, C, opaque(m));
    }
    team_arrive(slot + (k & 1), formal);
  };
  auto pass2 = [&](uint32_t y) {
    uint64_t *p, *po;
    int pid, pido;
    ad.at(y, tb.N, p, pid);
    ao.at(y, tb.N, po, pido);
    Epi e = epi;
    e.init(y, tb.N, tb.prime);
    const Sel sel = select(tb, pid, false);
    __syncthreads();
    if constexpr (FP) {
      double* swd = reinterpret_cast<double*>(sw);
      for (int i = threadIdx.x; i < S; i += TH) swd[i] = sel.twd[i];
      __syncthreads();
      fwd_rows_body<LS, L2Ld>(sel.fa, (const double*)swd, (const double*)swd, sel.tsd, sel.tsd, po, tile, C, e, opaque(m));
    } else {
      for (int i = threadIdx.x; i < S; i += TH) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
      __syncthreads();
      fwd_rows_body<LS, L2Ld>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, sel.ts, sel.tssh, po, tile, C, e, opaque(m));
    }
  };
  uint32_t y = team < ta.njobs && mine(team) ? team : next(team);
  uint32_t k = 0;
  if (y < ta.njobs) pass1(y, k);
  while (y < ta.njobs) {
    const uint32_t yn = next(y);
    if (yn < ta.njobs) pass1(yn, k + 1);
    if (!team_wait(ta, slot + (k & 1), ta.W * ++uses[k & 1], formal, &sflag)) return;
    pass2(y);
    y = yn;
    ++k;
  }
}

template <int LS, int TH, bool FP, class Addr, class AddrOut>
__global__ __launch_bounds__(TH) __attribute__((amdgpu_waves_per_eu(3))) void k_ntt_inv_team(Addr ad, AddrOut ao, NttTables tb, TeamArgs ta) {
  constexpr int S = 1 << LS, T = S / 16, CMAX = TH / T;
  constexpr int TW = CMAX * RowTile<LS>::PAD > CMAX * (S + 1) ? CMAX * RowTile<LS>::PAD : CMAX * (S + 1);
  __shared__ uint64_t tile[TW];
  __shared__ uint64_t sw[S], swp[S];
  __shared__ int sflag;
  uint32_t team, m;
  bool formal;
  if (!team_setup(ta, team, m, formal, &sflag)) return;
  const uint32_t N2 = tb.N >> LS;
  const int C = (int)(blockDim.x / T);
  auto mine = [&](uint32_t y) {
    uint64_t* p;
    int pid;
    ad.at(y, tb.N, p, pid);
    return (tb.prime[pid].fp64 != 0) == FP;
  };
  auto next = [&](uint32_t y) {
    for (y += ta.nteams; y < ta.njobs && !mine(y); y += ta.nteams) {}
    return y;
  };
  uint32_t* slot = ta.ctr + 3 * team + 1;
  uint32_t uses[2] = {0, 0};
  auto pass1 = [&](uint32_t y, uint32_t k) {
    uint64_t *p, *po;
    int pid, pido;
    ad.at(y, tb.N, p, pid);
    ao.at(y, tb.N, po, pido);
    const Sel sel = select(tb, pid, true);
    __syncthreads();
    if constexpr (FP) {
      double* swd = reinterpret_cast<double*>(sw);
      for (int i = threadIdx.x; i < S; i += TH) swd[i] = sel.twd[i];
      __syncthreads();
      inv_rows_body<LS>(sel.fa, (const double*)swd, (const double*)swd, sel.tsd, sel.tsd, p, po, tile, C, opaque(m));
    } else {
      for (int i = threadIdx.x; i < S; i += TH) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
      __syncthreads();
      inv_rows_body<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, sel.ts, sel.tssh, p, po, tile, C, opaque(m));
    }
    team_arrive(slot + (k & 1), formal);
  };
  auto pass2 = [&](uint32_t y) {
    uint64_t *p, *po;
    int pid, pido;
    ad.at(y, tb.N, p, pid);
    ao.at(y, tb.N, po, pido);
    const Sel sel = select(tb, pid, true);
    const PrimeConst& P = tb.prime[pid];
    __syncthreads();
    if constexpr (FP) {
      double* swd = reinterpret_cast<double*>(sw);
      for (int i = threadIdx.x; i < S; i += TH) swd[i] = sel.twd[i];
      __syncthreads();
      inv_cols_body<LS, L2Ld, OpaqueTid>(sel.fa, (const double*)swd, (const double*)swd, po, tile, N2, C, FpFin{sel.fa, P.ninv_d}, opaque(m));
    } else {
      for (int i = threadIdx.x; i < S; i += TH) { sw[i] = sel.tw[i]; swp[i] = sel.twsh[i]; }
      __syncthreads();
      inv_cols_body<LS, L2Ld, OpaqueTid>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, po, tile, N2, C, IntFin{sel.ia, P.ninv, P.ninv_sh},
                              opaque(m));
    }
  };
  uint32_t y = team < ta.njobs && mine(team) ? team : next(team);
  uint32_t k = 0;
  if (y < ta.njobs) pass1(y, k);
  while (y < ta.njobs) {
    const uint32_t yn = next(y);
    if (yn < ta.njobs) pass1(yn, k + 1);
    if (!team_wait(ta, slot + (k & 1), ta.W * ++uses[k & 1], formal, &sflag)) return;
    pass2(y);
    y = yn;
    ++k;
  }
}

// ------------------------------------------------------------------------------- launch
// A column tile is C whole columns (C = TH / (2^LS / 16), at most the number of columns); a row
// tile is R whole rows likewise. Both are powers of two dividing the column / row count.
template <int LS, int TH>
inline dim3 col_block(uint32_t N) {
  constexpr int T = (1 << LS) / 16;
  uint32_t N2 = N >> LS;
  uint32_t C = std::min<uint32_t>(TH / T, N2);
  return dim3(C * T);
}

template <int LS, int TH>
inline dim3 row_block(uint32_t N) {
  constexpr int T = (1 << LS) / 16;
  uint32_t S1 = N >> LS;  // number of rows
  uint32_t R = std::min<uint32_t>(TH / T, S1);
  return dim3(R * T);
}

// Offset addressing of a chunk of a larger job space (blockIdx.y + base).
template <class Addr>
struct Shift {
  Addr a;
  uint32_t base;
  __device__ __forceinline__ void at(uint32_t y, uint32_t N, uint64_t*& p, int& pid) const { a.at(y + base, N, p, pid); }
};

// Epilogues of a batch at mixed levels: one table entry per limb job (same index as the NTT job).
struct EpiModDownJobs {
  const EpiJob* jobs;
  const uint64_t* w;       // [P^-1]_{q_i} by prime, Shoup companions
  const uint64_t* wsh;
  const uint64_t* xi;
  uint64_t* oi;
  const uint64_t* ai;
  uint64_t q, wi, wish;
  __device__ __forceinline__ void init(uint32_t y, uint32_t, const PrimeConst* prime) {
    const EpiJob j = jobs[y];
    xi = j.x;
    oi = j.out;
    ai = j.add;
    q = prime[j.pid].q;
    wi = w[j.pid];
    wish = wsh[j.pid];
  }
  __device__ __forceinline__ void store(uint64_t*, uint32_t rbase, uint32_t jj, uint64_t v) const {
    uint64_t r = shoup(sub_mod(xi[rbase + jj], v, q), wi, wish, q);
    if (ai) r = add_mod(ai[rbase + jj], r, q);
    oi[rbase + jj] = r;
  }
};

// A job whose `add` is set rescales a ciphertext-plaintext product formed here: x_i = a_i pt_i,
// read from the ciphertext and the plaintext (job.x = a_i, job.add = pt_i) instead of from a
// materialized product (the ct x pt + Rescale fusion, src/NTTfusions.cuh NTT_MULTPT).
struct EpiRescaleJobs {
  const EpiJob* jobs;
  const uint64_t* w;       // [L1][L1] [q_l^-1]_{q_i}, Shoup companions
  const uint64_t* wsh;
  uint32_t L1;
  const uint64_t* xi;
  const uint64_t* pti;
  uint64_t* oi;
  uint64_t wi, wish;
  PrimeConst P;
  __device__ __forceinline__ void init(uint32_t y, uint32_t, const PrimeConst* prime) {
    const EpiJob j = jobs[y];
    xi = j.x;
    pti = j.add;
    oi = j.out;
    P = prime[j.pid];
    const size_t o = (size_t)j.l * L1 + j.pid;
    wi = w[o];
    wish = wsh[o];
  }
  __device__ __forceinline__ void store(uint64_t*, uint32_t rbase, uint32_t jj, uint64_t v) const {
    uint64_t x = xi[rbase + jj];
    if (pti) x = mul_mod(x, pti[rbase + jj], P);
    oi[rbase + jj] = add_mod(shoup(x, wi, wish, P.q), v, P.q);
  }
};

template <class Epi>
struct ShiftEpi {
  Epi e;
  uint32_t base;
  __device__ __forceinline__ void init(uint32_t y, uint32_t N, const PrimeConst* prime) { e.init(y + base, N, prime); }
  __device__ __forceinline__ void store(uint64_t* row, uint32_t rbase, uint32_t j, uint64_t v) const { e.store(row, rbase, j, v); }
};

// limbs per chunk: the intermediate of one chunk stays MALL-resident between the two launches
inline uint32_t chunk_limbs(const Context& c) {
  const size_t budget = (size_t)std::max(1, c.tune.ntt_chunk_mb) << 20;
  // at most 65535 limbs per launch (grid y limit)
  return (uint32_t)std::min<size_t>(65535, std::max<size_t>(1, budget / (8ull * c.N)));
}

// One launch of each pass kind, at workgroup size TH.
template <int LS, int TH, class Addr, class AddrOut>
void launch_fwd_cols(const Context& c, const Addr& a, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st) {
  const dim3 b = col_block<LS, TH>(c.N);
  const uint32_t C = b.x / ((1 << LS) / 16);
  ProfScope ps_(c, "ntt_fwd_cols", st);
  k_ntt_fwd_cols<LS, TH><<<dim3((c.N >> LS) / C, n), b, 0, st>>>(a, o, tb);
}
template <int LS, int TH, class AddrOut, class Epi>
void launch_fwd_rows(const Context& c, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st, const Epi& e, const char* name) {
  const dim3 b = row_block<LS, TH>(c.N);
  const uint32_t R = b.x / ((1 << LS) / 16);
  ProfScope ps_(c, name, st);
  k_ntt_fwd_rows<LS, TH><<<dim3((c.N >> LS) / R, n), b, 0, st>>>(o, tb, e);
}
template <int LS, int TH, class Addr, class AddrOut>
void launch_inv_rows(const Context& c, const Addr& a, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st) {
  const dim3 b = row_block<LS, TH>(c.N);
  const uint32_t R = b.x / ((1 << LS) / 16);
  ProfScope ps_(c, "ntt_inv_rows", st);
  k_ntt_inv_rows<LS, TH><<<dim3((c.N >> LS) / R, n), b, 0, st>>>(a, o, tb);
}
template <int LS, int TH, class AddrOut>
void launch_inv_cols(const Context& c, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st) {
  const dim3 b = col_block<LS, TH>(c.N);
  const uint32_t C = b.x / ((1 << LS) / 16);
  ProfScope ps_(c, "ntt_inv_cols", st);
  k_ntt_inv_cols<LS, TH><<<dim3((c.N >> LS) / C, n), b, 0, st>>>(o, tb);
}

// ---- team launches
__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// the team form applies to N = S x S with whole columns per tile (S >= 64), outside graph capture
// (the launch is cooperative)
template <int L1S, int L2S>
inline bool team_ok(const Context& c, hipStream_t st) {
  if constexpr (L1S != L2S || L1S < 6) return false;
  if (!c.tune.ntt_team) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  FH_HIP(hipStreamIsCapturing(st, &cs));
  return cs == hipStreamCaptureStatusNone;
}

// team count of a launch: a few teams per XCD (each holds one limb in flight in its L2), no more
// than the jobs need, and all of the launch's workgroups co-resident (cooperative launch)
inline bool team_check(const Context& c, uint32_t njobs, uint32_t W, int max_blocks_per_cu, TeamArgs& ta) {
  const uint32_t per_xcd = (uint32_t)std::max(1, c.tune.ntt_team_per_xcd);
  uint32_t nteams = std::min<uint32_t>(8 * per_xcd, (njobs + 7) / 8 * 8);
  const uint32_t resident = (uint32_t)std::max(0, max_blocks_per_cu) * (uint32_t)std::max(0, c.cu_count);
  nteams = std::min<uint32_t>(nteams, resident / W / 8 * 8);
  nteams = std::min<uint32_t>(nteams, Context::kMaxTeams);
  if (nteams < 8) return false;
  ta = TeamArgs{c.team_ctr(), c.team_xcc(), c.team_err(), njobs, nteams, W};
  return true;
}

// arrival counters of the launch back to zero (a fill kernel: capturable, unlike a memset node)
inline void team_reset(const Context& c, const TeamArgs& ta, hipStream_t st) {
  k_zero_u32<<<1, 256, 0, st>>>(c.team_ctr(), 3 * ta.nteams);
  FH_HIP(hipGetLastError());
}

inline const char* team_name(const char* rows_name) {
  if (!std::strcmp(rows_name, "ntt_fwd_rows_moddown")) return "ntt_fwd_team_moddown";
  if (!std::strcmp(rows_name, "ntt_fwd_rows_rescale")) return "ntt_fwd_team_rescale";
  return "ntt_fwd_team";
}

template <int LS, class Addr, class AddrOut, class Epi>
bool launch_fwd_team(const Context& c, const Addr& a, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st, const Epi& e,
                     const char* rows_name) {
  constexpr int TH = kNttThreads, T = (1 << LS) / 16, C = TH / T;
  constexpr uint32_t W = (1u << LS) / C;
  const void* fns[2] = {reinterpret_cast<const void*>(&k_ntt_fwd_team<LS, TH, true, Addr, AddrOut, Epi>),
                        reinterpret_cast<const void*>(&k_ntt_fwd_team<LS, TH, false, Addr, AddrOut, Epi>)};
  static int occ[2] = {-1, -1};
  TeamArgs ta[2];
  for (int k = 0; k < 2; ++k) {
    if (occ[k] < 0) FH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[k], fns[k], TH, 0));
    if (!team_check(c, n, W, occ[k], ta[k])) return false;
  }
  ProfScope ps_(c, team_name(rows_name), st);
  for (int k = 0; k < 2; ++k) {
    team_reset(c, ta[k], st);
    void* args[] = {(void*)&a, (void*)&o, (void*)&tb, (void*)&e, (void*)&ta[k]};
    FH_HIP(hipLaunchCooperativeKernel(fns[k], dim3(ta[k].nteams * W), dim3(TH), args, 0, st));
  }
  return true;
}

template <int LS, class Addr, class AddrOut>
bool launch_inv_team(const Context& c, const Addr& a, const AddrOut& o, uint32_t n, const NttTables& tb, hipStream_t st) {
  constexpr int TH = kNttThreads, T = (1 << LS) / 16, C = TH / T;
  constexpr uint32_t W = (1u << LS) / C;
  const void* fns[2] = {reinterpret_cast<const void*>(&k_ntt_inv_team<LS, TH, true, Addr, AddrOut>),
                        reinterpret_cast<const void*>(&k_ntt_inv_team<LS, TH, false, Addr, AddrOut>)};
  static int occ[2] = {-1, -1};
  TeamArgs ta[2];
  for (int k = 0; k < 2; ++k) {
    if (occ[k] < 0) FH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[k], fns[k], TH, 0));
    if (!team_check(c, n, W, occ[k], ta[k])) return false;
  }
  ProfScope ps_(c, "ntt_inv_team", st);
  for (int k = 0; k < 2; ++k) {
    team_reset(c, ta[k], st);
    void* args[] = {(void*)&a, (void*)&o, (void*)&tb, (void*)&ta[k]};
    FH_HIP(hipLaunchCooperativeKernel(fns[k], dim3(ta[k].nteams * W), dim3(TH), args, 0, st));
  }
  return true;
}

// WIDE: the wide (512-thread) tiles are built for this split (the default split of logN); the
// per-pass workgroup sizes come from the context's tunables.
// Chunks of a chunked transform alternate between the caller's stream and the context's second
// stream when the tunable ntt_streams is 2: a chunk's column pass then runs beside the previous
// chunk's row pass, so launch tails overlap and a chunk's intermediate can still be in the MALL
// when its second pass reads it. The second stream starts after everything issued before on the
// caller's stream, and the caller's stream waits for it at the end.
struct ChunkStreams {
  const Context& c;
  hipStream_t st;
  bool two;
  ChunkStreams(const Context& ctx, hipStream_t s, uint32_t chunk, uint32_t nlimbs)
      : c(ctx), st(s), two(ctx.tune.ntt_streams > 1 && chunk < nlimbs && ctx.aux_stream && s != ctx.aux_stream) {
    if (two) {
      FH_HIP(hipEventRecord(c.ev_fork, st));
      FH_HIP(hipStreamWaitEvent(c.aux_stream, c.ev_fork, 0));
    }
  }
  hipStream_t of(uint32_t k) const { return two && (k & 1) ? c.aux_stream : st; }
  void join() const {
    if (two) {
      FH_HIP(hipEventRecord(c.ev_join, c.aux_stream));
      FH_HIP(hipStreamWaitEvent(st, c.ev_join, 0));
    }
  }
};

template <class Addr, class AddrOut, int L1S, int L2S, bool WIDE, class Epi>
void fwd_impl(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st0, const Epi& epi, const char* rows_name) {
  NttTables tb = ntt_tables(c, L1S);
  if constexpr (L1S == L2S && L1S >= 6) {
    if (team_ok<L1S, L2S>(c, st0) && launch_fwd_team<L1S>(c, ad, ao, nlimbs, tb, st0, epi, rows_name)) return;
  }
  const bool wc = WIDE && c.tune.ntt_threads[0] == kNttWide, wr = WIDE && c.tune.ntt_threads[1] == kNttWide;
  const uint32_t chunk = chunk_limbs(c);
  const ChunkStreams cs(c, st0, chunk, nlimbs);
  for (uint32_t base = 0, k = 0; base < nlimbs; base += chunk, ++k) {
    const hipStream_t st = cs.of(k);
    uint32_t n = std::min(chunk, nlimbs - base);
    Shift<Addr> sa{ad, base};
    Shift<AddrOut> so{ao, base};
    ShiftEpi<Epi> se{epi, base};
    if constexpr (WIDE) {
      if (wc) launch_fwd_cols<L1S, kNttWide>(c, sa, so, n, tb, st);
      else launch_fwd_cols<L1S, kNttThreads>(c, sa, so, n, tb, st);
      if (wr) launch_fwd_rows<L2S, kNttWide>(c, so, n, tb, st, se, rows_name);
      else launch_fwd_rows<L2S, kNttThreads>(c, so, n, tb, st, se, rows_name);
    } else {
      launch_fwd_cols<L1S, kNttThreads>(c, sa, so, n, tb, st);
      launch_fwd_rows<L2S, kNttThreads>(c, so, n, tb, st, se, rows_name);
    }
  }
  cs.join();
}

template <class Addr, class AddrOut, int L1S, int L2S, bool WIDE>
void inv_impl(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st0) {
  NttTables tb = ntt_tables(c, L1S);
  if constexpr (L1S == L2S && L1S >= 6) {
    if (team_ok<L1S, L2S>(c, st0) && launch_inv_team<L1S>(c, ad, ao, nlimbs, tb, st0)) return;
  }
  const bool wr = WIDE && c.tune.ntt_threads[2] == kNttWide, wc = WIDE && c.tune.ntt_threads[3] == kNttWide;
  const uint32_t chunk = chunk_limbs(c);
  const ChunkStreams cs(c, st0, chunk, nlimbs);
  for (uint32_t base = 0, k = 0; base < nlimbs; base += chunk, ++k) {
    const hipStream_t st = cs.of(k);
    uint32_t n = std::min(chunk, nlimbs - base);
    Shift<Addr> sa{ad, base};
    Shift<AddrOut> so{ao, base};
    if constexpr (WIDE) {
      if (wr) launch_inv_rows<L2S, kNttWide>(c, sa, so, n, tb, st);
      else launch_inv_rows<L2S, kNttThreads>(c, sa, so, n, tb, st);
      if (wc) launch_inv_cols<L1S, kNttWide>(c, so, n, tb, st);
      else launch_inv_cols<L1S, kNttThreads>(c, so, n, tb, st);
    } else {
      launch_inv_rows<L2S, kNttThreads>(c, sa, so, n, tb, st);
      launch_inv_cols<L1S, kNttThreads>(c, so, n, tb, st);
    }
  }
  cs.join();
}

template <class Addr, class AddrOut, class Epi = EpiStore>
void dispatch(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st, bool fwd, const Epi& epi = Epi{},
              const char* rows_name = "ntt_fwd_rows") {
  if (nlimbs == 0) return;
  const int l1 = c.tune.ntt_split ? c.tune.ntt_split : ntt_default_split(c.logN);
  // (logN, l1) pairs accepted by ntt_split_ok; the default split of each logN also has wide tiles
#define FH_NTT_CASE(LN, A, WIDE)                                                                     \
  case LN * 16 + A:                                                                                  \
    if (fwd) fwd_impl<Addr, AddrOut, A, LN - A, WIDE, Epi>(c, ad, ao, nlimbs, st, epi, rows_name);   \
    else inv_impl<Addr, AddrOut, A, LN - A, WIDE>(c, ad, ao, nlimbs, st);                            \
    break;
  switch (c.logN * 16 + l1) {
    FH_NTT_CASE(10, 4, false)
    FH_NTT_CASE(10, 5, true)
    FH_NTT_CASE(10, 6, false)
    FH_NTT_CASE(11, 5, true)
    FH_NTT_CASE(11, 6, false)
    FH_NTT_CASE(12, 5, false)
    FH_NTT_CASE(12, 6, true)
    FH_NTT_CASE(12, 7, false)
    FH_NTT_CASE(13, 6, true)
    FH_NTT_CASE(13, 7, false)
    FH_NTT_CASE(14, 6, false)
    FH_NTT_CASE(14, 7, true)
    FH_NTT_CASE(14, 8, false)
    FH_NTT_CASE(15, 7, true)
    FH_NTT_CASE(15, 8, false)
    FH_NTT_CASE(16, 8, true)
    default:
      throw Error(-1, "unsupported ring dimension or radix split");
  }
#undef FH_NTT_CASE
  FH_HIP(hipGetLastError());
}

}  // namespace
}  // namespace fh
