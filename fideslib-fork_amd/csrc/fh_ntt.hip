// Negacyclic NTT / INTT over 64-bit RNS limbs for gfx950.
//
// The transform is OpenFHE's: forward Cooley-Tukey with the bit-reversed table
// tw[k] = psi^brev(k) (stage m = 1..N/2 uses tw[m + i] for group i), output in
// bit-reversed EVALUATION order; inverse Gentleman-Sande with psi^-brev(k) and
// a final N^-1. Reference: src/NTT.cu:42-310 (INTT), 331-756 (NTT),
// src/CKKS/LimbPartition.cu:433-550 (launch), src/NTThelper.cuh:21-72 (butterflies).
//
// MI355X design: N = S1 * S2 (S1 = 2^l1 "column" points, S2 = 2^l2 "row"
// points). The first l1 stages only couple elements of the same column
// (stride S2), the last l2 stages only elements of the same row (contiguous),
// so a limb is transformed in two launches. Inside a launch each 64-wide
// wavefront lane owns 16 elements of one sub-transform: four radix-2 stages run
// in registers, one LDS exchange re-partitions the sub-transform, and the
// remaining stages run in registers again. Global loads and stores are
// coalesced in both passes (rows are staged through LDS where the register
// layout is not). Arithmetic is Harvey-lazy Shoup: forward values live in
// [0, 4q), inverse values in [0, 2q); the last stage of each transform
// returns canonical residues.
#include "fh_kernels.hpp"

namespace fh {

namespace {

__device__ __forceinline__ void fwd_bfly(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t wp, uint64_t q, uint64_t q2) {
  uint64_t x = X >= q2 ? X - q2 : X;        // [0, 2q)
  uint64_t t = shoup_lazy(Y, w, wp, q);     // [0, 2q)
  X = x + t;                                // [0, 4q)
  Y = x - t + q2;                           // [0, 4q)
}

__device__ __forceinline__ void inv_bfly(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t wp, uint64_t q, uint64_t q2) {
  uint64_t u = X + Y;                       // X, Y in [0, 2q)
  u = u >= q2 ? u - q2 : u;                 // [0, 2q)
  uint64_t v = X - Y + q2;                  // [0, 4q)
  X = u;
  Y = shoup_lazy(v, w, wp, q);              // [0, 2q)
}

__device__ __forceinline__ uint64_t canon4(uint64_t x, uint64_t q, uint64_t q2) {
  x = x >= q2 ? x - q2 : x;
  return x >= q ? x - q : x;
}

constexpr int kThreads = 256;

// ---------------------------------------------------------------- forward, columns (first l1 stages)
template <int LS, class Addr, class AddrOut>
__global__ __launch_bounds__(kThreads) void k_ntt_fwd_cols(Addr ad, AddrOut ao, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, CMAX = kThreads / T;
  __shared__ uint64_t tile[CMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t *p, *po;
  int pid, pido;
  ad.at(blockIdx.y, tb.N, p, pid);
  ao.at(blockIdx.y, tb.N, po, pido);
  const uint64_t q = tb.prime[pid].q, q2 = 2 * q;
  const uint64_t* tw = tb.tw + (size_t)pid * tb.N;
  const uint64_t* twp = tb.tw_sh + (size_t)pid * tb.N;
  const int tid = threadIdx.x, C = blockDim.x / T;
  for (int i = tid; i < S; i += blockDim.x) { sw[i] = tw[i]; swp[i] = twp[i]; }
  const uint32_t N2 = tb.N >> LS;
  const int cc = tid % C, r0 = tid / C;
  const uint32_t col = blockIdx.x * C + cc;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = p[(size_t)(r0 + T * k) * N2 + col];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      fwd_bfly(x[k], x[k + h], sw[idx], swp[idx], q, q2);
    }
  }
  if constexpr (LS > 4) {
    uint64_t* t = tile + cc * PADS;
#pragma unroll
    for (int k = 0; k < 16; ++k) t[r0 + T * k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[16 * r0 + k];
#pragma unroll
    for (int s = 4; s < LS; ++s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * r0 + k) >> (LS - s));
        fwd_bfly(x[k], x[k + h], sw[idx], swp[idx], q, q2);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) po[(size_t)(16 * r0 + k) * N2 + col] = x[k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) po[(size_t)(r0 + T * k) * N2 + col] = x[k];
  }
}

// ---------------------------------------------------------------- forward, rows (last l2 stages)
template <int LS, class Addr>
__global__ __launch_bounds__(kThreads) void k_ntt_fwd_rows(Addr ad, NttTables tb, int l1) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, RMAX = kThreads / T;
  __shared__ uint64_t tile[RMAX * PADS];
  uint64_t* p;
  int pid;
  ad.at(blockIdx.y, tb.N, p, pid);
  const uint64_t q = tb.prime[pid].q, q2 = 2 * q;
  const uint64_t* tw = tb.tw + (size_t)pid * tb.N;
  const uint64_t* twp = tb.tw_sh + (size_t)pid * tb.N;
  const int tid = threadIdx.x, R = blockDim.x / T;
  const int rr = tid / T, c0 = tid % T;
  const uint32_t rho = blockIdx.x * R + rr;
  uint64_t* row = p + (size_t)rho * S;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = row[c0 + T * k];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const uint32_t idx = (1u << (l1 + s)) + (rho << s) + (k >> (4 - s));
      fwd_bfly(x[k], x[k + h], tw[idx], twp[idx], q, q2);
    }
  }
  uint64_t* t = tile + rr * PADS;
  if constexpr (LS > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) t[c0 + T * k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[16 * c0 + k];
#pragma unroll
    for (int s = 4; s < LS; ++s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const uint32_t idx = (1u << (l1 + s)) + (rho << s) + ((16 * c0 + k) >> (LS - s));
        fwd_bfly(x[k], x[k + h], tw[idx], twp[idx], q, q2);
      }
    }
    // each thread rewrites exactly the LDS words it read: no barrier needed before the store
#pragma unroll
    for (int k = 0; k < 16; ++k) t[16 * c0 + k] = canon4(x[k], q, q2);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) row[c0 + T * k] = t[c0 + T * k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) row[c0 + T * k] = canon4(x[k], q, q2);
  }
}

// ---------------------------------------------------------------- inverse, rows (first: last l2 stages reversed)
template <int LS, class Addr, class AddrOut>
__global__ __launch_bounds__(kThreads) void k_ntt_inv_rows(Addr ad, AddrOut ao, NttTables tb, int l1) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, RMAX = kThreads / T;
  __shared__ uint64_t tile[RMAX * PADS];
  uint64_t *p, *po;
  int pid, pido;
  ad.at(blockIdx.y, tb.N, p, pid);
  ao.at(blockIdx.y, tb.N, po, pido);
  const uint64_t q = tb.prime[pid].q, q2 = 2 * q;
  const uint64_t* tw = tb.itw + (size_t)pid * tb.N;
  const uint64_t* twp = tb.itw_sh + (size_t)pid * tb.N;
  const int tid = threadIdx.x, R = blockDim.x / T;
  const int rr = tid / T, c0 = tid % T;
  const uint32_t rho = blockIdx.x * R + rr;
  const uint64_t* row = p + (size_t)rho * S;
  uint64_t* rowo = po + (size_t)rho * S;
  uint64_t* t = tile + rr * PADS;
  uint64_t x[16];
  if constexpr (LS > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) t[c0 + T * k] = row[c0 + T * k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[16 * c0 + k];
#pragma unroll
    for (int s = LS - 1; s >= 4; --s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const uint32_t idx = (1u << (l1 + s)) + (rho << s) + ((16 * c0 + k) >> (LS - s));
        inv_bfly(x[k], x[k + h], tw[idx], twp[idx], q, q2);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) t[16 * c0 + k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[c0 + T * k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = row[c0 + T * k];
  }
#pragma unroll
  for (int s = 3; s >= 0; --s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const uint32_t idx = (1u << (l1 + s)) + (rho << s) + (k >> (4 - s));
      inv_bfly(x[k], x[k + h], tw[idx], twp[idx], q, q2);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) rowo[c0 + T * k] = x[k];
}

// ---------------------------------------------------------------- inverse, columns (second: first l1 stages reversed, N^-1)
template <int LS, class Addr>
__global__ __launch_bounds__(kThreads) void k_ntt_inv_cols(Addr ad, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = S + 1, CMAX = kThreads / T;
  __shared__ uint64_t tile[CMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  uint64_t* p;
  int pid;
  ad.at(blockIdx.y, tb.N, p, pid);
  const PrimeConst& P = tb.prime[pid];
  const uint64_t q = P.q, q2 = 2 * q, ninv = P.ninv, ninv_sh = P.ninv_sh;
  const uint64_t* tw = tb.itw + (size_t)pid * tb.N;
  const uint64_t* twp = tb.itw_sh + (size_t)pid * tb.N;
  const int tid = threadIdx.x, C = blockDim.x / T;
  for (int i = tid; i < S; i += blockDim.x) { sw[i] = tw[i]; swp[i] = twp[i]; }
  const uint32_t N2 = tb.N >> LS;
  const int cc = tid % C, r0 = tid / C;
  const uint32_t col = blockIdx.x * C + cc;
  uint64_t x[16];
  uint64_t* t = tile + cc * PADS;
  if constexpr (LS > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = p[(size_t)(16 * r0 + k) * N2 + col];
    __syncthreads();
#pragma unroll
    for (int s = LS - 1; s >= 4; --s) {
      const int h = S >> (s + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & h) continue;
        const int idx = (1 << s) + ((16 * r0 + k) >> (LS - s));
        inv_bfly(x[k], x[k + h], sw[idx], swp[idx], q, q2);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) t[16 * r0 + k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[r0 + T * k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = p[(size_t)(r0 + T * k) * N2 + col];
    __syncthreads();
  }
#pragma unroll
  for (int s = 3; s >= 0; --s) {
    const int h = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & h) continue;
      const int idx = (1 << s) + (k >> (4 - s));
      inv_bfly(x[k], x[k + h], sw[idx], swp[idx], q, q2);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    uint64_t v = shoup_lazy(x[k], ninv, ninv_sh, q);
    p[(size_t)(r0 + T * k) * N2 + col] = v >= q ? v - q : v;
  }
}

template <int LS>
inline dim3 col_block(uint32_t N) {
  constexpr int T = (1 << LS) / 16;
  uint32_t N2 = N >> LS;
  uint32_t C = std::min<uint32_t>(kThreads / T, N2);
  return dim3(C * T);
}

template <int LS>
inline dim3 row_block(uint32_t N) {
  constexpr int T = (1 << LS) / 16;
  uint32_t S1 = N >> LS;  // number of rows
  uint32_t R = std::min<uint32_t>(kThreads / T, S1);
  return dim3(R * T);
}

template <class Addr, class AddrOut, int L1S, int L2S>
void fwd_impl(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st) {
  NttTables tb = ntt_tables(c);
  dim3 b1 = col_block<L1S>(c.N);
  uint32_t C = b1.x / ((1 << L1S) / 16);
  dim3 g1((c.N >> L1S) / C, nlimbs);
  k_ntt_fwd_cols<L1S, Addr, AddrOut><<<g1, b1, 0, st>>>(ad, ao, tb);
  dim3 b2 = row_block<L2S>(c.N);
  uint32_t R = b2.x / ((1 << L2S) / 16);
  dim3 g2((c.N >> L2S) / R, nlimbs);
  k_ntt_fwd_rows<L2S, AddrOut><<<g2, b2, 0, st>>>(ao, tb, L1S);
}

template <class Addr, class AddrOut, int L1S, int L2S>
void inv_impl(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st) {
  NttTables tb = ntt_tables(c);
  dim3 b2 = row_block<L2S>(c.N);
  uint32_t R = b2.x / ((1 << L2S) / 16);
  dim3 g2((c.N >> L2S) / R, nlimbs);
  k_ntt_inv_rows<L2S, Addr, AddrOut><<<g2, b2, 0, st>>>(ad, ao, tb, L1S);
  dim3 b1 = col_block<L1S>(c.N);
  uint32_t C = b1.x / ((1 << L1S) / 16);
  dim3 g1((c.N >> L1S) / C, nlimbs);
  k_ntt_inv_cols<L1S, AddrOut><<<g1, b1, 0, st>>>(ao, tb);
}

template <class Addr, class AddrOut>
void dispatch(const Context& c, const Addr& ad, const AddrOut& ao, uint32_t nlimbs, hipStream_t st, bool fwd) {
  if (nlimbs == 0) return;
  if (nlimbs > 65535u) throw Error(-1, "too many limbs in one NTT launch");
#define FH_NTT_CASE(LN, A, B)                                   \
  case LN:                                                      \
    if (fwd) fwd_impl<Addr, AddrOut, A, B>(c, ad, ao, nlimbs, st); \
    else inv_impl<Addr, AddrOut, A, B>(c, ad, ao, nlimbs, st);  \
    break;
  switch (c.logN) {
    FH_NTT_CASE(10, 5, 5)
    FH_NTT_CASE(11, 5, 6)
    FH_NTT_CASE(12, 6, 6)
    FH_NTT_CASE(13, 6, 7)
    FH_NTT_CASE(14, 7, 7)
    FH_NTT_CASE(15, 7, 8)
    FH_NTT_CASE(16, 8, 8)
    default:
      throw Error(-1, "unsupported ring dimension");
  }
#undef FH_NTT_CASE
  FH_HIP(hipGetLastError());
}

}  // namespace

void ntt_fwd(const Context& c, const LimbSet& s, hipStream_t st) { dispatch(c, SetAddr{s}, SetAddr{s}, s.nlimbs * s.npolys, st, true); }
void ntt_inv(const Context& c, const LimbSet& s, hipStream_t st) { dispatch(c, SetAddr{s}, SetAddr{s}, s.nlimbs * s.npolys, st, false); }
void ntt_fwd(const Context& c, const LimbSet& in, const LimbSet& out, hipStream_t st) {
  if (in.nlimbs * in.npolys != out.nlimbs * out.npolys) throw Error(-1, "NTT in/out shapes differ");
  dispatch(c, SetAddr{in}, SetAddr{out}, in.nlimbs * in.npolys, st, true);
}
void ntt_inv(const Context& c, const LimbSet& in, const LimbSet& out, hipStream_t st) {
  if (in.nlimbs * in.npolys != out.nlimbs * out.npolys) throw Error(-1, "NTT in/out shapes differ");
  dispatch(c, SetAddr{in}, SetAddr{out}, in.nlimbs * in.npolys, st, false);
}
void ntt_fwd_jobs(const Context& c, const LimbJob* j, uint32_t n, hipStream_t st) { dispatch(c, JobAddr{j}, JobAddr{j}, n, st, true); }
void ntt_inv_jobs(const Context& c, const LimbJob* j, uint32_t n, hipStream_t st) { dispatch(c, JobAddr{j}, JobAddr{j}, n, st, false); }
void ntt_fwd_modup(const Context& c, const ModUpAddr& a, uint32_t n, hipStream_t st) { dispatch(c, a, a, n, st, true); }

}  // namespace fh
