// CKKS encoding on the device: the special inverse FFT of the slot values in fp64, rounding to
// integer coefficients, their residues modulo each requested prime, and the forward NTT into the
// EVALUATION layout, for a batch of plaintexts in one pass. Only the slot values cross PCIe
// (16 B per slot instead of 8N B per limb).
// Reference: src/CKKS/DeviceEncodeBatch.cu:143-226 (fit_to_native_batch_, LaunchEncodeNtt),
// :435-529 (ifft_load_/ifft_stage_/ifft_finish_), :529-665 (EncodeChunkSlots), :665-860
// (loadCoefficientsBatch / loadSlotsBatch).
//
// The arithmetic is the host encoder's (special_ifft of the C++ facade, OpenFHE's FFTSpecialInv
// order): Gentleman-Sande stages len = n .. 2 with twiddle zeta^((4 len - 5^j mod 4 len) M / 4 len),
// bit reversal, division by n, multiplication by the scale, llround; every product and sum is
// rounded once (no contraction), so the device result equals the host result bit for bit.
#include <algorithm>
#include <cmath>

#include "fh_ops.hpp"

#pragma clang fp contract(off)

namespace fh {

namespace {

constexpr int kEncThreads = 256;
constexpr uint32_t kLdsElems = 2048;  // complex values per block in the shared-memory stages

struct D2 {
  double re, im;
};

__device__ __forceinline__ void butterfly(D2& a, D2& b, const double2 w) {
  const double ur = a.re + b.re, ui = a.im + b.im;
  const double dr = a.re - b.re, di = a.im - b.im;
  const double p0 = dr * w.x, p1 = di * w.y, p2 = dr * w.y, p3 = di * w.x;
  a = D2{ur, ui};
  b = D2{p0 - p1, p2 + p3};
}

// (lenq - (5^j mod lenq)) * (M / lenq) with lenq = 4 len = 2^(loglen + 2), M = 2^logM: the
// twiddle index of butterfly j in a stage of length len
__device__ __forceinline__ uint32_t tw_index(const uint32_t* rot, uint32_t j, int loglen, int logM) {
  const uint32_t lenq = 4u << loglen;
  return (lenq - (rot[j] & (lenq - 1))) << (logM - loglen - 2);
}

// values (count x slots complex, interleaved) -> work
__global__ __launch_bounds__(kEncThreads) void k_ifft_load(const double* __restrict__ vals, D2* __restrict__ work, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t o = (size_t)blockIdx.y * n + k;
  work[o] = D2{vals[2 * o], vals[2 * o + 1]};
}

// one stage of length len over every plaintext (blockIdx.y); one thread per butterfly
__global__ __launch_bounds__(kEncThreads) void k_ifft_stage(D2* __restrict__ work, const uint32_t* __restrict__ rot, const double2* __restrict__ w,
                                                            int logM, uint32_t n, int loglen) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (n >> 1)) return;
  const uint32_t lenh = 1u << (loglen - 1), j = t & (lenh - 1), i = (t >> (loglen - 1)) << loglen;
  D2* a = work + (size_t)blockIdx.y * n;
  D2 x = a[i + j], y = a[i + j + lenh];
  butterfly(x, y, w[tw_index(rot, j, loglen, logM)]);
  a[i + j] = x;
  a[i + j + lenh] = y;
}

// the stages len = top .. 2 of one block of `top` values in LDS (top <= kLdsElems)
__global__ __launch_bounds__(kEncThreads) void k_ifft_stages_lds(D2* __restrict__ work, const uint32_t* __restrict__ rot,
                                                                 const double2* __restrict__ w, int logM, uint32_t n, int logtop) {
  __shared__ D2 sh[kLdsElems];
  const uint32_t top = 1u << logtop;
  D2* a = work + (size_t)blockIdx.y * n + (size_t)blockIdx.x * top;
  for (uint32_t e = threadIdx.x; e < top; e += blockDim.x) sh[e] = a[e];
  __syncthreads();
  for (int loglen = logtop; loglen >= 1; --loglen) {
    const uint32_t lenh = 1u << (loglen - 1);
    for (uint32_t t = threadIdx.x; t < (top >> 1); t += blockDim.x) {
      const uint32_t j = t & (lenh - 1), i = (t >> (loglen - 1)) << loglen;
      D2 x = sh[i + j], y = sh[i + j + lenh];
      butterfly(x, y, w[tw_index(rot, j, loglen, logM)]);
      sh[i + j] = x;
      sh[i + j + lenh] = y;
    }
    __syncthreads();
  }
  for (uint32_t e = threadIdx.x; e < top; e += blockDim.x) a[e] = sh[e];
}

// coefficient j of plaintext `poly`: bit reversal, / n, * scale, llround, the gap scatter
// (gap = 2^loggap)
__device__ __forceinline__ int64_t coeff_at(const D2* work, uint32_t j, uint32_t N, uint32_t n, int logn, int loggap, double scale) {
  const uint32_t half = N >> 1;
  const bool im = j >= half;
  const uint32_t jj = im ? j - half : j;
  if (jj & ((1u << loggap) - 1)) return 0;
  const uint32_t k = jj >> loggap;
  const uint32_t r = logn ? brev_bits(k, logn) : 0u;
  const D2 x = work[r];
  const double v = (im ? x.im : x.re) / (double)n;
  return (int64_t)llround(v * scale);
}

__device__ __forceinline__ uint64_t signed_residue(int64_t v, const PrimeConst& P) {
  if (v >= 0) return barrett128(0, (uint64_t)v, P);
  uint64_t r = barrett128(0, (uint64_t)(-(v + 1)), P) + 1;
  if (r == P.q) r = 0;
  return r ? P.q - r : 0;
}

// job y = (limb, plaintext) of the output set: the residues of the rounded coefficients
__global__ __launch_bounds__(kEncThreads) void k_encode_finish(const D2* __restrict__ work, LimbSet out, const PrimeConst* pr, uint32_t N, uint32_t n,
                                                               int logn, int loggap, double scale) {
  uint64_t* p;
  int pid;
  SetAddr{out}.at(blockIdx.y, N, p, pid);
  const uint32_t poly = blockIdx.y % out.npolys;
  const PrimeConst P = pr[pid];
  const D2* wk = work + (size_t)poly * n;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x)
    p[j] = signed_residue(coeff_at(wk, j, N, n, logn, loggap, scale), P);
}

// integer coefficients (count x N, device) -> residues of every limb of the output set
__global__ __launch_bounds__(kEncThreads) void k_encode_coeffs(const int64_t* __restrict__ m, LimbSet out, const PrimeConst* pr, uint32_t N) {
  uint64_t* p;
  int pid;
  SetAddr{out}.at(blockIdx.y, N, p, pid);
  const uint32_t poly = blockIdx.y % out.npolys;
  const PrimeConst P = pr[pid];
  const int64_t* mp = m + (size_t)poly * N;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) p[j] = signed_residue(mp[j], P);
}

inline dim3 enc_grid(uint32_t N, uint32_t jobs) {
  uint32_t gx = (N + kEncThreads - 1) / kEncThreads;
  if (gx > 64) gx = 64;
  return dim3(gx, jobs);
}

}  // namespace

// zeta^idx for idx < M = 2N and 5^j mod M for j < N/2, computed as the host encoder computes them
static void ensure_ifft_tables(Context& c) {
  if (c.d_ifft_w) return;
  const uint64_t M = 2ull * c.N;
  constexpr double kPi = 3.14159265358979323846;
  std::vector<double> w(2 * M);
  for (uint64_t idx = 0; idx < M; ++idx) {
    const double ang = 2.0 * kPi * (double)idx / (double)M;
    // one sincos call per angle, as the host encoder computes its twiddles (glibc's sincos and
    // separate sin / cos calls differ in the last bit for some angles)
    ::sincos(ang, &w[2 * idx + 1], &w[2 * idx]);
  }
  std::vector<uint32_t> rot(c.N / 2);
  uint64_t g = 1;
  for (uint32_t j = 0; j < c.N / 2; ++j) {
    rot[j] = (uint32_t)g;
    g = (g * 5) % M;
  }
  double* dw = nullptr;
  uint32_t* dr = nullptr;
  FH_HIP(hipMalloc(&dw, w.size() * sizeof(double)));
  c.owned.push_back(dw);
  FH_HIP(hipMalloc(&dr, rot.size() * sizeof(uint32_t)));
  c.owned.push_back(dr);
  FH_HIP(hipMemcpy(dw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(dr, rot.data(), rot.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  c.d_ifft_w = dw;
  c.d_ifft_rot = dr;
}

void encode_slots(Context& c, const double* values, int count, uint32_t slots, double scale, const LimbSet& out, hipStream_t st) {
  if (count < 1) return;
  if (slots < 1 || (slots & (slots - 1)) || slots > c.N / 2) throw Error(-1, "slots must be a power of two <= N/2");
  if ((uint32_t)count != out.npolys) throw Error(-1, "encode: output set and count differ");
  if (count > 65535 || (size_t)out.nlimbs * count > 65535) throw Error(-1, "encode: too many plaintexts in one call");
  if (!(scale > 0) || !std::isfinite(scale)) throw Error(-1, "encode: the scale must be positive and finite");
  // Every coefficient is at most max |z| in magnitude (the inverse transform averages, up to a
  // rounding error of a few ulps). Above 2^61 (OpenFHE's MAX_BITS_IN_WORD) OpenFHE's
  // CKKSPackedEncoding::Encode switches to an approximate path with different bits, and the
  // reference's device encoder rejects such coefficients (src/CKKS/DeviceEncodeBatch.cu:628-652);
  // this encoder rounds exactly, so it rejects every input that could reach 2^61: |re| + |im| times
  // the scale must stay below 2^61 (1 - 2^-40). One comparison per value also rejects NaN and infinity.
  const double lim = 0x1p61 * (1.0 - 0x1p-40) / scale;
  bool bad = false;
  for (size_t i = 0; i < (size_t)count * slots; ++i) bad |= !(std::fabs(values[2 * i]) + std::fabs(values[2 * i + 1]) < lim);
  if (bad) throw Error(-1, "encode: a slot value is not finite, or times the scale it reaches 2^61 (OpenFHE's MAX_BITS_IN_WORD)");
  ensure_ifft_tables(c);
  const uint32_t n = slots;
  int logn = 0;
  while ((1u << logn) < n) ++logn;
  const int logM = c.logN + 1, loggap = c.logN - 1 - logn;
  DBuf vals(c.pool, (size_t)count * n * 2), work(c.pool, (size_t)count * n * 2);
  const double2* tw = reinterpret_cast<const double2*>(c.d_ifft_w);
  // chunks of about 8 MB of slot values: the upload of one chunk (synchronous for pageable host
  // memory) overlaps the transforms of the previous one
  const int chunk = std::max(1, std::min(count, (int)((8u << 20) / (16u * n))));
  for (int b0 = 0; b0 < count; b0 += chunk) {
    const int nb = std::min(chunk, count - b0);
    double* v = reinterpret_cast<double*>(vals.p) + (size_t)b0 * n * 2;
    D2* wk = reinterpret_cast<D2*>(work.p) + (size_t)b0 * n;
    FH_HIP(hipMemcpyAsync(v, values + (size_t)b0 * n * 2, (size_t)nb * n * 2 * sizeof(double), hipMemcpyHostToDevice, st));
    LimbSet part = out;
    part.base = out.base + (size_t)b0 * out.poly_stride;
    part.npolys = (uint32_t)nb;
    ProfScope ps_(c, "encode", st);
    k_ifft_load<<<dim3((n + kEncThreads - 1) / kEncThreads, nb), kEncThreads, 0, st>>>(v, wk, n);
    FH_HIP(hipGetLastError());
    int loglen = logn;
    for (; (1u << loglen) > kLdsElems; --loglen) {
      k_ifft_stage<<<dim3((n / 2 + kEncThreads - 1) / kEncThreads, nb), kEncThreads, 0, st>>>(wk, c.d_ifft_rot, tw, logM, n, loglen);
      FH_HIP(hipGetLastError());
    }
    if (loglen >= 1) {
      k_ifft_stages_lds<<<dim3(n >> loglen, nb), kEncThreads, 0, st>>>(wk, c.d_ifft_rot, tw, logM, n, loglen);
      FH_HIP(hipGetLastError());
    }
    k_encode_finish<<<enc_grid(c.N, part.nlimbs * part.npolys), kEncThreads, 0, st>>>(wk, part, c.d_prime, c.N, n, logn, loggap, scale);
    FH_HIP(hipGetLastError());
  }
  ntt_fwd(c, out, st);
}

void encode_coeffs(Context& c, const int64_t* coeffs, int count, const LimbSet& out, hipStream_t st) {
  if (count < 1) return;
  if ((uint32_t)count != out.npolys) throw Error(-1, "encode: output set and count differ");
  if (count > 65535 || (size_t)out.nlimbs * count > 65535) throw Error(-1, "encode: too many plaintexts in one call");
  DBuf m(c.pool, (size_t)count * c.N);
  FH_HIP(hipMemcpyAsync(m.p, coeffs, (size_t)count * c.N * sizeof(int64_t), hipMemcpyHostToDevice, st));
  {
  ProfScope ps_(c, "encode", st);
  k_encode_coeffs<<<enc_grid(c.N, out.nlimbs * out.npolys), kEncThreads, 0, st>>>(reinterpret_cast<const int64_t*>(m.p), out, c.d_prime, c.N);
  FH_HIP(hipGetLastError());
  }
  ntt_fwd(c, out, st);
}

}  // namespace fh
