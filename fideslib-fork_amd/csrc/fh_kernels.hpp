// Launch descriptors shared by the HIP kernels and their host launchers.
//
// A kernel addresses the limbs it works on through an Addr object passed by
// value in the kernel arguments (no per-call pointer-table upload):
//   SetAddr - npolys polynomials of nlimbs limbs each, equally strided, one prime
//             map shared by all of them (blockIdx.y = poly * nlimbs + limb);
//   JobAddr - a device-resident table of (limb pointer, prime) jobs, used for
//             batches of ciphertexts at different levels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fh_arith.hpp"
#include "fh_context.hpp"

namespace fh {

struct LimbSet {
  uint64_t* base;
  uint64_t poly_stride;  // words between consecutive polynomials
  uint32_t nlimbs;
  uint32_t npolys;
  uint8_t pid[kMaxLimbs];
};

struct LimbJob {
  uint64_t* ptr;
  uint32_t pid;
  uint32_t pad;
};

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

// automorphism gather index in bit-reversed EVALUATION order:
// out[j] = in[map(j)], map(brev(i)) = brev(((2i+1) k mod 2N) >> 1)   (OpenFHE PrecomputeAutoMap)
__device__ __forceinline__ uint32_t auto_src(uint32_t j, uint32_t k, int logN) {
  uint32_t i = brev_bits(j, logN);
  uint32_t e = ((2 * i + 1) * k) & ((2u << logN) - 1);
  return brev_bits(e >> 1, logN);
}

// Epilogue entry of a fused forward NTT over a job table (same index as the NTT job):
// ModDown: out = (x - v) [P^-1]_{q_pid} (+ add); Rescale: out = x [q_l^-1]_{q_pid} + v.
struct EpiJob {
  const uint64_t* x;
  uint64_t* out;
  const uint64_t* add;
  uint32_t pid;
  uint32_t l;
};

struct SetAddr {
  LimbSet s;
  // limb-major job order: consecutive blocks work on the same prime (its twiddles stay in L2)
  __device__ __forceinline__ void at(uint32_t y, uint32_t N, uint64_t*& p, int& pid) const {
    uint32_t limb = y / s.npolys, poly = y % s.npolys;
    p = s.base + (size_t)poly * s.poly_stride + (size_t)limb * N;
    pid = s.pid[limb];
  }
};

struct JobAddr {
  const LimbJob* jobs;
  __device__ __forceinline__ void at(uint32_t y, uint32_t, uint64_t*& p, int& pid) const {
    p = jobs[y].ptr;
    pid = (int)jobs[y].pid;
  }
};

// Target limbs of a ModUp output (npolys polynomials, each beta digits over Q_l ∪ P, ext = l+1+K
// limbs per digit). Jobs run prime-major: every (poly, digit) target of extended position 0, then
// position 1, ..., so consecutive workgroups share one prime's twiddle table in L2. A Q position
// is a target of every digit but its own (beta - 1 jobs per poly), a P position of all beta.
struct ModUpAddr {
  uint64_t* base;
  uint32_t ext, lq, alpha, L1, beta;
  uint32_t npolys;
  uint64_t poly_stride;   // words between consecutive polynomials' digit blocks
  uint32_t mix = 0;       // 1: spread the P-position jobs evenly among the Q-position jobs
  __device__ __forceinline__ void at(uint32_t y, uint32_t N, uint64_t*& p, int& pid) const {
    const uint32_t qjobs = npolys * (beta - 1);  // jobs per Q position
    const uint32_t y1 = qjobs * lq;
    if (mix) {
      // position y holds a P job iff the proportional P count steps between y and y + 1
      const uint64_t total = (uint64_t)npolys * (beta * ext - lq), nb = total - y1;
      const uint32_t kb = (uint32_t)(((uint64_t)y * nb) / total), kb1 = (uint32_t)(((uint64_t)(y + 1) * nb) / total);
      y = kb1 > kb ? y1 + kb : y - kb;
    }
    uint32_t pos, r, d;
    if (y < y1) {
      pos = y / qjobs;
      r = y - pos * qjobs;
      const uint32_t b = r / (beta - 1), dd = r - b * (beta - 1), own = pos / alpha;
      d = dd < own ? dd : dd + 1;
      p = base + b * poly_stride + ((size_t)d * ext + pos) * N;
    } else {
      const uint32_t pj = npolys * beta;           // jobs per P position
      const uint32_t t = (y - y1) / pj;
      pos = lq + t;
      r = (y - y1) - t * pj;
      const uint32_t b = r / beta;
      d = r - b * beta;
      p = base + b * poly_stride + ((size_t)d * ext + pos) * N;
    }
    pid = (int)(pos < lq ? pos : L1 + (pos - lq));
  }
  __host__ uint32_t count() const { return npolys * (beta * ext - lq); }
};

struct NttTables {
  const PrimeConst* prime;
  const uint64_t* tw;
  const uint64_t* tw_sh;
  const uint64_t* itw;
  const uint64_t* itw_sh;
  const double* twd;   // centred twiddles as doubles (fp64 path)
  const double* itwd;
  const uint64_t* twist;     // row twists [np][N] and their inverses (Shoup companions, doubles)
  const uint64_t* twist_sh;
  const uint64_t* itwist;
  const uint64_t* itwist_sh;
  const double* twistd;
  const double* itwistd;
  uint32_t N;
  uint32_t logN;
};

// tables of the two-pass NTT at radix split l1 (its row twists must have been built)
inline NttTables ntt_tables(const Context& c, int l1) {
  const Context::TwistSet& t = c.twists[l1];
  if (!t.tw) throw Error(-1, "row twists of this radix split are not built");
  return NttTables{c.d_prime, c.d_tw, c.d_tw_sh, c.d_itw, c.d_itw_sh, c.d_twd, c.d_itwd, t.tw, t.tw_sh,
                   t.itw, t.itw_sh, t.twd, t.itwd, c.N, (uint32_t)c.logN};
}

// NTT radix splits N = 2^l1 (column points) x 2^(logN - l1) (row points) the kernels are built
// for: both factors in [2^4, 2^8] and within a factor 4 of each other; floor(logN / 2) is the
// default.
inline int ntt_default_split(int logN) { return logN / 2; }
inline bool ntt_split_ok(int logN, int l1) {
  const int l2 = logN - l1;
  return l1 >= 4 && l1 <= 8 && l2 >= 4 && l2 <= 8 && (l1 - l2 <= 2) && (l2 - l1 <= 2);
}

inline LimbSet make_set(uint64_t* base, uint32_t nlimbs, const int* pids, uint32_t npolys = 1, uint64_t poly_stride = 0) {
  if (nlimbs > (uint32_t)kMaxLimbs) throw Error(-1, "too many limbs for one launch descriptor");
  LimbSet s{};
  s.base = base;
  s.nlimbs = nlimbs;
  s.npolys = npolys;
  s.poly_stride = poly_stride;
  for (uint32_t i = 0; i < nlimbs; ++i) s.pid[i] = (uint8_t)pids[i];
  return s;
}

// Launch KERNEL<SZ, M, RN> with the smallest compiled SZ >= n (n <= 16): the register-resident
// basis conversions are instantiated for source counts 2, 4, 8, 9, 12 and 16.
#define FH_SZ_DISPATCH(n, KERNEL, M, RN, LAUNCH, ARGS)                   \
  {                                                                     \
    const int n_ = (n);                                                 \
    if (n_ <= 2) KERNEL<2, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;           \
    else if (n_ <= 4) KERNEL<4, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;      \
    else if (n_ <= 8) KERNEL<8, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;      \
    else if (n_ <= 9) KERNEL<9, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;      \
    else if (n_ <= 12) KERNEL<12, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;    \
    else KERNEL<16, M, RN><<<FH_UNPACK LAUNCH>>> ARGS;                  \
  }
#define FH_UNPACK(...) __VA_ARGS__
// Pick the multiply-accumulate policy and the ReducedNoise variant of a conversion kernel.
#define FH_CONV_DISPATCH(ctx, n, KERNEL, LAUNCH, ARGS)                                          \
  {                                                                                            \
    if ((ctx).mac30_ok()) {                                                                    \
      if ((ctx).reduced_noise) FH_SZ_DISPATCH(n, KERNEL, Mac30, true, LAUNCH, ARGS)            \
      else FH_SZ_DISPATCH(n, KERNEL, Mac30, false, LAUNCH, ARGS)                               \
    } else {                                                                                   \
      if ((ctx).reduced_noise) FH_SZ_DISPATCH(n, KERNEL, MacWide, true, LAUNCH, ARGS)          \
      else FH_SZ_DISPATCH(n, KERNEL, MacWide, false, LAUNCH, ARGS)                             \
    }                                                                                          \
  }

// ---- launchers (fh_ntt.hip) ----
void ntt_fwd(const Context& c, const LimbSet& s, hipStream_t st);
void ntt_inv(const Context& c, const LimbSet& s, hipStream_t st);
void ntt_fwd(const Context& c, const LimbSet& in, const LimbSet& out, hipStream_t st);  // out-of-place
void ntt_inv(const Context& c, const LimbSet& in, const LimbSet& out, hipStream_t st);
void ntt_fwd_modup(const Context& c, const ModUpAddr& a, uint32_t n, hipStream_t st);
// the column pass alone over the ModUp targets: their row stages then run in the fused key inner
// product (rows_ksk, fh_rows_ksk.hip)
void ntt_fwd_modup_cols(const Context& c, const ModUpAddr& a, uint32_t n, hipStream_t st);

// NTT launch-shape autotuner: times forward + inverse transforms of `polys` synthetic
// polynomials over every prime of the context under candidate radix splits, per-pass workgroup
// sizes and limbs per launch pair, keeps the fastest in c.tune and reports it.
struct AutotuneResult {
  int split;            // log2 of the column-pass length
  int threads[4];       // forward cols, forward rows, inverse rows, inverse cols
  int chunk_mb;         // limb data per launch pair
  int candidates;       // configurations timed
  int limbs;            // limbs per timed transform
  double ms_default;    // forward + inverse at the default shape
  double ms_best;       // forward + inverse at the chosen shape
};
AutotuneResult autotune_ntt(Context& c, int polys);
void ntt_fwd_jobs(const Context& c, const LimbJob* d_jobs, uint32_t njobs, hipStream_t st);
// Forward NTT of the ModDown conversion (conv: npolys x lq limbs over Q_l, coefficient form, used as
// scratch) fused with the ModDown combination: out_i = (x_i - NTT(conv)_i) * [P^-1]_{q_i}.
void ntt_fwd_moddown(const Context& c, const LimbSet& conv, const uint64_t* x, uint64_t x_stride, uint64_t* out, uint64_t out_stride,
                     hipStream_t st);
// Forward NTT of the Rescale correction e (npolys x l limbs, coefficient form, scratch) fused with
// the Rescale combination: out_i = x_i * [q_l^-1]_{q_i} + NTT(e)_i.
void ntt_fwd_rescale(const Context& c, const LimbSet& e, const uint64_t* x, uint32_t l, uint64_t* out, hipStream_t st);
void ntt_inv_jobs(const Context& c, const LimbJob* d_jobs, uint32_t njobs, hipStream_t st);
// out-of-place job forms: job y reads in[y].ptr and writes out[y].ptr (same prime)
void ntt_fwd_jobs(const Context& c, const LimbJob* d_in, const LimbJob* d_out, uint32_t njobs, hipStream_t st);
void ntt_inv_jobs(const Context& c, const LimbJob* d_in, const LimbJob* d_out, uint32_t njobs, hipStream_t st);
// in-place forward NTTs of the job limbs (scratch) with the ModDown / Rescale epilogue of d_epi[y]
void ntt_fwd_jobs_moddown(const Context& c, const LimbJob* d_jobs, const EpiJob* d_epi, uint32_t njobs, hipStream_t st);
void ntt_fwd_jobs_rescale(const Context& c, const LimbJob* d_jobs, const EpiJob* d_epi, uint32_t njobs, hipStream_t st);

}  // namespace fh
