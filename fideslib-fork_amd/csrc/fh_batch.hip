// Mixed-level operation batches: job-table forms of the tensor product, addition, ModUp
// conversion, key inner product, ModDown conversion and Rescale expansion, and the planner
// that turns a list of EvalRotate / EvalMult / EvalAdd items into one launch sequence.
//
// Every stage launches once for the whole batch whatever the items' levels: each job table
// entry carries the pointers and level of one (item, limb) or (item, digit) unit, so a batch of
// 256 ciphertexts at 24 different levels fills the GPU as one launch per stage instead of
// hundreds of small launches. The arithmetic of each stage is that of the single-operation
// kernels (fh_keyswitch.hip), so results are bit-identical.
// Reference: api/engine/cuda/CudaEngine.cpp:656-702 (evalRotateMany), src/CKKS/Ciphertext.cpp:
// 1519-1599 (rotate_many), src/CKKS/ElemenwiseBatchKernels.cu:379-544 (batched KSK dot),
// src/CKKS/LimbPartitionMGPU.cu:669-892 (graph-captured key-switch chains).
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "fh_batch.hpp"

namespace fh {

namespace {

constexpr int kT = 256;
constexpr int kMaxDnumB = 8;

inline dim3 grid_tiles(uint32_t N, uint32_t y, uint32_t per_thread, uint32_t z = 1) {
  uint32_t gx = (N / per_thread + kT - 1) / kT;
  return dim3(std::max(1u, std::min(gx, 256u)), y, z);
}

// ------------------------------------------------------------------------------ element-wise
struct TensorJob {
  const uint64_t *a0, *a1, *b0, *b1;
  uint64_t *o0, *o1, *o2;
  uint32_t pid, pad;
};

// (a0 b0, a0 b1 + a1 b0, a1 b1) for one limb (src/CKKS/ElemenwiseBatchKernels.cu:1363-1400)
__global__ __launch_bounds__(kT) void k_tensor_jobs(const TensorJob* __restrict__ jobs, const PrimeConst* __restrict__ prime, uint32_t N) {
  const TensorJob j = jobs[blockIdx.y];
  const PrimeConst P = prime[j.pid];
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const uint64_t a0 = j.a0[n], a1 = j.a1[n], b0 = j.b0[n], b1 = j.b1[n];
    j.o0[n] = mul_mod(a0, b0, P);
    j.o1[n] = add_mod(mul_mod(a0, b1, P), mul_mod(a1, b0, P), P.q);
    j.o2[n] = mul_mod(a1, b1, P);
  }
}

struct AddJob {
  const uint64_t* a;
  const uint64_t* b;
  uint64_t* out;
  uint32_t pid, pad;
};

// out = a + b, or out = a when b is null (level reduction, the untouched component of AddPlain)
__global__ __launch_bounds__(kT) void k_add_jobs(const AddJob* __restrict__ jobs, const PrimeConst* __restrict__ prime, uint32_t N) {
  const AddJob j = jobs[blockIdx.y];
  const uint64_t q = prime[j.pid].q;
  if (j.b) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) j.out[n] = add_mod(j.a[n], j.b[n], q);
  } else {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) j.out[n] = j.a[n];
  }
}

struct MulPtJob {
  const uint64_t *a0, *a1, *pt;
  uint64_t *o0, *o1;
  uint32_t pid, pad;
};

// (a0 pt, a1 pt) for one limb: ciphertext times plaintext (src/CKKS/LimbPartition.cu:588-682 multElement)
__global__ __launch_bounds__(kT) void k_mulpt_jobs(const MulPtJob* __restrict__ jobs, const PrimeConst* __restrict__ prime, uint32_t N) {
  const MulPtJob j = jobs[blockIdx.y];
  const PrimeConst P = prime[j.pid];
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const uint64_t w = j.pt[n];
    j.o0[n] = mul_mod(j.a0[n], w, P);
    j.o1[n] = mul_mod(j.a1[n], w, P);
  }
}

// ------------------------------------------------------------------------------ ModUp conversion
struct ModUpJob {
  const uint64_t* coef;   // INTT of the key-switched polynomial (l+1 limbs)
  const uint64_t* eval;   // the polynomial itself (l+1 limbs), copied to the digit's own limbs; null: not copied
  uint64_t* out;          // digit d's block: l+1+K limbs
  uint32_t l, d;
};

struct ConvConst {
  uint32_t N, L1, K, alpha, dnum;
  const PrimeConst* prime;
  const uint64_t* qhinv;     // [L1][dnum][alpha]
  const uint64_t* qhinv_sh;
  const uint64_t* qhat_m;    // [L1][dnum][alpha][L1+K], Montgomery form
  const uint64_t* phinv;     // [K]
  const uint64_t* phinv_sh;
  const uint64_t* phat_m;    // [K][L1], Montgomery form
  const uint64_t* negq_m;    // ReducedNoise corrections (null = off), as in fh_keyswitch.hip
  const uint64_t* negp_m;
};

// k_modup_conv_r of fh_keyswitch.hip with the level and digit read from the job
template <int SZ, class M, bool RN>
__global__ __launch_bounds__(kT) void k_modup_conv_jobs(const ModUpJob* __restrict__ jobs, ConvConst cc) {
  using Op = typename M::Op;
  __shared__ Op s_qhat[SZ * kMaxPrimes];
  __shared__ uint64_t s_pq[kMaxPrimes], s_pn[kMaxPrimes];
  __shared__ uint64_t inv[SZ], invsh[SZ], qs[SZ];
  const ModUpJob j = jobs[blockIdx.y];
  const uint32_t N = cc.N, lq = j.l + 1, ext = lq + cc.K, E = cc.L1 + cc.K;
  const uint32_t start = j.d * cc.alpha;
  const uint32_t sz = min(cc.alpha, lq - start);
  const size_t tb = ((size_t)j.l * cc.dnum + j.d) * cc.alpha;
  for (uint32_t i = threadIdx.x; i < SZ * ext; i += blockDim.x) {
    const uint32_t si = i / ext, pos = i % ext;
    s_qhat[si * ext + pos] = si < sz ? M::prep(cc.qhat_m[(tb + si) * E + pos]) : M::zero_op();
  }
  __shared__ Op s_negq[RN ? kMaxPrimes : 1];
  for (uint32_t pos = threadIdx.x; pos < ext; pos += blockDim.x) {
    const PrimeConst& P = cc.prime[pos < lq ? pos : cc.L1 + (pos - lq)];
    s_pq[pos] = P.q;
    s_pn[pos] = P.mont_ninv;
    if constexpr (RN) s_negq[pos] = M::prep(cc.negq_m[((size_t)j.l * cc.dnum + j.d) * E + pos]);
  }
  for (uint32_t i = threadIdx.x; i < sz; i += blockDim.x) {
    inv[i] = cc.qhinv[tb + i];
    invsh[i] = cc.qhinv_sh[tb + i];
    qs[i] = cc.prime[start + i].q;
  }
  __syncthreads();
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < N; n += 2 * gridDim.x * blockDim.x) {
    Op y0[SZ], y1[SZ];
    uint32_t lift0 = 0, lift1 = 0;
#pragma unroll
    for (int i = 0; i < SZ; ++i) {
      if (i < (int)sz) {
        const size_t o = (size_t)(start + i) * N + n;
        const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(j.coef + o);
        const uint64_t v0 = shoup(c.x, inv[i], invsh[i], qs[i]), v1 = shoup(c.y, inv[i], invsh[i], qs[i]);
        if constexpr (RN) {
          lift0 += v0 > (qs[i] >> 1);
          lift1 += v1 > (qs[i] >> 1);
        }
        y0[i] = M::prep(v0);
        y1[i] = M::prep(v1);
        if (j.eval) *reinterpret_cast<ulonglong2*>(j.out + o) = *reinterpret_cast<const ulonglong2*>(j.eval + o);
      } else {
        y0[i] = M::zero_op();
        y1[i] = M::zero_op();
      }
    }
    const Op l0 = M::prep(lift0), l1 = M::prep(lift1);
    for (uint32_t pos = 0; pos < ext; ++pos) {
      if (pos == start) {
        pos += sz - 1;
        continue;
      }
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int i = 0; i < SZ; ++i) {
        const Op c = s_qhat[i * ext + pos];
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
      *reinterpret_cast<ulonglong2*>(j.out + (size_t)pos * N + n) = make_ulonglong2(redc(h0, l0, pq, pn), redc(h1, l1, pq, pn));
    }
  }
}

// ------------------------------------------------------------------------------ key inner product
struct KskItem {
  const uint64_t* dig;    // beta x (l+1+K) limbs
  const uint64_t* c0;     // rotation: c0 over Q_l, folded as P c0; null otherwise
  uint64_t* ext;          // 2 x (l+1+K) limbs
  const uint64_t* own;    // the key-switched polynomial (EVAL, l+1 limbs): digit d's own limbs
  uint32_t l, pad;
};

struct KskGroup {
  const uint64_t* key;    // shared by the group's items
  uint32_t gal_inv;       // scatter form of the automorphism (1 = none)
  uint32_t first, count;
  uint32_t max_lq;        // largest l+1 in the group
};

struct KskConst {
  uint32_t N, L1, K, alpha, dnum, logN;
  const PrimeConst* prime;
  const uint64_t* pmod;
  const uint64_t* pmod_sh;
};

// blockIdx.y = prime index p (Q then P), blockIdx.z = group. The key words of (p, n) are loaded
// once and applied to every item of the group for which p is active.
template <class M>
__global__ __launch_bounds__(kT) void k_ksk_jobs(const KskItem* __restrict__ items, const KskGroup* __restrict__ groups, KskConst kc) {
  using Op = typename M::Op;
  const uint32_t p = blockIdx.y;
  const KskGroup g = groups[blockIdx.z];
  if (p < kc.L1 && p >= g.max_lq) return;
  const PrimeConst P = kc.prime[p];
  const uint32_t E = kc.L1 + kc.K;
  const size_t N = kc.N;
  const uint32_t bmax = min(kc.dnum, (g.max_lq + kc.alpha - 1) / kc.alpha);
  const bool qprime = p < kc.L1;
  const uint64_t pm = qprime ? kc.pmod[p] : 0, pmsh = qprime ? kc.pmod_sh[p] : 0;
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < kc.N; n += 2 * gridDim.x * blockDim.x) {
    Op kb0[kMaxDnumB], kb1[kMaxDnumB], ka0[kMaxDnumB], ka1[kMaxDnumB];
#pragma unroll
    for (uint32_t d = 0; d < kMaxDnumB; ++d) {
      if (d < bmax) {
        const ulonglong2 b = *reinterpret_cast<const ulonglong2*>(g.key + (((size_t)d * 2 + 0) * E + p) * N + n);
        const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(g.key + (((size_t)d * 2 + 1) * E + p) * N + n);
        kb0[d] = M::prep(b.x);
        kb1[d] = M::prep(b.y);
        ka0[d] = M::prep(a.x);
        ka1[d] = M::prep(a.y);
      }
    }
    const uint32_t j0 = g.gal_inv != 1 ? auto_src(n, g.gal_inv, kc.logN) : n;
    const uint32_t j1 = g.gal_inv != 1 ? auto_src(n + 1, g.gal_inv, kc.logN) : n + 1;
    for (uint32_t t = 0; t < g.count; ++t) {
      const KskItem it = items[g.first + t];
      const uint32_t lq = it.l + 1;
      if (qprime && p >= lq) continue;
      const uint32_t ext = lq + kc.K;
      const uint32_t pos = qprime ? p : lq + (p - kc.L1);
      const uint32_t beta = min(kc.dnum, (lq + kc.alpha - 1) / kc.alpha);
      const uint64_t* dg = it.dig + (size_t)pos * N + n;
      const uint32_t own_d = it.own && qprime ? pos / kc.alpha : ~0u;
      M a0, a1, b0, b1;
      a0.zero();
      a1.zero();
      b0.zero();
      b1.zero();
#pragma unroll
      for (uint32_t d = 0; d < kMaxDnumB; ++d) {
        if (d < beta) {
          const ulonglong2 x = *reinterpret_cast<const ulonglong2*>(d == own_d ? it.own + (size_t)pos * N + n : dg + (size_t)d * ext * N);
          const Op x0 = M::prep(x.x), x1 = M::prep(x.y);
          a0.mac(x0, kb0[d]);
          a1.mac(x0, ka0[d]);
          b0.mac(x1, kb1[d]);
          b1.mac(x1, ka1[d]);
        }
      }
      uint64_t h, l;
      a0.get(h, l);
      uint64_t r0 = redc(h, l, P);
      a1.get(h, l);
      uint64_t r1 = redc(h, l, P);
      b0.get(h, l);
      uint64_t s0 = redc(h, l, P);
      b1.get(h, l);
      uint64_t s1 = redc(h, l, P);
      if (it.c0 && qprime) {
        const ulonglong2 c = *reinterpret_cast<const ulonglong2*>(it.c0 + (size_t)pos * N + n);
        r0 = add_mod(r0, shoup(c.x, pm, pmsh, P.q), P.q);
        s0 = add_mod(s0, shoup(c.y, pm, pmsh, P.q), P.q);
      }
      uint64_t* o = it.ext;
      o[(size_t)pos * N + j0] = r0;
      o[((size_t)ext + pos) * N + j0] = r1;
      o[(size_t)pos * N + j1] = s0;
      o[((size_t)ext + pos) * N + j1] = s1;
    }
  }
}

// ------------------------------------------------------------------------------ ModDown conversion
struct MdJob {
  const uint64_t* pc;     // INTT of the K special limbs of one polynomial
  uint64_t* conv;         // l+1 limbs, coefficient form
  uint32_t lq, pad;
};

// k_moddown_conv_r of fh_keyswitch.hip with l+1 read from the job
template <int KK, class M, bool RN>
__global__ __launch_bounds__(kT) void k_moddown_conv_jobs(const MdJob* __restrict__ jobs, ConvConst cc) {
  using Op = typename M::Op;
  __shared__ Op s_phat[KK * kMaxPrimes];
  __shared__ uint64_t s_q[kMaxPrimes], s_n[kMaxPrimes];
  __shared__ uint64_t inv[KK], invsh[KK], ps[KK];
  const MdJob j = jobs[blockIdx.y];
  const uint32_t N = cc.N, lq = j.lq, K = cc.K;
  for (uint32_t i = threadIdx.x; i < KK * lq; i += blockDim.x) {
    const uint32_t k = i / lq, t = i % lq;
    s_phat[k * lq + t] = k < K ? M::prep(cc.phat_m[(size_t)k * cc.L1 + t]) : M::zero_op();
  }
  __shared__ Op s_negp[RN ? kMaxPrimes : 1];
  for (uint32_t i = threadIdx.x; i < lq; i += blockDim.x) {
    s_q[i] = cc.prime[i].q;
    s_n[i] = cc.prime[i].mont_ninv;
    if constexpr (RN) s_negp[i] = M::prep(cc.negp_m[i]);
  }
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
    inv[k] = cc.phinv[k];
    invsh[k] = cc.phinv_sh[k];
    ps[k] = cc.prime[cc.L1 + k].q;
  }
  __syncthreads();
  for (uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x); n < N; n += 2 * gridDim.x * blockDim.x) {
    Op y0[KK], y1[KK];
    uint32_t lift0 = 0, lift1 = 0;
#pragma unroll
    for (int k = 0; k < KK; ++k) {
      if (k < (int)K) {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(j.pc + (size_t)k * N + n);
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
    for (uint32_t i = 0; i < lq; ++i) {
      M a0, a1;
      a0.zero();
      a1.zero();
#pragma unroll
      for (int k = 0; k < KK; ++k) {
        const Op c = s_phat[k * lq + i];
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
      *reinterpret_cast<ulonglong2*>(j.conv + (size_t)i * N + n) = make_ulonglong2(redc(h0, l0, q, nv), redc(h1, l1, q, nv));
    }
  }
}

// ------------------------------------------------------------------------------ Rescale expansion
struct RsJob {
  const uint64_t* top;    // INTT of limb l of one component
  uint64_t* e;            // limb i of the correction
  uint32_t l, i;
};

// k_rescale_expand of fh_keyswitch.hip with (l, i) read from the job
__global__ __launch_bounds__(kT) void k_rescale_expand_jobs(const RsJob* __restrict__ jobs, const PrimeConst* __restrict__ prime,
                                                           const uint64_t* __restrict__ QlQl, const uint64_t* __restrict__ QlQl_sh,
                                                           uint32_t L1, uint32_t N) {
  const RsJob j = jobs[blockIdx.y];
  const uint64_t ql = prime[j.l].q, qi = prime[j.i].q;
  const uint64_t half = ql >> 1;
  const size_t o = (size_t)j.l * L1 + j.i;
  const uint64_t f = QlQl[o], fsh = QlQl_sh[o];
  const uint64_t diff = ql > qi ? qi - (ql % qi) : qi - ql;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    uint64_t v = j.top[n];
    if (v > half) v += diff;
    if (v >= qi) v %= qi;
    j.e[n] = shoup(v, f, fsh, qi);
  }
}

// Reorder NTT job lists prime-major (stable), permuting a companion list the same way:
// consecutive workgroups then share one prime's twiddle table in L2.
template <class T>
void sort_by_prime(std::vector<LimbJob>& jobs, std::vector<T>* comp) {
  std::vector<size_t> idx(jobs.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return jobs[a].pid < jobs[b].pid; });
  std::vector<LimbJob> j2(jobs.size());
  for (size_t i = 0; i < idx.size(); ++i) j2[i] = jobs[idx[i]];
  jobs.swap(j2);
  if (comp) {
    std::vector<T> c2(comp->size());
    for (size_t i = 0; i < idx.size(); ++i) c2[i] = (*comp)[idx[i]];
    comp->swap(c2);
  }
}

// Spread the jobs of integer-arithmetic (60-bit) primes evenly among the fp64 ones, keeping the
// prime-major order within each class (jobs, comp: as after sort_by_prime).
template <class T>
void mix_prime_classes(const Context& c, std::vector<LimbJob>& jobs, std::vector<T>* comp) {
  std::vector<size_t> a, b;
  for (size_t i = 0; i < jobs.size(); ++i) (c.fp64[jobs[i].pid] ? a : b).push_back(i);
  if (a.empty() || b.empty()) return;
  const size_t total = jobs.size(), nb = b.size();
  std::vector<size_t> idx(total);
  for (size_t y = 0; y < total; ++y) {
    const size_t kb = y * nb / total, kb1 = (y + 1) * nb / total;
    idx[y] = kb1 > kb ? b[kb] : a[y - kb];
  }
  std::vector<LimbJob> j2(total);
  for (size_t y = 0; y < total; ++y) j2[y] = jobs[idx[y]];
  jobs.swap(j2);
  if (comp) {
    std::vector<T> c2(total);
    for (size_t y = 0; y < total; ++y) c2[y] = (*comp)[idx[y]];
    comp->swap(c2);
  }
}

// byte-addressed table builder: sections aligned to 64 bytes, uploaded in one copy
struct TableBuilder {
  std::vector<uint8_t> bytes;
  template <class T>
  size_t put(const std::vector<T>& v) {
    size_t off = (bytes.size() + 63) & ~size_t(63);
    bytes.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

}  // namespace

// ===================================================================================== planner

struct BatchPlan {
  struct Rot {
    const uint64_t* ct;
    int l;
    const uint64_t* key;
    uint32_t galois;
    uint64_t* out;
  };
  struct Mult {
    const uint64_t *a, *b;
    int l;
    const uint64_t* key;
    bool rescale;
    uint64_t* out;
    uint32_t ta, tb;  // limbs per component of a and b (their first l+1 limbs are read)
  };
  struct Add {
    const uint64_t *a, *b;
    int l, ncomp;
    uint64_t* out;
  };
  struct MultPt {
    const uint64_t *a, *pt;
    int l;
    bool rescale;
    uint64_t* out;
  };
  struct AddPt {
    const uint64_t *a, *pt;
    int l;
    uint64_t* out;
  };
  struct Drop {
    const uint64_t* a;
    int l_in, l_out, ncomp;
    uint64_t* out;
  };
  std::vector<Rot> rots;
  std::vector<Mult> mults;
  std::vector<Add> adds;
  std::vector<MultPt> mulpts;
  std::vector<AddPt> addpts;
  std::vector<Drop> drops;
  bool final = false;

  // scratch and tables (kept across reset() and regrown on demand)
  DBuf arena;
  DBuf tables;  // device job tables (from the context pool: kept across sessions)
  void* d_tables = nullptr;
  size_t tables_cap = 0;
  size_t scratch_words = 0;
  const MulPtJob* t_mulpt = nullptr;
  uint32_t n_mulpt = 0;
  // table sections (device pointers) and sizes
  const TensorJob* t_tensor = nullptr;
  uint32_t n_tensor = 0;
  const AddJob* t_add = nullptr;
  uint32_t n_add = 0;
  const LimbJob *t_ks_in = nullptr, *t_ks_out = nullptr;
  uint32_t n_ks_intt = 0;
  const ModUpJob* t_modup = nullptr;
  uint32_t n_modup = 0;
  const LimbJob* t_modup_ntt = nullptr;
  uint32_t n_modup_ntt = 0;
  const KskItem* t_ksk_items = nullptr;
  const KskGroup* t_ksk_groups = nullptr;
  uint32_t n_ksk_groups = 0;
  const LimbJob *t_md_in = nullptr, *t_md_out = nullptr;
  uint32_t n_md_intt = 0;
  const MdJob* t_md = nullptr;
  uint32_t n_md = 0;
  const LimbJob* t_md_ntt = nullptr;
  const EpiJob* t_md_epi = nullptr;
  uint32_t n_md_ntt = 0;
  const LimbJob *t_rs_in = nullptr, *t_rs_out = nullptr;
  uint32_t n_rs_intt = 0;
  const RsJob* t_rs = nullptr;
  uint32_t n_rs = 0;
  const LimbJob* t_rs_ntt = nullptr;
  const EpiJob* t_rs_epi = nullptr;
  uint32_t n_rs_ntt = 0;

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;

  hipEvent_t uploaded = nullptr;  // after the last table upload (the context's staging ring)
  // a run on a stream other than the context stream: the next upload waits for it on the GPU
  hipEvent_t run_ev = nullptr;
  bool run_other = false;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_stream = nullptr;
  }
  ~BatchPlan() {
    drop_graph();
    if (run_ev) {
      if (run_other) (void)hipEventSynchronize(run_ev);
      (void)hipEventDestroy(run_ev);
    }
  }
};

Batch::Batch(Context& c) : c_(c), p_(new BatchPlan) {}
Batch::~Batch() = default;

size_t Batch::items() const {
  return p_->rots.size() + p_->mults.size() + p_->adds.size() + p_->mulpts.size() + p_->addpts.size() + p_->drops.size();
}

void Batch::reset() {
  BatchPlan& P = *p_;
  P.rots.clear();
  P.mults.clear();
  P.adds.clear();
  P.mulpts.clear();
  P.addpts.clear();
  P.drops.clear();
  P.drop_graph();
  P.final = false;
}
size_t Batch::scratch_bytes() const { return p_->scratch_words * 8; }

static void check_level(const Context& c, int l) {
  if (l < 0 || l >= c.L1) throw Error(-1, "level out of range");
}

void Batch::add_rotate(const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, uint64_t* out) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l);
  if (!ct || !key || !out) throw Error(-1, "null pointer in batch rotation");
  if ((galois & 1) == 0 || galois >= 2 * c_.N) throw Error(-1, "galois index must be odd and < 2N");
  p_->rots.push_back({ct, l, key, galois, out});
}

void Batch::add_mult(const uint64_t* a, const uint64_t* b, int l, const uint64_t* key, bool rescale, uint64_t* out, int towers_a,
                     int towers_b) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l);
  if (!a || !b || !key || !out) throw Error(-1, "null pointer in batch multiplication");
  if (rescale && l < 1) throw Error(-1, "cannot rescale a level-0 ciphertext");
  const int lq = l + 1;
  if (towers_a == 0) towers_a = lq;
  if (towers_b == 0) towers_b = lq;
  if (towers_a < lq || towers_b < lq || towers_a > c_.L1 || towers_b > c_.L1)
    throw Error(-1, "batch multiplication: operand towers must lie in [level + 1, number of Q primes]");
  p_->mults.push_back({a, b, l, key, rescale, out, (uint32_t)towers_a, (uint32_t)towers_b});
}

void Batch::add_add(const uint64_t* a, const uint64_t* b, int l, int ncomp, uint64_t* out) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l);
  if (!a || !b || !out) throw Error(-1, "null pointer in batch addition");
  if (ncomp < 1 || ncomp > 3) throw Error(-1, "ncomp must be 1, 2 or 3");
  p_->adds.push_back({a, b, l, ncomp, out});
}

void Batch::add_mult_plain(const uint64_t* a, const uint64_t* pt, int l, bool rescale, uint64_t* out) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l);
  if (!a || !pt || !out) throw Error(-1, "null pointer in batch plaintext multiplication");
  if (rescale && l < 1) throw Error(-1, "cannot rescale a level-0 ciphertext");
  p_->mulpts.push_back({a, pt, l, rescale, out});
}

void Batch::add_add_plain(const uint64_t* a, const uint64_t* pt, int l, uint64_t* out) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l);
  if (!a || !pt || !out) throw Error(-1, "null pointer in batch plaintext addition");
  p_->addpts.push_back({a, pt, l, out});
}

void Batch::add_level_reduce(const uint64_t* a, int l_in, int l_out, int ncomp, uint64_t* out) {
  if (p_->final) throw Error(-1, "batch already finalized");
  check_level(c_, l_in);
  if (l_out < 0 || l_out > l_in) throw Error(-1, "level reduction target out of range");
  if (!a || !out) throw Error(-1, "null pointer in batch level reduction");
  if (ncomp < 1 || ncomp > 3) throw Error(-1, "ncomp must be 1, 2 or 3");
  p_->drops.push_back({a, l_in, l_out, ncomp, out});
}

void Batch::finalize() {
  BatchPlan& P = *p_;
  if (P.final) return;
  Context& c = c_;
  const size_t N = c.N;
  const uint32_t K = c.K, L1 = c.L1;
  if (!P.rots.empty() || !P.mults.empty()) {
    if (K < 1) throw Error(-1, "key switching needs at least one special prime");
    if (!c.redc_conv_ok || c.alpha > 16 || K > 16) throw Error(-3, "batched key switching needs alpha, K <= 16");
    if (c.reduced_noise && !c.rn_conv_ok) throw Error(-3, "ReducedNoise conversions need (alpha + 1) * max prime <= 2^64");
    if (c.dnum > kMaxDnumB) throw Error(-3, "dnum above kernel limit");
  }

  // ---- key-switch items: rotations (c1, P c0 folded, automorphism) then multiplications (c2 of
  // the tensor, relinearization add fused into the ModDown epilogue) ----
  struct Ks {
    const uint64_t* x;
    const uint64_t* c0;
    const uint64_t* key;
    uint32_t galois;
    int l;
    const uint64_t* add0;   // added to the ModDown result (tensor c0, c1) or null
    const uint64_t* add1;
    uint64_t* out;          // 2 x (l+1) limbs
    // scratch offsets (words)
    size_t coef, dig, ext, pc, conv;
  };
  std::vector<Ks> ks;
  // scratch arena layout
  size_t off = 0;
  auto take = [&](size_t words) {
    size_t o = off;
    off += (words + 31) & ~size_t(31);  // 256-byte aligned sections
    return o;
  };
  struct MultScratch {
    size_t tensor, relin, top, e;
  };
  std::vector<MultScratch> ms(P.mults.size());
  for (size_t i = 0; i < P.mults.size(); ++i) {
    const auto& m = P.mults[i];
    const size_t lq = m.l + 1;
    ms[i].tensor = take(3 * lq * N);
    ms[i].relin = m.rescale ? take(2 * lq * N) : 0;
    ms[i].top = m.rescale ? take(2 * N) : 0;
    ms[i].e = m.rescale ? take(2 * (lq - 1) * N) : 0;
  }
  std::vector<MultScratch> ps(P.mulpts.size());
  for (size_t i = 0; i < P.mulpts.size(); ++i) {
    const auto& m = P.mulpts[i];
    const size_t lq = m.l + 1;
    ps[i].tensor = 0;
    ps[i].relin = m.rescale ? take(2 * lq * N) : 0;
    ps[i].top = m.rescale ? take(2 * N) : 0;
    ps[i].e = m.rescale ? take(2 * (lq - 1) * N) : 0;
  }
  for (const auto& r : P.rots) {
    const size_t lq = r.l + 1;
    ks.push_back(Ks{r.ct + lq * N, r.ct, r.key, r.galois, r.l, nullptr, nullptr, r.out, 0, 0, 0, 0, 0});
  }
  // tensor outputs and relinearization outputs are addressed after the arena is allocated
  const size_t first_mult_ks = ks.size();
  for (size_t i = 0; i < P.mults.size(); ++i) {
    const auto& m = P.mults[i];
    ks.push_back(Ks{nullptr, nullptr, m.key, 1, m.l, nullptr, nullptr, m.out, 0, 0, 0, 0, 0});
  }
  for (auto& k : ks) {
    const size_t lq = k.l + 1, ext = lq + K, beta = c.num_digits(k.l);
    k.coef = take(lq * N);
    k.dig = take(beta * ext * N);
    k.ext = take(2 * ext * N);
    k.pc = take(2 * K * N);
    k.conv = take(2 * lq * N);
  }
  P.scratch_words = off;
  if (P.arena.words < off || !P.arena.p) P.arena = DBuf(c.pool, std::max<size_t>(off + off / 4, 1));
  uint64_t* A = P.arena.p;

  // resolve multiplication pointers
  for (size_t i = 0; i < P.mults.size(); ++i) {
    const auto& m = P.mults[i];
    const size_t lq = m.l + 1;
    Ks& k = ks[first_mult_ks + i];
    uint64_t* t = A + ms[i].tensor;
    k.x = t + 2 * lq * N;
    k.add0 = t;
    k.add1 = t + lq * N;
    if (m.rescale) k.out = A + ms[i].relin;
  }

  // ---- job tables ----
  std::vector<TensorJob> tj;
  for (size_t i = 0; i < P.mults.size(); ++i) {
    const auto& m = P.mults[i];
    const size_t lq = m.l + 1;
    uint64_t* t = A + ms[i].tensor;
    for (size_t j = 0; j < lq; ++j)
      tj.push_back(TensorJob{m.a + j * N, m.a + (m.ta + j) * N, m.b + j * N, m.b + (m.tb + j) * N, t + j * N, t + (lq + j) * N,
                             t + (2 * lq + j) * N, (uint32_t)j, 0});
  }
  std::vector<AddJob> aj;
  for (const auto& a : P.adds) {
    const size_t lq = a.l + 1;
    for (int comp = 0; comp < a.ncomp; ++comp)
      for (size_t j = 0; j < lq; ++j) {
        const size_t o = (comp * lq + j) * N;
        aj.push_back(AddJob{a.a + o, a.b + o, a.out + o, (uint32_t)j, 0});
      }
  }
  for (const auto& a : P.addpts) {
    const size_t lq = a.l + 1;
    for (size_t j = 0; j < lq; ++j) {
      aj.push_back(AddJob{a.a + j * N, a.pt + j * N, a.out + j * N, (uint32_t)j, 0});
      aj.push_back(AddJob{a.a + (lq + j) * N, nullptr, a.out + (lq + j) * N, (uint32_t)j, 0});
    }
  }
  for (const auto& d : P.drops) {
    const size_t li = d.l_in + 1, lo = d.l_out + 1;
    for (int comp = 0; comp < d.ncomp; ++comp)
      for (size_t j = 0; j < lo; ++j) aj.push_back(AddJob{d.a + (comp * li + j) * N, nullptr, d.out + (comp * lo + j) * N, (uint32_t)j, 0});
  }
  // A plaintext product followed by a Rescale forms only its top limb here (the INTT's input);
  // the Rescale epilogue forms the other limbs' products as it reads them (unless the output
  // overlaps the input, when the whole product is materialized first).
  auto fused_mulpt = [&](const BatchPlan::MultPt& m) {
    if (!m.rescale) return false;
    const size_t lq = m.l + 1;
    const uint64_t *a0 = m.a, *a1 = m.a + 2 * lq * N, *o0 = m.out, *o1 = m.out + 2 * (lq - 1) * N;
    return o1 <= a0 || a1 <= o0;
  };
  std::vector<MulPtJob> pj;
  for (size_t i = 0; i < P.mulpts.size(); ++i) {
    const auto& m = P.mulpts[i];
    const size_t lq = m.l + 1;
    uint64_t* o = m.rescale ? A + ps[i].relin : m.out;
    for (size_t j = fused_mulpt(m) ? lq - 1 : 0; j < lq; ++j)
      pj.push_back(MulPtJob{m.a + j * N, m.a + (lq + j) * N, m.pt + j * N, o + j * N, o + (lq + j) * N, (uint32_t)j, 0});
  }
  std::vector<LimbJob> ks_in, ks_out, modup_ntt, md_in, md_out, md_ntt, rs_in, rs_out, rs_ntt;
  std::vector<ModUpJob> mj;
  std::vector<MdJob> mdj;
  std::vector<EpiJob> md_epi, rs_epi;
  std::vector<RsJob> rj;
  for (const auto& k : ks) {
    const uint32_t lq = k.l + 1, ext = lq + K, beta = c.num_digits(k.l);
    uint64_t* coef = A + k.coef;
    uint64_t* dig = A + k.dig;
    uint64_t* ex = A + k.ext;
    uint64_t* pc = A + k.pc;
    uint64_t* conv = A + k.conv;
    for (uint32_t j = 0; j < lq; ++j) {
      ks_in.push_back(LimbJob{const_cast<uint64_t*>(k.x) + (size_t)j * N, j, 0});
      ks_out.push_back(LimbJob{coef + (size_t)j * N, j, 0});
    }
    for (uint32_t d = 0; d < beta; ++d) {
      mj.push_back(ModUpJob{coef, nullptr, dig + (size_t)d * ext * N, (uint32_t)k.l, d});
      const uint32_t start = d * c.alpha, sz = std::min<uint32_t>(c.alpha, lq - start);
      for (uint32_t pos = 0; pos < ext; ++pos) {
        if (pos >= start && pos < start + sz) continue;
        modup_ntt.push_back(LimbJob{dig + ((size_t)d * ext + pos) * N, pos < lq ? pos : L1 + (pos - lq), 0});
      }
    }
    for (uint32_t poly = 0; poly < 2; ++poly) {
      for (uint32_t s = 0; s < K; ++s) {
        md_in.push_back(LimbJob{ex + ((size_t)poly * ext + lq + s) * N, L1 + s, 0});
        md_out.push_back(LimbJob{pc + ((size_t)poly * K + s) * N, L1 + s, 0});
      }
      mdj.push_back(MdJob{pc + (size_t)poly * K * N, conv + (size_t)poly * lq * N, lq, 0});
      const uint64_t* add = poly == 0 ? k.add0 : k.add1;
      for (uint32_t i = 0; i < lq; ++i) {
        md_ntt.push_back(LimbJob{conv + ((size_t)poly * lq + i) * N, i, 0});
        md_epi.push_back(EpiJob{ex + ((size_t)poly * ext + i) * N, k.out + ((size_t)poly * lq + i) * N, add ? add + (size_t)i * N : nullptr, i,
                                (uint32_t)k.l});
      }
    }
  }
  // groups of key-switch items sharing (key, galois), levels descending inside a group. A group is
  // cut into chunks of at most ksk_group_chunk items (each chunk reads the key once), so no workgroup
  // runs far longer than the others, and the chunks are launched largest first.
  const uint32_t kGroupChunk = (uint32_t)std::max(1, c.tune.ksk_group_chunk);
  std::vector<size_t> order(ks.size());
  for (size_t i = 0; i < ks.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return std::make_tuple(ks[a].key, ks[a].galois, -ks[a].l) < std::make_tuple(ks[b].key, ks[b].galois, -ks[b].l);
  });
  std::vector<KskItem> ki;
  std::vector<KskGroup> kg;
  for (size_t t = 0; t < order.size();) {
    const Ks& k0 = ks[order[t]];
    KskGroup g{k0.key, k0.galois == 1 ? 1u : (uint32_t)h_inv_odd(k0.galois, 2ull * c.N), (uint32_t)ki.size(), 0, 0};
    while (t < order.size() && ks[order[t]].key == k0.key && ks[order[t]].galois == k0.galois && g.count < kGroupChunk) {
      const Ks& k = ks[order[t]];
      ki.push_back(KskItem{A + k.dig, k.c0, A + k.ext, k.x, (uint32_t)k.l, 0});
      g.max_lq = std::max<uint32_t>(g.max_lq, k.l + 1);
      ++g.count;
      ++t;
    }
    kg.push_back(g);
  }
  std::stable_sort(kg.begin(), kg.end(), [](const KskGroup& a, const KskGroup& b) { return a.count * a.max_lq > b.count * b.max_lq; });
  // rescale sources: relinearized products and plaintext products that asked for a Rescale
  struct RsSrc {
    uint32_t l;
    uint64_t *R, *top, *e, *out;
    const uint64_t *a, *pt;   // fused plaintext product: the epilogue reads a_i and pt_i
  };
  std::vector<RsSrc> rss;
  for (size_t i = 0; i < P.mults.size(); ++i)
    if (P.mults[i].rescale)
      rss.push_back(RsSrc{(uint32_t)P.mults[i].l, A + ms[i].relin, A + ms[i].top, A + ms[i].e, P.mults[i].out, nullptr, nullptr});
  for (size_t i = 0; i < P.mulpts.size(); ++i) {
    const auto& m = P.mulpts[i];
    if (m.rescale)
      rss.push_back(RsSrc{(uint32_t)m.l, A + ps[i].relin, A + ps[i].top, A + ps[i].e, m.out, fused_mulpt(m) ? m.a : nullptr, m.pt});
  }
  for (const RsSrc& rs : rss) {
    const uint32_t l = rs.l, lq = l + 1;
    uint64_t* R = rs.R;
    uint64_t* top = rs.top;
    uint64_t* e = rs.e;
    uint64_t* rout = rs.out;
    for (uint32_t comp = 0; comp < 2; ++comp) {
      rs_in.push_back(LimbJob{R + ((size_t)comp * lq + l) * N, l, 0});
      rs_out.push_back(LimbJob{top + (size_t)comp * N, l, 0});
      for (uint32_t ii = 0; ii < l; ++ii) {
        rj.push_back(RsJob{top + (size_t)comp * N, e + ((size_t)comp * l + ii) * N, l, ii});
        rs_ntt.push_back(LimbJob{e + ((size_t)comp * l + ii) * N, ii, 0});
        if (rs.a)
          rs_epi.push_back(EpiJob{rs.a + ((size_t)comp * lq + ii) * N, rout + ((size_t)comp * l + ii) * N, rs.pt + (size_t)ii * N, ii, l});
        else
          rs_epi.push_back(EpiJob{R + ((size_t)comp * lq + ii) * N, rout + ((size_t)comp * l + ii) * N, nullptr, ii, l});
      }
    }
  }

  sort_by_prime(ks_in, &ks_out);
  sort_by_prime<LimbJob>(modup_ntt, nullptr);
  sort_by_prime(md_in, &md_out);
  sort_by_prime(md_ntt, &md_epi);
  sort_by_prime(rs_in, &rs_out);
  sort_by_prime(rs_ntt, &rs_epi);
  if (c.tune.ntt_mix) {
    mix_prime_classes(c, ks_in, &ks_out);
    mix_prime_classes<LimbJob>(c, modup_ntt, nullptr);
    mix_prime_classes(c, md_in, &md_out);
    mix_prime_classes(c, md_ntt, &md_epi);
    mix_prime_classes(c, rs_in, &rs_out);
    mix_prime_classes(c, rs_ntt, &rs_epi);
  }

  TableBuilder tb;
  const size_t o_mulpt = tb.put(pj);
  const size_t o_tensor = tb.put(tj), o_add = tb.put(aj), o_ks_in = tb.put(ks_in), o_ks_out = tb.put(ks_out), o_modup = tb.put(mj),
               o_modup_ntt = tb.put(modup_ntt), o_ki = tb.put(ki), o_kg = tb.put(kg), o_md_in = tb.put(md_in), o_md_out = tb.put(md_out),
               o_md = tb.put(mdj), o_md_ntt = tb.put(md_ntt), o_md_epi = tb.put(md_epi), o_rs_in = tb.put(rs_in), o_rs_out = tb.put(rs_out),
               o_rs = tb.put(rj), o_rs_ntt = tb.put(rs_ntt), o_rs_epi = tb.put(rs_epi);
  if (P.tables_cap < tb.bytes.size() || !P.d_tables) {
    // a run on another stream may still read the old tables; runs on the context stream are
    // ordered before any later use of the block returned to the pool
    if (P.d_tables && P.run_other) FH_HIP(hipEventSynchronize(P.run_ev));
    const size_t words = (std::max<size_t>(tb.bytes.size() + tb.bytes.size() / 4, 4096) + 7) / 8;
    P.tables = DBuf(c.pool, words);
    P.d_tables = P.tables.p;
    P.tables_cap = words * 8;
  }
  // Asynchronous copy on the context stream through the context's pinned staging ring: ordered
  // on the GPU after every earlier run of this batch on that stream (and, through run_ev, after
  // the last run on another stream), so the host never waits for the batch in flight.
  if (P.run_other) {
    FH_HIP(hipStreamWaitEvent(c.stream, P.run_ev, 0));
    P.run_other = false;
  }
  P.uploaded = c.uploads.upload(P.d_tables, tb.bytes.data(), tb.bytes.size(), c.stream);
  const uint8_t* D = static_cast<const uint8_t*>(P.d_tables);
  P.t_mulpt = reinterpret_cast<const MulPtJob*>(D + o_mulpt);
  P.n_mulpt = (uint32_t)pj.size();
  P.t_tensor = reinterpret_cast<const TensorJob*>(D + o_tensor);
  P.n_tensor = (uint32_t)tj.size();
  P.t_add = reinterpret_cast<const AddJob*>(D + o_add);
  P.n_add = (uint32_t)aj.size();
  P.t_ks_in = reinterpret_cast<const LimbJob*>(D + o_ks_in);
  P.t_ks_out = reinterpret_cast<const LimbJob*>(D + o_ks_out);
  P.n_ks_intt = (uint32_t)ks_in.size();
  P.t_modup = reinterpret_cast<const ModUpJob*>(D + o_modup);
  P.n_modup = (uint32_t)mj.size();
  P.t_modup_ntt = reinterpret_cast<const LimbJob*>(D + o_modup_ntt);
  P.n_modup_ntt = (uint32_t)modup_ntt.size();
  P.t_ksk_items = reinterpret_cast<const KskItem*>(D + o_ki);
  P.t_ksk_groups = reinterpret_cast<const KskGroup*>(D + o_kg);
  P.n_ksk_groups = (uint32_t)kg.size();
  P.t_md_in = reinterpret_cast<const LimbJob*>(D + o_md_in);
  P.t_md_out = reinterpret_cast<const LimbJob*>(D + o_md_out);
  P.n_md_intt = (uint32_t)md_in.size();
  P.t_md = reinterpret_cast<const MdJob*>(D + o_md);
  P.n_md = (uint32_t)mdj.size();
  P.t_md_ntt = reinterpret_cast<const LimbJob*>(D + o_md_ntt);
  P.t_md_epi = reinterpret_cast<const EpiJob*>(D + o_md_epi);
  P.n_md_ntt = (uint32_t)md_ntt.size();
  P.t_rs_in = reinterpret_cast<const LimbJob*>(D + o_rs_in);
  P.t_rs_out = reinterpret_cast<const LimbJob*>(D + o_rs_out);
  P.n_rs_intt = (uint32_t)rs_in.size();
  P.t_rs = reinterpret_cast<const RsJob*>(D + o_rs);
  P.n_rs = (uint32_t)rj.size();
  P.t_rs_ntt = reinterpret_cast<const LimbJob*>(D + o_rs_ntt);
  P.t_rs_epi = reinterpret_cast<const EpiJob*>(D + o_rs_epi);
  P.n_rs_ntt = (uint32_t)rs_ntt.size();
  if (P.n_modup > 65535 || P.n_md > 65535 || P.n_rs > 65535 || P.n_tensor > 65535 || P.n_add > 65535 || P.n_ksk_groups > 65535 ||
      P.n_mulpt > 65535)
    throw Error(-3, "batch too large for one launch per stage");
  P.final = true;
}

namespace {

void launch_all(Context& c, BatchPlan& P, hipStream_t st) {
  const uint32_t N = c.N;
  if (P.n_add) {
    ProfScope ps_(c, "batch_add", st);
    k_add_jobs<<<grid_tiles(N, P.n_add, 1), kT, 0, st>>>(P.t_add, c.d_prime, N);
  }
  if (P.n_tensor) {
    ProfScope ps_(c, "batch_tensor", st);
    k_tensor_jobs<<<grid_tiles(N, P.n_tensor, 1), kT, 0, st>>>(P.t_tensor, c.d_prime, N);
  }
  if (P.n_mulpt) {
    ProfScope ps_(c, "batch_mulpt", st);
    k_mulpt_jobs<<<grid_tiles(N, P.n_mulpt, 1), kT, 0, st>>>(P.t_mulpt, c.d_prime, N);
  }
  if (P.n_ks_intt) {
    ntt_inv_jobs(c, P.t_ks_in, P.t_ks_out, P.n_ks_intt, st);
    ConvConst cc{N, (uint32_t)c.L1, (uint32_t)c.K, (uint32_t)c.alpha, (uint32_t)c.dnum, c.d_prime, c.d_modup_qhinv, c.d_modup_qhinv_sh,
                 c.d_modup_qhat_m, c.d_md_phinv, c.d_md_phinv_sh, c.d_md_phat_m, c.reduced_noise ? c.d_modup_negq_m : nullptr,
                 c.reduced_noise ? c.d_md_negp_m : nullptr};
    {
      ProfScope ps_(c, "modup_conv", st);
      const dim3 g = grid_tiles(N, P.n_modup, 2);
      FH_CONV_DISPATCH(c, c.alpha, k_modup_conv_jobs, (g, kT, 0, st), (P.t_modup, cc))
    }
    FH_HIP(hipGetLastError());
    ntt_fwd_jobs(c, P.t_modup_ntt, P.n_modup_ntt, st);
    {
      ProfScope ps_(c, "ksk_batch", st);
      KskConst kc{N, (uint32_t)c.L1, (uint32_t)c.K, (uint32_t)c.alpha, (uint32_t)c.dnum, (uint32_t)c.logN, c.d_prime, c.d_pmod, c.d_pmod_sh};
      const dim3 g = grid_tiles(N, c.L1 + c.K, 2, P.n_ksk_groups);
      if (c.split_mac_ok) k_ksk_jobs<Mac30><<<g, kT, 0, st>>>(P.t_ksk_items, P.t_ksk_groups, kc);
      else k_ksk_jobs<MacWide><<<g, kT, 0, st>>>(P.t_ksk_items, P.t_ksk_groups, kc);
    }
    FH_HIP(hipGetLastError());
    ntt_inv_jobs(c, P.t_md_in, P.t_md_out, P.n_md_intt, st);
    {
      ProfScope ps_(c, "moddown_conv", st);
      const dim3 g = grid_tiles(N, P.n_md, 2);
      FH_CONV_DISPATCH(c, c.K, k_moddown_conv_jobs, (g, kT, 0, st), (P.t_md, cc))
    }
    FH_HIP(hipGetLastError());
    ntt_fwd_jobs_moddown(c, P.t_md_ntt, P.t_md_epi, P.n_md_ntt, st);
  }
  if (P.n_rs_intt) {
    ntt_inv_jobs(c, P.t_rs_in, P.t_rs_out, P.n_rs_intt, st);
    {
      ProfScope ps_(c, "rescale_expand", st);
      k_rescale_expand_jobs<<<grid_tiles(N, P.n_rs, 1), kT, 0, st>>>(P.t_rs, c.d_prime, c.d_rs_QlQl, c.d_rs_QlQl_sh, (uint32_t)c.L1, N);
    }
    FH_HIP(hipGetLastError());
    ntt_fwd_jobs_rescale(c, P.t_rs_ntt, P.t_rs_epi, P.n_rs_ntt, st);
  }
  FH_HIP(hipGetLastError());
}

}  // namespace

void Batch::run(hipStream_t st, bool graph) {
  BatchPlan& P = *p_;
  if (!P.final) finalize();
  const bool other = st != c_.stream;
  // the tables were uploaded on the context stream
  if (other) FH_HIP(hipStreamWaitEvent(st, P.uploaded, 0));
  struct Mark {  // after the launches on another stream: the next upload waits for them
    BatchPlan& P;
    hipStream_t st;
    bool on;
    ~Mark() {
      if (!on) return;
      if (!P.run_ev && hipEventCreateWithFlags(&P.run_ev, hipEventDisableTiming) != hipSuccess) return;
      if (hipEventRecord(P.run_ev, st) == hipSuccess) P.run_other = true;
    }
  } mark{P, st, other};
  if (!graph) {
    launch_all(c_, P, st);
    return;
  }
  if (P.exec && P.graph_stream != st) throw Error(-1, "a captured batch replays on the stream it was captured on");
  if (!P.exec) {
    // capture once: profiling event records are kept out of the graph
    const bool prof_on = c_.prof.on;
    c_.prof.on = false;
    FH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      launch_all(c_, P, st);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      c_.prof.on = prof_on;
      throw;
    }
    FH_HIP(hipStreamEndCapture(st, &P.graph));
    c_.prof.on = prof_on;
    FH_HIP(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
    P.graph_stream = st;
  }
  FH_HIP(hipGraphLaunch(P.exec, st));
}

}  // namespace fh
