// The key inner product fused into the row pass of the ModUp targets' forward NTT (the
// reference's NTT_KSK_DOT, src/NTTfusions.cuh:282-336, launched at src/CKKS/LimbPartition.cu:
// 1652-1662), for batched rotations.
//
// One workgroup per (rotation item, extended position, row tile): for each digit it finishes the
// digit's target limb on its rows (the row stages of the transform, the column pass having run
// over every target before) or reads the digit's own limb from the ciphertext, multiplies the
// canonical values by the key's words at the same positions and accumulates out_b / out_a in
// registers; then folds P c0 into the Q positions and writes both outputs through the
// automorphism. The transformed digits are never written to HBM and the key inner product does
// not read them back: per item and position that saves a write and a read of the digits' target
// limbs. The accumulation is mod p per digit (Montgomery products with the keys' Montgomery form),
// so the canonical results are those of the separate key inner product.
//
// Workgroups run in an XCD-aware order: the jobs of one (position, row tile) are consecutive and
// sorted by key, and consecutive jobs go to one XCD, so the items sharing a rotation key read its
// tile through that XCD's L2 at about the same time.
#include <algorithm>
#include <numeric>

#include "fh_keyswitch.hpp"
#include "fh_ntt_impl.hpp"

namespace fh {
namespace {

constexpr int kRkMaxGroups = 64;
constexpr int kRkMaxItems = 256;

struct RowsKskParams {
  uint32_t N, lq, ext, L1, K, beta, logN, alpha;
  uint32_t nitems, ntiles, total;   // jobs: total = ext x ntiles x nitems
  const uint64_t* digits;           // item b's digit d, position pos: digits + b dig_stride + (d ext + pos) N (column-pass output)
  uint64_t dig_stride;
  const uint64_t* own;              // item b's key-switched polynomial (EVAL, lq limbs): own + b own_stride
  uint64_t own_stride;
  const uint64_t* c0;               // optional: c0 of item b (EVAL, lq limbs) at c0 + b c0_stride, folded as P c0
  uint64_t c0_stride;
  const uint64_t* pmod;             // [P]_{q_i}, Shoup companions
  const uint64_t* pmod_sh;
  uint64_t* out;                    // item b's result (2 x ext limbs) at out + b out_stride
  uint64_t out_stride;
  const PrimeConst* prime;
};

struct RowsKskItems {
  const uint64_t* key[kRkMaxGroups];  // key of each (key, galois) group
  uint32_t gal_inv[kRkMaxGroups];     // inverse automorphism index (scatter form), 1 = none
  uint16_t item[kRkMaxItems];         // batch index of the i-th item in group order
  uint8_t grp[kRkMaxItems];           // its group
};

template <int LS, int TH>
__global__ __launch_bounds__(TH) void k_ntt_rows_ksk(RowsKskParams kp, RowsKskItems it, NttTables tb) {
  constexpr int S = 1 << LS, T = S / 16, PADS = RowTile<LS>::PAD, RMAX = TH / T;
  static_assert(LS > 4, "the fused row pass reads its outputs from the LDS tile");
  __shared__ uint64_t tile[RMAX * PADS];
  __shared__ uint64_t sw[S], swp[S];
  const uint32_t per = (kp.total + 7) / 8;
  const uint32_t J = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (J >= kp.total) return;  // the whole workgroup: J is uniform
  const uint32_t i = J % kp.nitems, pt = J / kp.nitems;
  const uint32_t tl = pt % kp.ntiles, pos = pt / kp.ntiles;
  const uint32_t b = it.item[i], g = it.grp[i];
  const uint32_t pid = pos < kp.lq ? pos : kp.L1 + (pos - kp.lq);
  const Sel sel = select(tb, (int)pid, false);
  const PrimeConst P = kp.prime[pid];
  const int tid = threadIdx.x, rr = tid / T, c0 = tid % T;
  const int R = blockDim.x / T;
  const uint32_t rbase = (tl * (uint32_t)R + (uint32_t)rr) * S;
  if (sel.fp) {
    double* swd = reinterpret_cast<double*>(sw);
    for (int k = tid; k < S; k += blockDim.x) swd[k] = sel.twd[k];
  } else {
    for (int k = tid; k < S; k += blockDim.x) {
      sw[k] = sel.tw[k];
      swp[k] = sel.twsh[k];
    }
  }
  __syncthreads();
  const size_t N = kp.N, E = (size_t)kp.L1 + kp.K;
  const uint32_t own_d = pos < kp.lq ? pos / kp.alpha : ~0u;
  const uint64_t* key = it.key[g];
  const uint64_t* t = tile + rr * PADS;
  uint64_t a0[16], a1[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) a0[k] = a1[k] = 0;
  for (uint32_t d = 0; d < kp.beta; ++d) {
    uint64_t x[16];
    if (d == own_d) {
      const uint64_t* ow = kp.own + b * kp.own_stride + (size_t)pos * N + rbase;
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = ow[c0 + T * k];
    } else {
      // every thread reads back only the tile words it wrote last (positions c0 + T k of its row
      // slot), so the next digit's first LDS writes need no barrier
      const uint64_t* row = kp.digits + b * kp.dig_stride + ((size_t)d * kp.ext + pos) * N + rbase;
      if (sel.fp)
        fwd_rows_core<LS>(sel.fa, (const double*)sw, (const double*)sw, sel.tsd, sel.tsd, row, rbase, tile, rr, c0, EpiStore{});
      else
        fwd_rows_core<LS>(sel.ia, (const uint64_t*)sw, (const uint64_t*)swp, sel.ts, sel.tssh, row, rbase, tile, rr, c0, EpiStore{});
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = t[rsw(c0 + T * k)];
    }
    const uint64_t* kb = key + ((size_t)(2 * d) * E + pid) * N + rbase;
    const uint64_t* ka = kb + E * N;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      a0[k] = add_mod(a0[k], mont_mul(x[k], kb[c0 + T * k], P), P.q);
      a1[k] = add_mod(a1[k], mont_mul(x[k], ka[c0 + T * k], P), P.q);
    }
  }
  const bool addc0 = kp.c0 != nullptr && pos < kp.lq;
  const uint64_t pm = addc0 ? kp.pmod[pos] : 0, pmsh = addc0 ? kp.pmod_sh[pos] : 0;
  const uint64_t* cz = addc0 ? kp.c0 + b * kp.c0_stride + (size_t)pos * N + rbase : nullptr;
  const uint32_t kinv = it.gal_inv[g];
  uint64_t* o0 = kp.out + b * kp.out_stride + (size_t)pos * N;
  uint64_t* o1 = o0 + (size_t)kp.ext * N;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t e = rbase + c0 + T * k;
    uint64_t r0 = a0[k];
    if (addc0) r0 = add_mod(r0, shoup(cz[c0 + T * k], pm, pmsh, P.q), P.q);
    const uint32_t jo = kinv != 1 ? auto_src(e, kinv, (int)kp.logN) : e;
    o0[jo] = r0;
    o1[jo] = a1[k];
  }
}

int split_of(const Context& c) { return c.tune.ntt_split ? c.tune.ntt_split : ntt_default_split(c.logN); }

// the column pass alone over the ModUp targets (up to 65535 limbs per launch)
template <int L1S, bool WIDE>
void modup_cols(const Context& c, const ModUpAddr& m, uint32_t n, hipStream_t st) {
  const NttTables tb = ntt_tables(c, L1S);
  for (uint32_t base = 0; base < n; base += 65535) {
    const uint32_t cnt = std::min<uint32_t>(65535, n - base);
    const Shift<ModUpAddr> sa{m, base};
    if (WIDE && c.tune.ntt_threads[0] == kNttWide) launch_fwd_cols<L1S, kNttWide>(c, sa, sa, cnt, tb, st);
    else launch_fwd_cols<L1S, kNttThreads>(c, sa, sa, cnt, tb, st);
  }
}

template <int L1S, int L2S>
void rows_ksk_launch(const Context& c, const RowsKskParams& kp0, const RowsKskItems& it, hipStream_t st) {
  if constexpr (L2S > 4 && L2S <= 8) {
    const NttTables tb = ntt_tables(c, L1S);
    const dim3 blk = row_block<L2S, kNttThreads>(c.N);
    const uint32_t R = blk.x / ((1u << L2S) / 16);
    RowsKskParams kp = kp0;
    kp.ntiles = (c.N >> L2S) / R;
    kp.total = kp.ext * kp.ntiles * kp.nitems;
    const uint32_t per = (kp.total + 7) / 8;
    ProfScope ps_(c, "ntt_rows_ksk", st);
    k_ntt_rows_ksk<L2S, kNttThreads><<<dim3(8 * per), blk, 0, st>>>(kp, it, tb);
  } else {
    throw Error(-1, "fused row pass: row length not supported");
  }
}

}  // namespace

bool rows_ksk_ok(const Context& c) {
  if (!c.tune.rot_fused || c.tune.ntt_team) return false;
  const int l2 = c.logN - split_of(c);
  return ntt_split_ok(c.logN, split_of(c)) && l2 > 4 && l2 <= 8;
}

void ntt_fwd_modup_cols(const Context& c, const ModUpAddr& a, uint32_t n, hipStream_t st) {
  if (n == 0) return;
  ModUpAddr m = a;
  m.mix = c.tune.ntt_mix ? 1u : 0u;
#define FH_COLS_CASE(LN, A, WIDE) \
  case LN * 16 + A:               \
    modup_cols<A, WIDE>(c, m, n, st); \
    break;
  switch (c.logN * 16 + split_of(c)) {
    FH_COLS_CASE(10, 4, false)
    FH_COLS_CASE(10, 5, true)
    FH_COLS_CASE(11, 5, true)
    FH_COLS_CASE(11, 6, false)
    FH_COLS_CASE(12, 5, false)
    FH_COLS_CASE(12, 6, true)
    FH_COLS_CASE(12, 7, false)
    FH_COLS_CASE(13, 6, true)
    FH_COLS_CASE(13, 7, false)
    FH_COLS_CASE(14, 6, false)
    FH_COLS_CASE(14, 7, true)
    FH_COLS_CASE(14, 8, false)
    FH_COLS_CASE(15, 7, true)
    FH_COLS_CASE(15, 8, false)
    FH_COLS_CASE(16, 8, true)
    default:
      throw Error(-1, "fused row pass: unsupported ring dimension or radix split");
  }
#undef FH_COLS_CASE
  FH_HIP(hipGetLastError());
}

void rows_ksk(Context& c, int nb, const uint64_t* digits, uint64_t dig_stride, const uint64_t* c0s, uint64_t c0_stride, bool add_c0,
              const uint64_t* own, uint64_t own_stride, int l, const uint64_t* const* keys, const uint32_t* galois, uint64_t* out,
              uint64_t out_stride, hipStream_t st) {
  if (nb <= 0) return;
  if (!rows_ksk_ok(c)) throw Error(-1, "fused row pass: not enabled for this context");
  const uint32_t lq = l + 1, ext = lq + c.K, beta = c.num_digits(l);
  std::vector<int> idx(nb);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (keys[a] != keys[b]) return keys[a] < keys[b];
    return galois[a] < galois[b];
  });
  RowsKskParams kp{c.N, lq, ext, (uint32_t)c.L1, (uint32_t)c.K, beta, (uint32_t)c.logN, (uint32_t)c.alpha, 0, 0, 0,
                   digits, dig_stride, own, own_stride, add_c0 ? c0s : nullptr, c0_stride, c.d_pmod, c.d_pmod_sh, out, out_stride,
                   c.d_prime};
  int i = 0;
  while (i < nb) {
    // one launch takes at most kRkMaxGroups groups and kRkMaxItems items (a group may continue)
    RowsKskItems it{};
    int ng = 0, n = 0;
    while (i < nb && ng < kRkMaxGroups && n < kRkMaxItems) {
      const uint64_t* k = keys[idx[i]];
      const uint32_t gl = galois[idx[i]];
      if (!k) throw Error(-1, "null key");
      if ((gl & 1) == 0 || gl >= 2 * c.N) throw Error(-1, "galois index must be odd and < 2N");
      it.key[ng] = k;
      it.gal_inv[ng] = gl == 1 ? 1 : (uint32_t)h_inv_odd(gl, 2ull * c.N);
      while (i < nb && n < kRkMaxItems && keys[idx[i]] == k && galois[idx[i]] == gl) {
        it.item[n] = (uint16_t)idx[i];
        it.grp[n] = (uint8_t)ng;
        ++n;
        ++i;
      }
      ++ng;
    }
    kp.nitems = (uint32_t)n;
#define FH_ROWS_CASE(LN, A)                                   \
  case LN * 16 + A:                                           \
    rows_ksk_launch<A, LN - A>(c, kp, it, st);                \
    break;
    switch (c.logN * 16 + split_of(c)) {
      FH_ROWS_CASE(10, 4)
      FH_ROWS_CASE(10, 5)
      FH_ROWS_CASE(11, 5)
      FH_ROWS_CASE(11, 6)
      FH_ROWS_CASE(12, 5)
      FH_ROWS_CASE(12, 6)
      FH_ROWS_CASE(12, 7)
      FH_ROWS_CASE(13, 6)
      FH_ROWS_CASE(13, 7)
      FH_ROWS_CASE(14, 6)
      FH_ROWS_CASE(14, 7)
      FH_ROWS_CASE(14, 8)
      FH_ROWS_CASE(15, 7)
      FH_ROWS_CASE(15, 8)
      FH_ROWS_CASE(16, 8)
      default:
        throw Error(-1, "fused row pass: unsupported ring dimension or radix split");
    }
#undef FH_ROWS_CASE
    FH_HIP(hipGetLastError());
  }
}

}  // namespace fh
