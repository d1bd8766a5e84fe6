/*
 * CPU oracle for the RNS-CKKS evaluation path — TEST INFRASTRUCTURE ONLY.
 * See ckks_oracle.h for the layout conventions. Each function cites the
 * reference code (paths relative to the FIDESlib-fork tree) or the OpenFHE
 * routine it restates. OpenFHE itself is not available here; where the
 * reference defers to OpenFHE, the published OpenFHE algorithm is restated and
 * the parity status is stated at the routine.
 */
#define _GNU_SOURCE /* sincos */
#include "ckks_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;
typedef uint64_t u64;

/* ------------------------------------------------------------------------- */
/* scalar number theory                                                      */
/* ------------------------------------------------------------------------- */

u64 orc_mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }

u64 orc_powmod(u64 a, u64 e, u64 q) {
  u64 r = 1 % q;
  a %= q;
  while (e) {
    if (e & 1) r = orc_mulmod(r, a, q);
    a = orc_mulmod(a, a, q);
    e >>= 1;
  }
  return r;
}

u64 orc_invmod(u64 a, u64 q) {
  /* extended Euclid over signed 128-bit */
  i128 t = 0, nt = 1, r = q, nr = a % q;
  while (nr) {
    i128 qq = r / nr, tmp;
    tmp = t - qq * nt; t = nt; nt = tmp;
    tmp = r - qq * nr; r = nr; nr = tmp;
  }
  if (r != 1) return 0;
  if (t < 0) t += q;
  return (u64)t;
}

static u64 modadd(u64 a, u64 b, u64 q) { u64 s = a + b; return s >= q ? s - q : s; }
static u64 modsub(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

/* Deterministic Miller-Rabin for 64-bit inputs (bases cover all n < 2^64).
 * OpenFHE's MillerRabinPrimalityTest is probabilistic; on primes both agree. */
int orc_is_prime(u64 n) {
  static const u64 bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return 0;
  for (int i = 0; i < 12; ++i) {
    if (n == bases[i]) return 1;
    if (n % bases[i] == 0) return 0;
  }
  u64 d = n - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  for (int i = 0; i < 12; ++i) {
    u64 x = orc_powmod(bases[i], d, n);
    if (x == 1 || x == n - 1) continue;
    int comp = 1;
    for (int r = 1; r < s; ++r) {
      x = orc_mulmod(x, x, n);
      if (x == n - 1) { comp = 0; break; }
    }
    if (comp) return 0;
  }
  return 1;
}

/* OpenFHE FirstPrime(nBits, m): smallest prime = 1 mod m above 2^nBits. */
u64 orc_first_prime(int bits, u64 m) {
  u64 q = (u64)1 << bits;
  u64 r = q % m;
  u64 qn = q + 1 - r;
  if (r > 0) qn += m;
  while (!orc_is_prime(qn)) qn += m;
  return qn;
}

/* OpenFHE LastPrime(nBits, m): largest prime = 1 mod m below 2^nBits. */
u64 orc_last_prime(int bits, u64 m) {
  u64 q = (u64)1 << bits;
  u64 r = q % m;
  u64 qn = q + 1 - r - m;
  while (!orc_is_prime(qn)) qn -= m;
  return qn;
}

u64 orc_previous_prime(u64 q, u64 m) {
  u64 qn = q - m;
  while (!orc_is_prime(qn)) qn -= m;
  return qn;
}

u64 orc_next_prime(u64 q, u64 m) {
  u64 qn = q + m;
  while (!orc_is_prime(qn)) qn += m;
  return qn;
}

/* OpenFHE RootOfUnity(m, q): the MINIMUM primitive m-th root of unity mod q
 * (any generator raised to (q-1)/m, then the smallest of its odd powers). */
u64 orc_root_of_unity(u64 m, u64 q) {
  if ((q - 1) % m) return 0;
  /* factor q-1 to test generators */
  u64 f[64];
  int nf = 0;
  u64 t = q - 1;
  for (u64 p = 2; p * p <= t && p < 2000000; ++p) {
    if (t % p == 0) {
      f[nf++] = p;
      while (t % p == 0) t /= p;
    }
  }
  if (t > 1) f[nf++] = t; /* remaining cofactor: prime for our moduli sizes (checked) */
  u64 g = 2;
  for (;; ++g) {
    int ok = 1;
    for (int i = 0; i < nf; ++i)
      if (orc_powmod(g, (q - 1) / f[i], q) == 1) { ok = 0; break; }
    if (ok) break;
  }
  u64 w = orc_powmod(g, (q - 1) / m, q);
  /* m is a power of two here: the primitive roots are w^k for odd k < m */
  u64 w2 = orc_mulmod(w, w, q);
  u64 best = w, cur = w;
  for (u64 k = 3; k < m; k += 2) {
    cur = orc_mulmod(cur, w2, q);
    if (cur < best) best = cur;
  }
  return best;
}

/* OpenFHE ParameterGenerationCKKSRNS (FIXEDMANUAL): the top modulus is
 * FirstPrime(scale_bits), then alternately PreviousPrime / NextPrime going down
 * the chain, q_0 = LastPrime(first_bits). Special primes (CryptoParametersRNS::
 * PrecomputeCRTTables, HYBRID): PreviousPrime from FirstPrime(aux_bits), skipping
 * any prime already in Q. Reference consumer: src/CKKS/Parameters.cu:12-18. */
int orc_gen_primes(int logN, int L1, int first_bits, int scale_bits, int K, int aux_bits, u64* out) {
  u64 m = (u64)2 << logN;
  if (L1 < 1) return -1;
  u64 q = orc_first_prime(scale_bits, m);
  out[L1 - 1] = q;
  u64 qnext = q, qprev = q;
  int cnt = 0;
  for (int i = L1 - 2; i >= 1; --i) {
    if ((cnt % 2) == 0) { qprev = orc_previous_prime(qprev, m); out[i] = qprev; }
    else { qnext = orc_next_prime(qnext, m); out[i] = qnext; }
    ++cnt;
  }
  if (L1 > 1) {
    if (first_bits == scale_bits) out[0] = orc_next_prime(qnext, m);
    else out[0] = orc_last_prime(first_bits, m);
  }
  u64 pprev = orc_first_prime(aux_bits, m);
  for (int k = 0; k < K; ++k) {
    int inq;
    do {
      pprev = orc_previous_prime(pprev, m);
      inq = 0;
      for (int j = 0; j < L1; ++j) if (out[j] == pprev) inq = 1;
    } while (inq);
    out[L1 + k] = pprev;
  }
  return 0;
}

/* bit length of the product of q[a..b) via log2 sum with exact integer tail check */
int orc_default_num_special(const u64* q, int L1, int dnum, int aux_bits) {
  int alpha = (L1 + dnum - 1) / dnum;
  int maxbits = 0;
  for (int d = 0; d < dnum; ++d) {
    /* exact bit length using a small multiprecision product */
    u64 acc[ORC_MAXP + 1];
    int n = 1;
    acc[0] = 1;
    for (int i = d * alpha; i < (d + 1) * alpha && i < L1; ++i) {
      u128 carry = 0;
      for (int j = 0; j < n; ++j) {
        u128 v = (u128)acc[j] * q[i] + carry;
        acc[j] = (u64)v;
        carry = v >> 64;
      }
      if (carry) acc[n++] = (u64)carry;
    }
    int bits = 64 * (n - 1) + (64 - __builtin_clzll(acc[n - 1]));
    if (bits > maxbits) maxbits = bits;
  }
  return (maxbits + aux_bits - 1) / aux_bits;
}

/* ------------------------------------------------------------------------- */
/* tiny multiprecision helpers (little-endian u64 limbs)                     */
/* ------------------------------------------------------------------------- */
typedef struct { u64 w[ORC_MAXP + 2]; int n; } mp_t;

static void mp_set(mp_t* a, u64 v) { a->w[0] = v; a->n = 1; }
static void mp_mul_u64(mp_t* a, u64 v) {
  u128 carry = 0;
  for (int j = 0; j < a->n; ++j) {
    u128 t = (u128)a->w[j] * v + carry;
    a->w[j] = (u64)t;
    carry = t >> 64;
  }
  if (carry) a->w[a->n++] = (u64)carry;
}
static u64 mp_mod_u64(const mp_t* a, u64 m) {
  u128 r = 0;
  for (int j = a->n - 1; j >= 0; --j) r = ((r << 64) | a->w[j]) % m;
  return (u64)r;
}
static void mp_div_u64(mp_t* a, u64 m) { /* floor division in place */
  u128 r = 0;
  for (int j = a->n - 1; j >= 0; --j) {
    u128 cur = (r << 64) | a->w[j];
    a->w[j] = (u64)(cur / m);
    r = cur % m;
  }
  while (a->n > 1 && a->w[a->n - 1] == 0) --a->n;
}

/* ------------------------------------------------------------------------- */
/* context                                                                   */
/* ------------------------------------------------------------------------- */
struct orc_ctx {
  int logN;
  u64 N;
  int L1, K, dnum, alpha, np;
  u64 q[ORC_MAXP];
  u64 psi[ORC_MAXP];
  u64 ninv[ORC_MAXP];
  u64* tw[ORC_MAXP];
  u64* itw[ORC_MAXP];
  /* ModDown: [P^_k^-1]_{p_k}, [P^_k]_{q_i}, [P^-1]_{q_i}, [P]_{q_i} */
  u64 PhatInv[ORC_MAXP];
  u64* PhatModq; /* [K][L1] */
  u64 PInv[ORC_MAXP];
  u64 Pmod[ORC_MAXP];
  /* Rescale: [q_l^-1]_{q_i}, QlQlInvModqlDivqlModq[l][i] */
  u64* qlInv;   /* [L1][L1] */
  u64* QlQlInv; /* [L1][L1] */
  /* ReducedNoise basis conversions (src/CKKS/ReducedNoise.cuh:12-37): a prescaled source residue
   * above q/2 counts as v - q, i.e. the sum gains -count * Q_src */
  int reduced_noise;
};

static int brev(u64 x, int bits) {
  u64 r = 0;
  for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
  return (int)r;
}

/* OpenFHE ChineseRemainderTransformFTTNat::PreCompute: Table[brev(i)] = psi^i,
 * i.e. tw[k] = psi^brev(k); the inverse table holds psi^-brev(k).
 * Reference: src/CKKS/openfhe-interface/RawCiphertext.cu:278-300, src/ConstantsGPU.cu:259-285. */
void orc_make_tables(u64 N, u64 q, u64 psi, u64* tw, u64* itw) {
  int logN = 0;
  while (((u64)1 << logN) < N) ++logN;
  u64 psiinv = orc_invmod(psi, q);
  u64 x = 1, xi = 1;
  for (u64 i = 0; i < N; ++i) {
    int r = brev(i, logN);
    tw[r] = x;
    itw[r] = xi;
    x = orc_mulmod(x, psi, q);
    xi = orc_mulmod(xi, psiinv, q);
  }
}

orc_ctx* orc_create(int logN, int L1, int K, int dnum, const u64* primes) {
  if (L1 + K > ORC_MAXP || L1 < 1 || dnum < 1) return NULL;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->logN = logN;
  c->N = (u64)1 << logN;
  c->L1 = L1;
  c->K = K;
  c->dnum = dnum;
  c->alpha = (L1 + dnum - 1) / dnum;
  c->np = L1 + K;
  u64 m = 2 * c->N;
  for (int i = 0; i < c->np; ++i) {
    c->q[i] = primes[i];
    c->psi[i] = orc_root_of_unity(m, primes[i]);
    c->ninv[i] = orc_invmod(c->N % primes[i], primes[i]);
    c->tw[i] = (u64*)malloc(c->N * 8);
    c->itw[i] = (u64*)malloc(c->N * 8);
    orc_make_tables(c->N, c->q[i], c->psi[i], c->tw[i], c->itw[i]);
  }
  /* ModDown constants (OpenFHE CryptoParametersRNS: PHatInvModp, PHatModq, PInvModq) */
  c->PhatModq = (u64*)calloc((size_t)(K ? K : 1) * L1, 8);
  for (int k = 0; k < K; ++k) {
    u64 pk = c->q[L1 + k];
    u64 prod = 1;
    for (int j = 0; j < K; ++j) if (j != k) prod = orc_mulmod(prod, c->q[L1 + j] % pk, pk);
    c->PhatInv[k] = orc_invmod(prod, pk);
    for (int i = 0; i < L1; ++i) {
      u64 qi = c->q[i];
      u64 pr = 1;
      for (int j = 0; j < K; ++j) if (j != k) pr = orc_mulmod(pr, c->q[L1 + j] % qi, qi);
      c->PhatModq[(size_t)k * L1 + i] = pr;
    }
  }
  for (int i = 0; i < L1; ++i) {
    u64 qi = c->q[i];
    u64 pr = 1;
    for (int j = 0; j < K; ++j) pr = orc_mulmod(pr, c->q[L1 + j] % qi, qi);
    c->Pmod[i] = pr;
    c->PInv[i] = orc_invmod(pr, qi);
  }
  /* Rescale constants (OpenFHE CryptoParametersCKKSRNS::PrecomputeCRTTables):
   * for dropping limb l, QlInvModql = [Q_{l-1}^-1]_{q_l},
   * result = floor(QlInvModql * Q_{l-1} / q_l), table[l][i] = result mod q_i. */
  c->qlInv = (u64*)calloc((size_t)L1 * L1, 8);
  c->QlQlInv = (u64*)calloc((size_t)L1 * L1, 8);
  for (int l = 1; l < L1; ++l) {
    u64 ql = c->q[l];
    mp_t Q;
    mp_set(&Q, 1);
    for (int i = 0; i < l; ++i) mp_mul_u64(&Q, c->q[i]);
    u64 QmodQl = mp_mod_u64(&Q, ql);
    u64 u = orc_invmod(QmodQl, ql);
    mp_mul_u64(&Q, u);
    mp_div_u64(&Q, ql);
    for (int i = 0; i < l; ++i) {
      c->QlQlInv[(size_t)l * L1 + i] = mp_mod_u64(&Q, c->q[i]);
      c->qlInv[(size_t)l * L1 + i] = orc_invmod(ql % c->q[i], c->q[i]);
    }
  }
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  for (int i = 0; i < c->np; ++i) { free(c->tw[i]); free(c->itw[i]); }
  free(c->PhatModq);
  free(c->qlInv);
  free(c->QlQlInv);
  free(c);
}

u64 orc_prime(const orc_ctx* c, int idx) { return c->q[idx]; }
u64 orc_psi(const orc_ctx* c, int idx) { return c->psi[idx]; }
int orc_alpha(const orc_ctx* c) { return c->alpha; }
const u64* orc_twiddles(const orc_ctx* c, int idx) { return c->tw[idx]; }
const u64* orc_inv_twiddles(const orc_ctx* c, int idx) { return c->itw[idx]; }
u64 orc_rescale_ql_inv(const orc_ctx* c, int l, int i) { return c->qlInv[(size_t)l * c->L1 + i]; }
u64 orc_rescale_QlQlInv(const orc_ctx* c, int l, int i) { return c->QlQlInv[(size_t)l * c->L1 + i]; }
u64 orc_P_mod(const orc_ctx* c, int i) { return c->Pmod[i]; }
u64 orc_P_inv_mod(const orc_ctx* c, int i) { return c->PInv[i]; }

/* ------------------------------------------------------------------------- */
/* NTT                                                                       */
/* ------------------------------------------------------------------------- */

/* OpenFHE NumberTheoreticTransformNat::ForwardTransformToBitReverseInPlace:
 * Cooley-Tukey, stage m = 1..N/2 (t = N/2m), group i uses tw[m + i]. Output
 * A[j] = sum_i a_i psi^((2 brev(j) + 1) i), canonical in [0, q).
 * Reference: src/NTT.cu:331-695 (two-pass GPU form of the same transform). */
void orc_ntt_fwd_raw(u64* a, u64 N, u64 q, const u64* tw) {
  u64 t = N >> 1;
  for (u64 m = 1; m < N; m <<= 1, t >>= 1) {
    for (u64 i = 0; i < m; ++i) {
      u64 w = tw[m + i];
      u64 j1 = 2 * i * t;
      for (u64 j = j1; j < j1 + t; ++j) {
        u64 u = a[j];
        u64 v = orc_mulmod(a[j + t], w, q);
        a[j] = modadd(u, v, q);
        a[j + t] = modsub(u, v, q);
      }
    }
  }
}

/* OpenFHE InverseTransformFromBitReverseInPlace: Gentleman-Sande with the
 * inverse table, then scaling by N^-1. Reference: src/NTT.cu:42-234. */
void orc_ntt_inv_raw(u64* a, u64 N, u64 q, const u64* itw, u64 ninv) {
  u64 t = 1;
  for (u64 m = N >> 1; m >= 1; m >>= 1, t <<= 1) {
    for (u64 i = 0; i < m; ++i) {
      u64 w = itw[m + i];
      u64 j1 = 2 * i * t;
      for (u64 j = j1; j < j1 + t; ++j) {
        u64 u = a[j], v = a[j + t];
        a[j] = modadd(u, v, q);
        a[j + t] = orc_mulmod(modsub(u, v, q), w, q);
      }
    }
  }
  for (u64 j = 0; j < N; ++j) a[j] = orc_mulmod(a[j], ninv, q);
}

void orc_ntt_fwd(const orc_ctx* c, u64* a, const int* pids, int nlimbs) {
#pragma omp parallel for schedule(dynamic)
  for (int l = 0; l < nlimbs; ++l) {
    int p = pids[l];
    orc_ntt_fwd_raw(a + (size_t)l * c->N, c->N, c->q[p], c->tw[p]);
  }
}

void orc_ntt_inv(const orc_ctx* c, u64* a, const int* pids, int nlimbs) {
#pragma omp parallel for schedule(dynamic)
  for (int l = 0; l < nlimbs; ++l) {
    int p = pids[l];
    orc_ntt_inv_raw(a + (size_t)l * c->N, c->N, c->q[p], c->itw[p], c->ninv[p]);
  }
}

static void ntt1(const orc_ctx* c, u64* a, int p) { orc_ntt_fwd_raw(a, c->N, c->q[p], c->tw[p]); }
static void intt1(const orc_ctx* c, u64* a, int p) { orc_ntt_inv_raw(a, c->N, c->q[p], c->itw[p], c->ninv[p]); }

/* ------------------------------------------------------------------------- */
/* element-wise ops (limb i uses prime i). Reference: src/AddSub.cuh:12-50,   */
/* src/CKKS/ElemenwiseBatchKernels.cu:1363-1400 (tensor product).            */
/* ------------------------------------------------------------------------- */
void orc_add(const orc_ctx* c, const u64* a, const u64* b, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = modadd(a[o + j], b[o + j], q);
  }
}
void orc_sub(const orc_ctx* c, const u64* a, const u64* b, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = modsub(a[o + j], b[o + j], q);
  }
}
void orc_mul(const orc_ctx* c, const u64* a, const u64* b, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = orc_mulmod(a[o + j], b[o + j], q);
  }
}
void orc_neg(const orc_ctx* c, const u64* a, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = a[o + j] ? q - a[o + j] : 0;
  }
}
void orc_mul_scalar(const orc_ctx* c, const u64* a, const u64* s, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = orc_mulmod(a[o + j], s[l] % q, q);
  }
}
void orc_add_scalar(const orc_ctx* c, const u64* a, const u64* s, u64* out, int nl) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    size_t o = (size_t)l * c->N;
    for (u64 j = 0; j < c->N; ++j) out[o + j] = modadd(a[o + j], s[l] % q, q);
  }
}

/* OpenFHE LeveledSHECKKSRNS::EvalMultCore: (c0 d0, c0 d1 + c1 d0, c1 d1). */
void orc_tensor(const orc_ctx* c, const u64* ct1, const u64* ct2, u64* out3, int nl) {
  size_t P = (size_t)nl * c->N;
  const u64 *c0 = ct1, *c1 = ct1 + P, *d0 = ct2, *d1 = ct2 + P;
#pragma omp parallel for
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[l];
    for (u64 j = 0; j < c->N; ++j) {
      size_t x = (size_t)l * c->N + j;
      out3[x] = orc_mulmod(c0[x], d0[x], q);
      out3[P + x] = modadd(orc_mulmod(c0[x], d1[x], q), orc_mulmod(c1[x], d0[x], q), q);
      out3[2 * P + x] = orc_mulmod(c1[x], d1[x], q);
    }
  }
}

void orc_square(const orc_ctx* c, const u64* ct1, u64* out3, int nl) { orc_tensor(c, ct1, ct1, out3, nl); }

/* ------------------------------------------------------------------------- */
/* Rescale                                                                   */
/* ------------------------------------------------------------------------- */

/* OpenFHE NativeVector::SwitchModulus: centred lift of v (mod om) into mod nm. */
static u64 switch_modulus(u64 v, u64 om, u64 nm) {
  u64 half = om >> 1;
  if (nm > om) {
    if (v > half) v += nm - om;
    return v;
  }
  u64 diff = nm - (om % nm);
  if (v > half) v += diff;
  if (v >= nm) v %= nm;
  return v;
}

/* OpenFHE DCRTPoly::DropLastElementAndScale, as the reference fuses it:
 * src/CKKS/LimbPartition.cu:711-797, src/NTTfusions.cuh:214-235, src/Rescale.cuh:44-64.
 * out_i = x_i * [q_l^-1]_{q_i} + NTT(SM_{q_l->q_i}(INTT(x_l)) * QlQlInv[l][i]). */
void orc_rescale_poly(const orc_ctx* c, const u64* x, int l, u64* out) {
  u64 N = c->N;
  u64* top = (u64*)malloc(N * 8);
  memcpy(top, x + (size_t)l * N, N * 8);
  intt1(c, top, l);
  u64 ql = c->q[l];
#pragma omp parallel for
  for (int i = 0; i < l; ++i) {
    u64 qi = c->q[i];
    u64* e = (u64*)malloc(N * 8);
    u64 f = c->QlQlInv[(size_t)l * c->L1 + i];
    for (u64 j = 0; j < N; ++j) e[j] = orc_mulmod(switch_modulus(top[j], ql, qi), f, qi);
    ntt1(c, e, i);
    u64 s = c->qlInv[(size_t)l * c->L1 + i];
    const u64* xi = x + (size_t)i * N;
    u64* oi = out + (size_t)i * N;
    for (u64 j = 0; j < N; ++j) oi[j] = modadd(orc_mulmod(xi[j], s, qi), e[j], qi);
    free(e);
  }
  free(top);
}

void orc_rescale_ct(const orc_ctx* c, const u64* ct, int l, int ncomp, u64* out) {
  for (int k = 0; k < ncomp; ++k)
    orc_rescale_poly(c, ct + (size_t)k * (l + 1) * c->N, l, out + (size_t)k * l * c->N);
}

/* CKKS bootstrapping ModRaise, the raise step of src/CKKS/Bootstrap.cu:708-930 (tower 0 of each
 * component, coefficient form, SwitchModulus to every tower: the broadcast of
 * src/CKKS/ElemenwiseBatchKernels.cu:140-148), then back to EVALUATION form. ct: ncomp components
 * of l_in + 1 limbs; out: ncomp components of l + 1 limbs. */
void orc_mod_raise(const orc_ctx* c, const u64* ct, int l_in, int ncomp, int l, u64* out) {
  u64 N = c->N;
  u64* t = (u64*)malloc(N * 8);
  for (int k = 0; k < ncomp; ++k) {
    memcpy(t, ct + (size_t)k * (l_in + 1) * N, N * 8);
    intt1(c, t, 0);
    for (int i = 0; i <= l; ++i) {
      u64* o = out + ((size_t)k * (l + 1) + i) * N;
      for (u64 j = 0; j < N; ++j) o[j] = i == 0 ? t[j] : switch_modulus(t[j], c->q[0], c->q[i]);
      ntt1(c, o, i);
    }
  }
  free(t);
}

/* ------------------------------------------------------------------------- */
/* hybrid key switching                                                      */
/* ------------------------------------------------------------------------- */
int orc_num_digits(const orc_ctx* c, int l) {
  int b = (l + 1 + c->alpha - 1) / c->alpha;
  return b > c->dnum ? c->dnum : b;
}

/* extended-basis limb position -> prime index */
static int ext_prime(const orc_ctx* c, int l, int pos) { return pos <= l ? pos : c->L1 + (pos - l - 1); }

/* OpenFHE KeySwitchHYBRID::EvalKeySwitchPrecomputeCore + DCRTPoly::ApproxSwitchCRTBasis.
 * Reference: src/CKKS/LimbPartition.cu:851-968 (modup), src/CKKS/Conv.cu:274-430
 * (DecompAndModUpConv), constants src/ConstantsGPU.cu:539-632.
 * Digit d holds primes [d*alpha, min((d+1)*alpha, l+1)). For each coefficient:
 * y_i = [x_i * Qhat_i^-1]_{q_i}; target j: sum_i y_i * [Qhat_i]_{p_j} mod p_j. */
void orc_modup(const orc_ctx* c, const u64* x, int l, u64* out) {
  u64 N = c->N;
  int beta = orc_num_digits(c, l);
  int ext = l + 1 + c->K;
  for (int d = 0; d < beta; ++d) {
    int start = d * c->alpha;
    int end = start + c->alpha;
    if (end > l + 1) end = l + 1;
    int sz = end - start;
    u64* y = (u64*)malloc((size_t)sz * N * 8);
    u64 qhinv[ORC_MAXP];
    for (int i = 0; i < sz; ++i) {
      int pi = start + i;
      u64 qi = c->q[pi];
      u64 prod = 1;
      for (int j = start; j < end; ++j) if (j != pi) prod = orc_mulmod(prod, c->q[j] % qi, qi);
      qhinv[i] = orc_invmod(prod, qi);
      memcpy(y + (size_t)i * N, x + (size_t)pi * N, N * 8);
      intt1(c, y + (size_t)i * N, pi);
      for (u64 k = 0; k < N; ++k) y[(size_t)i * N + k] = orc_mulmod(y[(size_t)i * N + k], qhinv[i], qi);
    }
    u64* od = out + (size_t)d * ext * N;
    /* ReducedNoise: per coefficient, the number of prescaled sources above q_i/2 */
    unsigned char* lifted = NULL;
    if (c->reduced_noise) {
      lifted = (unsigned char*)calloc(N, 1);
      for (int i = 0; i < sz; ++i) {
        u64 half = c->q[start + i] >> 1;
        for (u64 k = 0; k < N; ++k) lifted[k] += y[(size_t)i * N + k] > half;
      }
    }
#pragma omp parallel for schedule(dynamic)
    for (int pos = 0; pos < ext; ++pos) {
      int p = ext_prime(c, l, pos);
      u64* o = od + (size_t)pos * N;
      if (pos >= start && pos < end) {
        memcpy(o, x + (size_t)pos * N, N * 8);
        continue;
      }
      u64 pj = c->q[p];
      u64 qh[ORC_MAXP];
      for (int i = 0; i < sz; ++i) {
        u64 pr = 1;
        for (int j = start; j < end; ++j) if (j != start + i) pr = orc_mulmod(pr, c->q[j] % pj, pj);
        qh[i] = pr;
      }
      u64 negQ = 0;  /* [-Q_d]_{p_j}, Q_d the product of the digit's active primes */
      if (lifted) {
        u64 qd = 1;
        for (int j = start; j < end; ++j) qd = orc_mulmod(qd, c->q[j] % pj, pj);
        negQ = qd ? pj - qd : 0;
      }
      for (u64 k = 0; k < N; ++k) {
        u128 s = 0;
        for (int i = 0; i < sz; ++i) s += (u128)y[(size_t)i * N + k] * qh[i];
        if (lifted) s += (u128)lifted[k] * negQ;
        o[k] = (u64)(s % pj);
      }
      ntt1(c, o, p);
    }
    free(lifted);
    free(y);
  }
}

/* OpenFHE KeySwitchHYBRID::EvalFastKeySwitchCoreExt.
 * Reference: src/CKKS/ElemenwiseBatchKernels.cu:202-273 (fusedDotKSK_2_),
 * key layout src/CKKS/KeySwitchingKey.cu:20-35. */
void orc_ksk_inner(const orc_ctx* c, const u64* digits, int l, const u64* key, u64* out0, u64* out1) {
  u64 N = c->N;
  int beta = orc_num_digits(c, l);
  int ext = l + 1 + c->K;
  int full = c->L1 + c->K;
#pragma omp parallel for
  for (int pos = 0; pos < ext; ++pos) {
    int p = ext_prime(c, l, pos);
    u64 q = c->q[p];
    u64* o0 = out0 + (size_t)pos * N;
    u64* o1 = out1 + (size_t)pos * N;
    for (u64 k = 0; k < N; ++k) { o0[k] = 0; o1[k] = 0; }
    for (int d = 0; d < beta; ++d) {
      const u64* dg = digits + ((size_t)d * ext + pos) * N;
      const u64* kb = key + (((size_t)d * 2 + 0) * full + p) * N;
      const u64* ka = key + (((size_t)d * 2 + 1) * full + p) * N;
      for (u64 k = 0; k < N; ++k) {
        o0[k] = modadd(o0[k], orc_mulmod(dg[k], kb[k], q), q);
        o1[k] = modadd(o1[k], orc_mulmod(dg[k], ka[k], q), q);
      }
    }
  }
}

/* OpenFHE DCRTPoly::ApproxModDown (CKKS: no t factor).
 * Reference: src/CKKS/LimbPartition.cu:2145-2217, src/CKKS/Conv.cu:27-133 (ModDown2),
 * src/NTTfusions.cuh:237-252 (moddown_fusion), constants src/ConstantsGPU.cu:480-537.
 * y_k = [INTT(x_Pk) * Phat_k^-1]_{p_k}; conv_i = sum_k y_k [Phat_k]_{q_i};
 * out_i = (x_i - NTT(conv_i)) * [P^-1]_{q_i}. */
void orc_moddown(const orc_ctx* c, const u64* x, int l, u64* out) {
  u64 N = c->N;
  int K = c->K;
  u64* y = (u64*)malloc((size_t)(K ? K : 1) * N * 8);
  for (int k = 0; k < K; ++k) {
    int p = c->L1 + k;
    u64 pk = c->q[p];
    memcpy(y + (size_t)k * N, x + (size_t)(l + 1 + k) * N, N * 8);
    intt1(c, y + (size_t)k * N, p);
    for (u64 j = 0; j < N; ++j) y[(size_t)k * N + j] = orc_mulmod(y[(size_t)k * N + j], c->PhatInv[k], pk);
  }
  unsigned char* lifted = NULL;
  if (c->reduced_noise) {
    lifted = (unsigned char*)calloc(N, 1);
    for (int k = 0; k < K; ++k) {
      u64 half = c->q[c->L1 + k] >> 1;
      for (u64 j = 0; j < N; ++j) lifted[j] += y[(size_t)k * N + j] > half;
    }
  }
#pragma omp parallel for
  for (int i = 0; i <= l; ++i) {
    u64 qi = c->q[i];
    u64 negP = c->Pmod[i] ? qi - c->Pmod[i] : 0;  /* [-P]_{q_i} */
    u64* conv = (u64*)malloc(N * 8);
    for (u64 j = 0; j < N; ++j) {
      u128 s = 0;
      for (int k = 0; k < K; ++k) s += (u128)y[(size_t)k * N + j] * c->PhatModq[(size_t)k * c->L1 + i];
      if (lifted) s += (u128)lifted[j] * negP;
      conv[j] = (u64)(s % qi);
    }
    ntt1(c, conv, i);
    const u64* xi = x + (size_t)i * N;
    u64* oi = out + (size_t)i * N;
    for (u64 j = 0; j < N; ++j) oi[j] = orc_mulmod(modsub(xi[j], conv[j], qi), c->PInv[i], qi);
    free(conv);
  }
  free(lifted);
  free(y);
}

void orc_set_reduced_noise(orc_ctx* c, int on) { c->reduced_noise = on != 0; }

/* OpenFHE PrecomputeAutoMap + AutomorphismTransform (EVALUATION form, gather):
 * out[brev(j)] = x[brev(((2j+1) k mod 2N) >> 1)].
 * Reference: src/Rotation.cuh:14-34 writes the same permutation as a scatter with
 * the inverse index (k = 5^(2N-r), src/CKKS/RNSPoly.cpp:656-658). */
void orc_automorph(const orc_ctx* c, const u64* x, int nl, uint32_t k, u64* out) {
  u64 N = c->N;
  int logN = c->logN;
  u64 mask = 2 * N - 1;
  int* map = (int*)malloc(N * sizeof(int));
  for (u64 j = 0; j < N; ++j) {
    u64 jt = 2 * j + 1;
    u64 idx = ((jt * k) & mask) >> 1;
    map[brev(j, logN)] = brev(idx, logN);
  }
  for (int l = 0; l < nl; ++l) {
    const u64* xi = x + (size_t)l * N;
    u64* oi = out + (size_t)l * N;
    for (u64 j = 0; j < N; ++j) oi[j] = xi[map[j]];
  }
  free(map);
}

/* OpenFHE FindAutomorphismIndex2nComplex: 5^r mod 2N (5^-1 for negative r). */
uint32_t orc_rot_index(const orc_ctx* c, int r) {
  u64 m = 2 * c->N;
  if (r == 0) return 1;
  u64 g = 5;
  if (r < 0) { g = orc_invmod(5, m); r = -r; }
  return (uint32_t)orc_powmod(g, (u64)r, m);
}

/* ------------------------------------------------------------------------- */
/* ciphertext-level ops                                                      */
/* ------------------------------------------------------------------------- */

/* KeySwitch-first relinearization: ks = ModDown(<ModUp(c2), evk>); c0 += ks0; c1 += ks1.
 * Reference: src/CKKS/Ciphertext.cpp:618-622, 846-866. */
void orc_relinearize(const orc_ctx* c, const u64* ct3, int l, const u64* key, u64* out2) {
  u64 N = c->N;
  size_t P = (size_t)(l + 1) * N;
  int ext = l + 1 + c->K;
  int beta = orc_num_digits(c, l);
  u64* dig = (u64*)malloc((size_t)beta * ext * N * 8);
  u64* t0 = (u64*)malloc((size_t)ext * N * 8);
  u64* t1 = (u64*)malloc((size_t)ext * N * 8);
  u64* k0 = (u64*)malloc(P * 8);
  u64* k1 = (u64*)malloc(P * 8);
  orc_modup(c, ct3 + 2 * P, l, dig);
  orc_ksk_inner(c, dig, l, key, t0, t1);
  orc_moddown(c, t0, l, k0);
  orc_moddown(c, t1, l, k1);
  orc_add(c, ct3, k0, out2, l + 1);
  orc_add(c, ct3 + P, k1, out2 + P, l + 1);
  free(dig); free(t0); free(t1); free(k0); free(k1);
}

void orc_eval_mult(const orc_ctx* c, const u64* ct1, const u64* ct2, int l, const u64* key, u64* out2) {
  size_t P = (size_t)(l + 1) * c->N;
  u64* t = (u64*)malloc(3 * P * 8);
  orc_tensor(c, ct1, ct2, t, l + 1);
  orc_relinearize(c, t, l, key, out2);
  free(t);
}

/* Hoisted rotation formulation (OpenFHE fork EvalAutomorphismCore / EvalFastRotationExt
 * with addFirst, then KeySwitchDown): t0 = sum digit*b + P*c0 (Q limbs), t1 = sum digit*a,
 * both automorphed in Q_l∪P. Reference: src/CKKS/Ciphertext.cpp:1102-1168,
 * src/CKKS/ElemenwiseBatchKernels.cu:277-377 (P*c0 term :315-319). */
/* Rotation core. c0_mode 0: no c0; 1: c0 over Q_l folded as [P]_{q_i} c0; 2: c0 already in the
 * P-scaled extended form over Q_l∪P, added on every limb (reference hoistedRotateDotKSK_2_,
 * src/CKKS/ElemenwiseBatchKernels.cu:302-307, the c0_modup branch). */
static void rotate_core(const orc_ctx* c, const u64* c0, int c0_mode, const u64* c1, int l, const u64* key, uint32_t k, u64* out_ext) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  int beta = orc_num_digits(c, l);
  u64* dig = (u64*)malloc((size_t)beta * ext * N * 8);
  u64* t0 = (u64*)malloc((size_t)ext * N * 8);
  u64* t1 = (u64*)malloc((size_t)ext * N * 8);
  orc_modup(c, c1, l, dig);
  orc_ksk_inner(c, dig, l, key, t0, t1);
  if (c0_mode == 1) {
    for (int i = 0; i <= l; ++i) {
      u64 qi = c->q[i];
      for (u64 j = 0; j < N; ++j) {
        size_t x = (size_t)i * N + j;
        t0[x] = modadd(t0[x], orc_mulmod(c0[x], c->Pmod[i], qi), qi);
      }
    }
  } else if (c0_mode == 2) {
    for (int pos = 0; pos < ext; ++pos) {
      u64 q = c->q[ext_prime(c, l, pos)];
      for (u64 j = 0; j < N; ++j) {
        size_t x = (size_t)pos * N + j;
        t0[x] = modadd(t0[x], c0[x], q);
      }
    }
  }
  orc_automorph(c, t0, ext, k, out_ext);
  orc_automorph(c, t1, ext, k, out_ext + (size_t)ext * N);
  free(dig); free(t0); free(t1);
}

void orc_eval_rotate_ext(const orc_ctx* c, const u64* ct, int l, const u64* key, uint32_t k, int add_c0, u64* out_ext) {
  rotate_core(c, ct, add_c0 ? 1 : 0, ct + (size_t)(l + 1) * c->N, l, key, k, out_ext);
}

/* BSGS linear transform in the reference's extended-basis formulation
 * (src/CKKS/LinearTransform.cu:65-254, FUSED branch, ONLY_C1 = true):
 *   fast[0] = (P c0, P c1) over Q_l∪P (Ciphertext::extend, src/CKKS/Ciphertext.cpp:1087-1100),
 *   fast[i] = hoisted rotation by i*stride kept in Q_l∪P (no ModDown);
 *   R_j = sum_{i, bStep j + i < rowSize} fast[i] (.) diag[bStep j + i]   (MAC, both components);
 *   Horner from j = gStep-1 down: R_j += acc; for j > 0 and a non-trivial giant step, ModDown c1
 *   only and rotate by bStep*stride with the extended c0 folded directly; at j = 0 rotate by
 *   offset (then ModDown both) or ModDown both.
 * diags: rowSize plaintexts over Q_l∪P (l+1+K limbs each). baby_keys: bStep keys (entry i used
 * for i*stride != 0 mod slots), each dnum*2*(L1+K) limbs. Output: 2 x (l+1) limbs. */
void orc_linear_transform(const orc_ctx* c, const u64* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                          const u64* diags, const u64* baby_keys, const uint32_t* baby_k, const u64* giant_key, uint32_t giant_k,
                          const u64* off_key, uint32_t off_k, u64* out) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  size_t E = (size_t)ext * N, Q = (size_t)(l + 1) * N;
  size_t key_words = (size_t)c->dnum * 2 * c->np * N;
  int gStep = (rowSize + bStep - 1) / bStep;
  u64* fast = (u64*)calloc((size_t)bStep * 2 * E, 8);
  for (int i = 0; i < bStep; ++i) {
    u64* f = fast + (size_t)i * 2 * E;
    if (((long long)i * stride) % slots == 0) {
      for (int comp = 0; comp < 2; ++comp)
        for (int p = 0; p <= l; ++p) {
          u64 qp = c->q[p];
          for (u64 j = 0; j < N; ++j) f[comp * E + (size_t)p * N + j] = orc_mulmod(ct[comp * Q + (size_t)p * N + j], c->Pmod[p], qp);
        }
    } else {
      rotate_core(c, ct, 1, ct + Q, l, baby_keys + (size_t)i * key_words, baby_k[i], f);
    }
  }
  u64* acc = (u64*)calloc(2 * E, 8);
  u64* R = (u64*)malloc(2 * E * 8);
  u64* c1q = (u64*)malloc(Q * 8);
  int have_acc = 0;
  for (int j = gStep - 1; j >= 0; --j) {
    memset(R, 0, 2 * E * 8);
    for (int i = 0; i < bStep && bStep * j + i < rowSize; ++i) {
      const u64* d = diags + (size_t)(bStep * j + i) * E;
      const u64* f = fast + (size_t)i * 2 * E;
      for (int comp = 0; comp < 2; ++comp)
#pragma omp parallel for
        for (int pos = 0; pos < ext; ++pos) {
          u64 q = c->q[ext_prime(c, l, pos)];
          for (u64 n = 0; n < N; ++n) {
            size_t x = (size_t)pos * N + n;
            R[comp * E + x] = modadd(R[comp * E + x], orc_mulmod(f[comp * E + x], d[x], q), q);
          }
        }
    }
    if (have_acc) {
      for (int comp = 0; comp < 2; ++comp)
        for (int pos = 0; pos < ext; ++pos) {
          u64 q = c->q[ext_prime(c, l, pos)];
          for (u64 n = 0; n < N; ++n) {
            size_t x = comp * E + (size_t)pos * N + n;
            R[x] = modadd(R[x], acc[x], q);
          }
        }
    }
    if (j > 0) {
      if (((long long)bStep * stride) % slots != 0) {
        orc_moddown(c, R + E, l, c1q);
        rotate_core(c, R, 2, c1q, l, giant_key, giant_k, acc);
      } else {
        memcpy(acc, R, 2 * E * 8);
      }
      have_acc = 1;
    } else if (offset != 0) {
      orc_moddown(c, R + E, l, c1q);
      u64* t = (u64*)malloc(2 * E * 8);
      rotate_core(c, R, 2, c1q, l, off_key, off_k, t);
      orc_moddown(c, t, l, out);
      orc_moddown(c, t + E, l, out + Q);
      free(t);
    } else {
      orc_moddown(c, R, l, out);
      orc_moddown(c, R + E, l, out + Q);
    }
  }
  free(fast); free(acc); free(R); free(c1q);
}

/* The same transform with the giant steps as independent rotations: R_j (j >= 1) is rotated by
 * j*bStep*stride with its own key (ModDown of c1, the extended c0 folded directly, as a Horner
 * step does) and the results are added to R_0, which then takes the offset rotation or the
 * ModDown of both components. This build's own sequence for its bootstrap's transforms (it has
 * no reference counterpart; the reference's transform is the Horner form above): the same
 * linear map, different rounding. giant_keys: gStep-1 keys back to back (entry j-1 for
 * j*bStep*stride), giant_k their automorphism indices; an R_j whose multiple is 0 mod slots is
 * added unrotated (its entry unused). */
void orc_linear_transform_flat(const orc_ctx* c, const u64* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                               const u64* diags, const u64* baby_keys, const uint32_t* baby_k, const u64* giant_keys,
                               const uint32_t* giant_k, const u64* off_key, uint32_t off_k, u64* out) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  size_t E = (size_t)ext * N, Q = (size_t)(l + 1) * N;
  size_t key_words = (size_t)c->dnum * 2 * c->np * N;
  int gStep = (rowSize + bStep - 1) / bStep;
  int giant = ((long long)bStep * stride) % slots != 0;
  u64* fast = (u64*)calloc((size_t)bStep * 2 * E, 8);
  for (int i = 0; i < bStep; ++i) {
    u64* f = fast + (size_t)i * 2 * E;
    if (((long long)i * stride) % slots == 0) {
      for (int comp = 0; comp < 2; ++comp)
        for (int p = 0; p <= l; ++p) {
          u64 qp = c->q[p];
          for (u64 j = 0; j < N; ++j) f[comp * E + (size_t)p * N + j] = orc_mulmod(ct[comp * Q + (size_t)p * N + j], c->Pmod[p], qp);
        }
    } else {
      rotate_core(c, ct, 1, ct + Q, l, baby_keys + (size_t)i * key_words, baby_k[i], f);
    }
  }
  u64* acc = (u64*)calloc(2 * E, 8);
  u64* R = (u64*)malloc(2 * E * 8);
  u64* T = (u64*)malloc(2 * E * 8);
  u64* c1q = (u64*)malloc(Q * 8);
  for (int j = 0; j < gStep; ++j) {
    memset(R, 0, 2 * E * 8);
    for (int i = 0; i < bStep && bStep * j + i < rowSize; ++i) {
      const u64* d = diags + (size_t)(bStep * j + i) * E;
      const u64* f = fast + (size_t)i * 2 * E;
      for (int comp = 0; comp < 2; ++comp)
#pragma omp parallel for
        for (int pos = 0; pos < ext; ++pos) {
          u64 q = c->q[ext_prime(c, l, pos)];
          for (u64 n = 0; n < N; ++n) {
            size_t x = (size_t)pos * N + n;
            R[comp * E + x] = modadd(R[comp * E + x], orc_mulmod(f[comp * E + x], d[x], q), q);
          }
        }
    }
    const u64* add = R;
    if (j > 0 && giant && ((long long)j * bStep * stride) % slots != 0) {
      orc_moddown(c, R + E, l, c1q);
      rotate_core(c, R, 2, c1q, l, giant_keys + (size_t)(j - 1) * key_words, giant_k[j - 1], T);
      add = T;
    }
    for (int comp = 0; comp < 2; ++comp)
      for (int pos = 0; pos < ext; ++pos) {
        u64 q = c->q[ext_prime(c, l, pos)];
        for (u64 n = 0; n < N; ++n) {
          size_t x = comp * E + (size_t)pos * N + n;
          acc[x] = modadd(acc[x], add[x], q);
        }
      }
  }
  if (offset != 0) {
    orc_moddown(c, acc + E, l, c1q);
    rotate_core(c, acc, 2, c1q, l, off_key, off_k, T);
    orc_moddown(c, T, l, out);
    orc_moddown(c, T + E, l, out + Q);
  } else {
    orc_moddown(c, acc, l, out);
    orc_moddown(c, acc + E, l, out + Q);
  }
  free(fast); free(acc); free(R); free(T); free(c1q);
}

/* Rotation-fold sum from a start factor (the cascade: CKKS::Accumulate(ct, bStep, stride, size,
   startFactor), src/CKKS/AccumulateBroadcast.cu:42-84, behind CudaEngine::accumulateSumInPlace(ct,
   slots, stride, start), api/engine/cuda/CudaEngine.cpp:780-783): for s = start, start*radix, ... <
   size, the rotations by i*stride*s (0 < i < radix, i*s < size) of the current ciphertext are formed
   in the extended basis (one shared decomposition, P*c0 folded in), added to P*ct
   (Ciphertext::extend), and BOTH components take one ModDown at every level.
   keys / ks: one key and Galois element per rotation, in fold order. out may alias ct. */
void orc_accumulate_sum(const orc_ctx* c, const u64* ct, int l, int radix, int stride, int size, int start,
                        const u64* keys, const uint32_t* ks, u64* out) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  size_t E = (size_t)ext * N, Q = (size_t)(l + 1) * N;
  size_t key_words = (size_t)c->dnum * 2 * c->np * N;
  (void)stride;
  if (out != ct) memcpy(out, ct, 2 * Q * 8);
  if (start <= 0 || size <= 0) return;
  u64* acc = (u64*)malloc(2 * E * 8);
  u64* t = (u64*)malloc(2 * E * 8);
  int used = 0;
  for (long long s = start; s < size; s *= radix) {
    memset(acc, 0, 2 * E * 8);
    for (int comp = 0; comp < 2; ++comp)
      for (int p = 0; p <= l; ++p)
        for (u64 j = 0; j < N; ++j)
          acc[comp * E + (size_t)p * N + j] = orc_mulmod(out[comp * Q + (size_t)p * N + j], c->Pmod[p], c->q[p]);
    for (long long i = 1; i < radix && i * s < size; ++i, ++used) {
      rotate_core(c, out, 1, out + Q, l, keys + (size_t)used * key_words, ks[used], t);
      for (int comp = 0; comp < 2; ++comp)
#pragma omp parallel for
        for (int pos = 0; pos < ext; ++pos) {
          u64 q = c->q[ext_prime(c, l, pos)];
          for (u64 n = 0; n < N; ++n) {
            size_t x = comp * E + (size_t)pos * N + n;
            acc[x] = modadd(acc[x], t[x], q);
          }
        }
    }
    orc_moddown(c, acc, l, out);
    orc_moddown(c, acc + E, l, out + Q);
  }
  free(acc); free(t);
}

/* The lazy rotation-fold sum (CKKS::Accumulate(ct, bStep, stride, size),
   src/CKKS/AccumulateBroadcast.cu:122-160, behind CudaEngine::accumulateSum and
   accumulateSumInPlace(ct, slots, stride), api/engine/cuda/CudaEngine.cpp:755-777). For s = 1,
   radix, radix^2, ... < size: the rotations by i*stride*s (0 < i < radix, i*s < size) share one
   decomposition of c1 and stay in Q_l∪P (Ciphertext::rotate_hoisted with ext, src/CKKS/Ciphertext.cpp
   :1221-1350). c0 enters the first level as P*c0 over Q_l (0 over P) and from then on stays in
   Q_l∪P: every later rotation adds it on every limb before the automorphism (the c0_modup branch of
   hoistedRotateDotKSK_2_, src/CKKS/ElemenwiseBatchKernels.cu:302-307). The rotations are added to
   (c0, P*c1) (RNSPoly::add of an extended operand into a Q_l one scales it by P first,
   src/CKKS/RNSPoly.cpp:272-286 and src/AddSub.cu:57-70), c1 takes its ModDown at the end of every
   level, and c0 takes ONE ModDown after the last level. keys / ks: one per rotation in fold order
   (orc_accumulate_sum's index order with start 1). out may alias ct. */
void orc_accumulate_lazy(const orc_ctx* c, const u64* ct, int l, int radix, int size, const u64* keys, const uint32_t* ks,
                         u64* out) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  size_t E = (size_t)ext * N, Q = (size_t)(l + 1) * N;
  size_t key_words = (size_t)c->dnum * 2 * c->np * N;
  if (out != ct) memcpy(out, ct, 2 * Q * 8);
  if (size <= 1 || radix < 2) return;
  u64* c0e = (u64*)malloc(E * 8);     /* c0 over Q_l∪P once the first level has run */
  u64* a0 = (u64*)malloc(E * 8);
  u64* a1 = (u64*)malloc(E * 8);
  u64* t = (u64*)malloc(2 * E * 8);
  int c0_ext = 0, used = 0;
  for (long long s = 1; s < size; s *= radix) {
    /* accumulators: c0 (as P*c0 on the first level, extended afterwards) and P*c1 */
    memset(a1, 0, E * 8);
    for (int p = 0; p <= l; ++p)
      for (u64 j = 0; j < N; ++j) a1[(size_t)p * N + j] = orc_mulmod(out[Q + (size_t)p * N + j], c->Pmod[p], c->q[p]);
    if (c0_ext) {
      memcpy(a0, c0e, E * 8);
    } else {
      memset(a0, 0, E * 8);
      for (int p = 0; p <= l; ++p)
        for (u64 j = 0; j < N; ++j) a0[(size_t)p * N + j] = orc_mulmod(out[(size_t)p * N + j], c->Pmod[p], c->q[p]);
    }
    /* every rotation of this level reads the level's input (c0 or its extension, c1) */
    for (long long i = 1; i < radix && i * s < size; ++i, ++used) {
      rotate_core(c, c0_ext ? c0e : out, c0_ext ? 2 : 1, out + Q, l, keys + (size_t)used * key_words, ks[used], t);
#pragma omp parallel for
      for (int pos = 0; pos < ext; ++pos) {
        u64 q = c->q[ext_prime(c, l, pos)];
        for (u64 n = 0; n < N; ++n) {
          size_t x = (size_t)pos * N + n;
          a0[x] = modadd(a0[x], t[x], q);
          a1[x] = modadd(a1[x], t[E + x], q);
        }
      }
    }
    orc_moddown(c, a1, l, out + Q);
    memcpy(c0e, a0, E * 8);
    c0_ext = 1;
  }
  orc_moddown(c, c0e, l, out);
  free(c0e); free(a0); free(a1); free(t);
}

void orc_eval_rotate(const orc_ctx* c, const u64* ct, int l, const u64* key, uint32_t k, u64* out) {
  u64 N = c->N;
  int ext = l + 1 + c->K;
  u64* e = (u64*)malloc((size_t)2 * ext * N * 8);
  orc_eval_rotate_ext(c, ct, l, key, k, 1, e);
  orc_moddown(c, e, l, out);
  orc_moddown(c, e + (size_t)ext * N, l, out + (size_t)(l + 1) * N);
  free(e);
}

/* ------------------------------------------------------------------------- */
/* PRNG: mt19937_64 (same stream as std::mt19937_64)                         */
/* ------------------------------------------------------------------------- */
#define MT_N 312
#define MT_M 156
typedef struct { u64 mt[MT_N]; int i; } mt64;

static void mt_seed(mt64* s, u64 seed) {
  s->mt[0] = seed;
  for (int i = 1; i < MT_N; ++i) s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (u64)i;
  s->i = MT_N;
}

static u64 mt_next(mt64* s) {
  static const u64 A = 0xB5026F5AA96619E9ULL, UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  if (s->i >= MT_N) {
    int i;
    u64 x;
    for (i = 0; i < MT_N - MT_M; ++i) {
      x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
      s->mt[i] = s->mt[i + MT_M] ^ (x >> 1) ^ ((x & 1) ? A : 0);
    }
    for (; i < MT_N - 1; ++i) {
      x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
      s->mt[i] = s->mt[i + (MT_M - MT_N)] ^ (x >> 1) ^ ((x & 1) ? A : 0);
    }
    x = (s->mt[MT_N - 1] & UM) | (s->mt[0] & LM);
    s->mt[MT_N - 1] = s->mt[MT_M - 1] ^ (x >> 1) ^ ((x & 1) ? A : 0);
    s->i = 0;
  }
  u64 x = s->mt[s->i++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

u64 orc_mt_first(u64 seed, int skip) {
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  for (int i = 0; i < skip; ++i) mt_next(s);
  u64 v = mt_next(s);
  free(s);
  return v;
}

static u64 mt_uniform_mod(mt64* s, u64 q) {
  int bits = 64 - __builtin_clzll(q);
  for (;;) {
    u64 r = mt_next(s) >> (64 - bits);
    if (r < q) return r;
  }
}

static double mt_unit(mt64* s) { return (double)(mt_next(s) >> 11) * (1.0 / 9007199254740992.0); }

/* rounded Gaussian, sigma 3.19 (OpenFHE's default error distribution width) */
static int64_t mt_gauss(mt64* s) {
  double u1 = mt_unit(s), u2 = mt_unit(s);
  if (u1 < 1e-300) u1 = 1e-300;
  double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  return (int64_t)llround(z * 3.19);
}

static u64 signed_mod(int64_t v, u64 q) {
  if (v >= 0) return (u64)v % q;
  u64 r = (u64)(-(v + 1)) % q; /* avoid overflow on INT64_MIN */
  r = (r + 1) % q;
  return r ? q - r : 0;
}

void orc_random_poly(const orc_ctx* c, u64 seed, const int* pids, int nlimbs, u64* out) {
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  for (int l = 0; l < nlimbs; ++l) {
    u64 q = c->q[pids[l]];
    for (u64 j = 0; j < c->N; ++j) out[(size_t)l * c->N + j] = mt_uniform_mod(s, q);
  }
  free(s);
}

/* small signed coefficients -> EVAL over the given primes */
static void small_to_eval(const orc_ctx* c, const int64_t* v, const int* pids, int nl, u64* out) {
  for (int l = 0; l < nl; ++l) {
    u64 q = c->q[pids[l]];
    for (u64 j = 0; j < c->N; ++j) out[(size_t)l * c->N + j] = signed_mod(v[j], q);
  }
  orc_ntt_fwd(c, out, pids, nl);
}

void orc_gen_secret(const orc_ctx* c, u64 seed, int h, u64* s_out) {
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  int64_t* v = (int64_t*)calloc(c->N, sizeof(int64_t));
  if (h <= 0) {
    for (u64 j = 0; j < c->N; ++j) v[j] = (int64_t)(mt_next(s) % 3) - 1;
  } else {
    int placed = 0;
    while (placed < h) {
      u64 pos = mt_next(s) % c->N;
      if (v[pos]) continue;
      v[pos] = (mt_next(s) & 1) ? 1 : -1;
      ++placed;
    }
  }
  int pids[ORC_MAXP];
  for (int i = 0; i < c->np; ++i) pids[i] = i;
  small_to_eval(c, v, pids, c->np, s_out);
  free(v);
  free(s);
}

/* OpenFHE KeySwitchHYBRID::KeySwitchGenInternal: for digit d,
 * b_d = -a_d s_new + e_d + [P]_{q_i} s_old on the digit's Q primes, -a_d s_new + e_d elsewhere.
 * Layout [dnum][b, a][L1+K][N]. */
void orc_gen_switch_key(const orc_ctx* c, const u64* s_old, const u64* s_new, u64 seed, u64* key) {
  u64 N = c->N;
  int full = c->np;
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  int pids[ORC_MAXP];
  for (int i = 0; i < full; ++i) pids[i] = i;
  int64_t* ev = (int64_t*)malloc(N * sizeof(int64_t));
  u64* e = (u64*)malloc((size_t)full * N * 8);
  for (int d = 0; d < c->dnum; ++d) {
    u64* b = key + ((size_t)d * 2 + 0) * full * N;
    u64* a = key + ((size_t)d * 2 + 1) * full * N;
    for (int i = 0; i < full; ++i)
      for (u64 j = 0; j < N; ++j) a[(size_t)i * N + j] = mt_uniform_mod(s, c->q[i]);
    for (u64 j = 0; j < N; ++j) ev[j] = mt_gauss(s);
    small_to_eval(c, ev, pids, full, e);
    int start = d * c->alpha, end = start + c->alpha;
    if (end > c->L1) end = c->L1;
    for (int i = 0; i < full; ++i) {
      u64 q = c->q[i];
      for (u64 j = 0; j < N; ++j) {
        size_t x = (size_t)i * N + j;
        u64 v = modsub(e[x], orc_mulmod(a[x], s_new[x], q), q);
        if (i >= start && i < end) v = modadd(v, orc_mulmod(c->Pmod[i], s_old[x], q), q);
        b[x] = v;
      }
    }
  }
  free(ev); free(e); free(s);
}

void orc_gen_relin_key(const orc_ctx* c, const u64* sk, u64 seed, u64* key) {
  size_t n = (size_t)c->L1 * c->N;
  u64* s2 = (u64*)malloc(n * 8);
  orc_mul(c, sk, sk, s2, c->L1);
  orc_gen_switch_key(c, s2, sk, seed, key);
  free(s2);
}

/* OpenFHE EvalAutomorphismKeyGen: switch from s to sigma_{k^-1}(s). */
void orc_gen_rot_key(const orc_ctx* c, const u64* sk, uint32_t k, u64 seed, u64* key) {
  size_t n = (size_t)c->np * c->N;
  u64* sp = (u64*)malloc(n * 8);
  uint32_t kinv = (uint32_t)orc_invmod(k, 2 * c->N);
  orc_automorph(c, sk, c->np, kinv, sp);
  orc_gen_switch_key(c, sk, sp, seed, key);
  free(sp);
}

/* ------------------------------------------------------------------------- */
/* CKKS encoding: the special FFT of Cheon-Kim-Kim-Song (HEAAN), slots at     */
/* zeta^(5^j); coefficient i*gap gets Re, N/2 + i*gap gets Im.               */
/* ------------------------------------------------------------------------- */
static void bitrev_cplx(double complex* v, int n) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { double complex t = v[i]; v[i] = v[j]; v[j] = t; }
  }
}

static void special_fft(double complex* v, int n, u64 M) {
  bitrev_cplx(v, n);
  for (int len = 2; len <= n; len <<= 1) {
    int lenh = len >> 1;
    u64 lenq = (u64)len << 2;
    for (int i = 0; i < n; i += len) {
      u64 g = 1;
      for (int j = 0; j < lenh; ++j) {
        u64 idx = (g % lenq) * (M / lenq);
        double ang = 2.0 * M_PI * (double)idx / (double)M;
        double sn, cs;
        sincos(ang, &sn, &cs); /* one call for both, as the encoders on the device and host do */
        double complex w = cs + I * sn;
        double complex u = v[i + j], t = v[i + j + lenh] * w;
        v[i + j] = u + t;
        v[i + j + lenh] = u - t;
        g = (g * 5) % M;
      }
    }
  }
}

static void special_ifft(double complex* v, int n, u64 M) {
  for (int len = n; len >= 1; len >>= 1) {
    int lenh = len >> 1;
    u64 lenq = (u64)len << 2;
    for (int i = 0; i < n; i += len) {
      u64 g = 1;
      for (int j = 0; j < lenh; ++j) {
        u64 idx = (lenq - (g % lenq)) * (M / lenq);
        double ang = 2.0 * M_PI * (double)idx / (double)M;
        double sn, cs;
        sincos(ang, &sn, &cs); /* one call for both, as the encoders on the device and host do */
        double complex w = cs + I * sn;
        double complex u = v[i + j] + v[i + j + lenh];
        double complex t = (v[i + j] - v[i + j + lenh]) * w;
        v[i + j] = u;
        v[i + j + lenh] = t;
        g = (g * 5) % M;
      }
    }
  }
  bitrev_cplx(v, n);
  for (int i = 0; i < n; ++i) v[i] /= n;
}

static void encode_coeffs(const orc_ctx* c, const double* slots, int n, double scale, int64_t* m) {
  u64 N = c->N;
  double complex* v = (double complex*)malloc(n * sizeof(double complex));
  for (int i = 0; i < n; ++i) v[i] = slots[2 * i] + I * slots[2 * i + 1];
  special_ifft(v, n, 2 * N);
  memset(m, 0, N * sizeof(int64_t));
  u64 gap = (N / 2) / n;
  for (int i = 0; i < n; ++i) {
    m[i * gap] = (int64_t)llround(creal(v[i]) * scale);
    m[N / 2 + i * gap] = (int64_t)llround(cimag(v[i]) * scale);
  }
  free(v);
}

void orc_encode(const orc_ctx* c, const double* slots, int n, double scale, int l, u64* out) {
  int64_t* m = (int64_t*)malloc(c->N * sizeof(int64_t));
  encode_coeffs(c, slots, n, scale, m);
  int pids[ORC_MAXP];
  for (int i = 0; i <= l; ++i) pids[i] = i;
  small_to_eval(c, m, pids, l + 1, out);
  free(m);
}

void orc_encode_ext(const orc_ctx* c, const double* slots, int n, double scale, int l, u64* out) {
  int64_t* m = (int64_t*)malloc(c->N * sizeof(int64_t));
  encode_coeffs(c, slots, n, scale, m);
  int pids[ORC_MAXP];
  for (int i = 0; i < l + 1 + c->K; ++i) pids[i] = ext_prime(c, l, i);
  small_to_eval(c, m, pids, l + 1 + c->K, out);
  free(m);
}

void orc_decode(const orc_ctx* c, const u64* mp, int l, int n, double scale, double* slots) {
  u64 N = c->N;
  int nl = l >= 1 ? 2 : 1;
  u64* t = (u64*)malloc((size_t)nl * N * 8);
  memcpy(t, mp, (size_t)nl * N * 8);
  int pids[2] = {0, 1};
  orc_ntt_inv(c, t, pids, nl);
  double* coef = (double*)malloc(N * sizeof(double));
  u64 q0 = c->q[0];
  if (nl == 1) {
    for (u64 j = 0; j < N; ++j) coef[j] = t[j] > q0 / 2 ? -(double)(q0 - t[j]) : (double)t[j];
  } else {
    u64 q1 = c->q[1];
    u64 q0inv = orc_invmod(q0 % q1, q1);
    u128 Q = (u128)q0 * q1;
    for (u64 j = 0; j < N; ++j) {
      u64 a = t[j], b = t[N + j];
      u64 k = orc_mulmod(modsub(b, a % q1, q1), q0inv, q1);
      u128 x = (u128)a + (u128)q0 * k;
      coef[j] = x > Q / 2 ? -(double)(Q - x) : (double)x;
    }
  }
  double complex* v = (double complex*)malloc(n * sizeof(double complex));
  u64 gap = (N / 2) / n;
  for (int i = 0; i < n; ++i) v[i] = (coef[i * gap] + I * coef[N / 2 + i * gap]) / scale;
  special_fft(v, n, 2 * N);
  for (int i = 0; i < n; ++i) { slots[2 * i] = creal(v[i]); slots[2 * i + 1] = cimag(v[i]); }
  free(v); free(coef); free(t);
}

/* secret-key encryption: c1 = a, c0 = -a s + e + m (over Q_l, EVAL) */
void orc_encrypt(const orc_ctx* c, const u64* m, int l, const u64* sk, u64 seed, u64* ct) {
  u64 N = c->N;
  size_t P = (size_t)(l + 1) * N;
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  u64* a = ct + P;
  for (int i = 0; i <= l; ++i)
    for (u64 j = 0; j < N; ++j) a[(size_t)i * N + j] = mt_uniform_mod(s, c->q[i]);
  int64_t* ev = (int64_t*)malloc(N * sizeof(int64_t));
  for (u64 j = 0; j < N; ++j) ev[j] = mt_gauss(s);
  u64* e = (u64*)malloc(P * 8);
  int pids[ORC_MAXP];
  for (int i = 0; i <= l; ++i) pids[i] = i;
  small_to_eval(c, ev, pids, l + 1, e);
  for (int i = 0; i <= l; ++i) {
    u64 q = c->q[i];
    for (u64 j = 0; j < N; ++j) {
      size_t x = (size_t)i * N + j;
      ct[x] = modadd(modsub(e[x], orc_mulmod(a[x], sk[x], q), q), m[x], q);
    }
  }
  free(ev); free(e); free(s);
}

void orc_decrypt(const orc_ctx* c, const u64* ct, int l, int ncomp, const u64* sk, u64* m) {
  u64 N = c->N;
  size_t P = (size_t)(l + 1) * N;
  for (int i = 0; i <= l; ++i) {
    u64 q = c->q[i];
    for (u64 j = 0; j < N; ++j) {
      size_t x = (size_t)i * N + j;
      u64 s = sk[x];
      u64 v = modadd(ct[x], orc_mulmod(ct[P + x], s, q), q);
      if (ncomp == 3) v = modadd(v, orc_mulmod(ct[2 * P + x], orc_mulmod(s, s, q), q), q);
      m[x] = v;
    }
  }
}

/* Public-key encryption as in OpenFHE 1.5.1 PKERNS::EncryptZeroCore (an un-vendored dependency,
 * reached from api/CryptoContext.cpp:900-921): v ternary, e0, e1 Gaussian; c0 = b v + e0 + m,
 * c1 = a v + e1 over Q_l. The sampler streams are this build's (mt19937_64), not OpenFHE's. */
void orc_encrypt_pk(const orc_ctx* c, const u64* m, int l, const u64* pk_b, const u64* pk_a, u64 seed, u64* ct) {
  u64 N = c->N;
  size_t P = (size_t)(l + 1) * N;
  mt64* s = (mt64*)malloc(sizeof(mt64));
  mt_seed(s, seed);
  int64_t* w = (int64_t*)malloc(3 * N * sizeof(int64_t));
  for (u64 j = 0; j < N; ++j) w[j] = (int64_t)(mt_next(s) % 3) - 1;
  for (u64 j = 0; j < 2 * N; ++j) w[N + j] = mt_gauss(s);
  u64* ev = (u64*)malloc(3 * P * 8);
  int pids[ORC_MAXP];
  for (int i = 0; i <= l; ++i) pids[i] = i;
  for (int t = 0; t < 3; ++t) small_to_eval(c, w + (size_t)t * N, pids, l + 1, ev + (size_t)t * P);
  const u64 *v = ev, *e0 = ev + P, *e1 = ev + 2 * P;
  for (int i = 0; i <= l; ++i) {
    u64 q = c->q[i];
    for (u64 j = 0; j < N; ++j) {
      size_t x = (size_t)i * N + j;
      ct[x] = modadd(modadd(orc_mulmod(pk_b[x], v[x], q), e0[x], q), m[x], q);
      ct[P + x] = modadd(orc_mulmod(pk_a[x], v[x], q), e1[x], q);
    }
  }
  free(ev); free(w); free(s);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
