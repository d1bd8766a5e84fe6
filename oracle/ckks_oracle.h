/*
 * CPU oracle for the RNS-CKKS evaluation path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain-C restatement of the arithmetic that the reference
 * (FIDESlib, and the patched OpenFHE 1.5.1 fork it is bit-compatible with) runs
 * on the hot path: negacyclic NTT/INTT in OpenFHE's bit-reversed EVALUATION
 * order, RNS modular arithmetic, Rescale, hybrid key switching (ModUp, key inner
 * product, ModDown), automorphisms and the hoisted rotation formulation, and the
 * BSGS linear transform. Only the test suite, the smoke check and the CPU
 * baseline of the benchmark load it; the product never links or calls it.
 *
 * Layout: every polynomial is limb-major, one contiguous N-word array per limb,
 * values canonical in [0, q). A ciphertext is its components back to back.
 * Prime indices: 0..L1-1 are the Q chain (q_0 first), L1..L1+K-1 the special
 * primes P. A polynomial "over Q_l" has l+1 limbs (primes 0..l); "over Q_l∪P"
 * has l+1+K limbs (primes 0..l then L1..L1+K-1).
 *
 * Parity status (the checks live in tests/test_oracle.py, tests/test_bootstrap_model.py and the
 * GPU tests' closed-form checks):
 *   NTT / INTT            pinned: the reference's q=17, N=8 known-answer round trip, the closed
 *                         form A_j = sum a_i psi^((2 brev(j)+1) i) by brute force, negacyclic
 *                         convolution
 *   ModUp / ModDown       pinned: big-int fast base conversion and its error bounds, plain and
 *                         ReducedNoise
 *   Rescale               pinned: OpenFHE's QlQlInvModqlDivqlModq by the big-int floor formula,
 *                         output by CRT
 *   automorphism          pinned: coefficient-domain a(X^k) mod X^N + 1
 *   key switching, rotation, relinearization, linear transform, encryption, ModRaise,
 *   rotation-fold sum     pinned by decryption invariants and closed forms (the fold sum: its
 *                         decryption is the sum of the slot rotations, and each fold level equals
 *                         P*ct plus the extended rotations under one ModDown)
 *   prime chain, roots    pinned to OpenFHE's published FirstPrime / PreviousPrime / root rules by
 *                         an independent Python walk and brute force; parity unpinned against
 *                         OpenFHE itself (not available here)
 *   encode / decode       pinned by the round trip and the host and device encoders (same
 *                         rounding sequence, sincos twiddles); parity unpinned against OpenFHE's
 *                         FFTSpecialInv floating-point order
 *   key generation, sampling (mt19937_64 streams)
 *                         this build's own generators: the standard's known 10000th output pins
 *                         the generator; parity unpinned against OpenFHE's samplers
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXP 72

typedef struct orc_ctx orc_ctx;

/* ---- scalar number theory ---- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t q);
uint64_t orc_invmod(uint64_t a, uint64_t q);
int orc_is_prime(uint64_t n);
uint64_t orc_first_prime(int bits, uint64_t m);
uint64_t orc_last_prime(int bits, uint64_t m);
uint64_t orc_previous_prime(uint64_t q, uint64_t m);
uint64_t orc_next_prime(uint64_t q, uint64_t m);
uint64_t orc_root_of_unity(uint64_t m, uint64_t q);
/* FIXEDMANUAL CKKS modulus chain (L1 primes) followed by K special primes. */
int orc_gen_primes(int logN, int L1, int first_bits, int scale_bits, int K, int aux_bits, uint64_t* out);
/* OpenFHE's rule for the special-prime count: ceil(max digit bit length / aux_bits). */
int orc_default_num_special(const uint64_t* q, int L1, int dnum, int aux_bits);

/* ---- context ---- */
orc_ctx* orc_create(int logN, int L1, int K, int dnum, const uint64_t* primes);
/* ReducedNoise basis conversions in ModUp / ModDown (prescaled residues above q/2 read as v - q;
 * src/CKKS/ReducedNoise.cuh:12-37, src/CKKS/Conv.cu:76-89, 356-381); off by default */
void orc_set_reduced_noise(orc_ctx* c, int on);
void orc_destroy(orc_ctx* c);
uint64_t orc_prime(const orc_ctx* c, int idx);
uint64_t orc_psi(const orc_ctx* c, int idx);
int orc_alpha(const orc_ctx* c);
const uint64_t* orc_twiddles(const orc_ctx* c, int idx);     /* psi^brev(k), k < N */
const uint64_t* orc_inv_twiddles(const orc_ctx* c, int idx); /* psi^-brev(k) */
/* ModUp/ModDown/Rescale tables, exposed for pinning */
uint64_t orc_rescale_ql_inv(const orc_ctx* c, int l, int i);      /* [q_l^-1]_{q_i}, i<l */
uint64_t orc_rescale_QlQlInv(const orc_ctx* c, int l, int i);     /* OpenFHE QlQlInvModqlDivqlModq */
uint64_t orc_P_mod(const orc_ctx* c, int i);                      /* [P]_{q_i} */
uint64_t orc_P_inv_mod(const orc_ctx* c, int i);                  /* [P^-1]_{q_i} */

/* ---- NTT (in place, per limb; pids[i] is the prime index of limb i) ---- */
void orc_ntt_fwd(const orc_ctx* c, uint64_t* a, const int* pids, int nlimbs);
void orc_ntt_inv(const orc_ctx* c, uint64_t* a, const int* pids, int nlimbs);
/* single-prime versions with explicit tables (used by the q=17 known-answer test) */
void orc_ntt_fwd_raw(uint64_t* a, uint64_t N, uint64_t q, const uint64_t* tw);
void orc_ntt_inv_raw(uint64_t* a, uint64_t N, uint64_t q, const uint64_t* itw, uint64_t ninv);
void orc_make_tables(uint64_t N, uint64_t q, uint64_t psi, uint64_t* tw, uint64_t* itw);

/* ---- element-wise (over Q_l: prime i for limb i) ---- */
void orc_add(const orc_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int nlimbs);
void orc_sub(const orc_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int nlimbs);
void orc_mul(const orc_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int nlimbs);
void orc_neg(const orc_ctx* c, const uint64_t* a, uint64_t* out, int nlimbs);
/* multiply limb i by scalars[i] */
void orc_mul_scalar(const orc_ctx* c, const uint64_t* a, const uint64_t* scalars, uint64_t* out, int nlimbs);
void orc_add_scalar(const orc_ctx* c, const uint64_t* a, const uint64_t* scalars, uint64_t* out, int nlimbs);
/* tensor: (c0,c1)x(d0,d1) -> (c0d0, c0d1+c1d0, c1d1) */
void orc_tensor(const orc_ctx* c, const uint64_t* ct1, const uint64_t* ct2, uint64_t* out3, int nlimbs);
void orc_square(const orc_ctx* c, const uint64_t* ct1, uint64_t* out3, int nlimbs);

/* ---- Rescale: poly over Q_l (EVAL) -> over Q_{l-1} (EVAL) ---- */
void orc_rescale_poly(const orc_ctx* c, const uint64_t* x, int l, uint64_t* out);

/* ---- hybrid key switching ---- */
int orc_num_digits(const orc_ctx* c, int l);
/* ModUp of x over Q_l (EVAL): out = beta digits, each over Q_l∪P (EVAL) */
void orc_modup(const orc_ctx* c, const uint64_t* x, int l, uint64_t* out);
/* key inner product. key layout: [dnum][2 (b then a)][L1+K limbs][N] (EVAL).
 * out0 = sum_d digit_d * b_d, out1 = sum_d digit_d * a_d, both over Q_l∪P */
void orc_ksk_inner(const orc_ctx* c, const uint64_t* digits, int l, const uint64_t* key, uint64_t* out0, uint64_t* out1);
/* ModDown of x over Q_l∪P (EVAL) -> over Q_l (EVAL) */
void orc_moddown(const orc_ctx* c, const uint64_t* x, int l, uint64_t* out);
/* automorphism X -> X^k in EVALUATION form, gather form of OpenFHE's PrecomputeAutoMap */
void orc_automorph(const orc_ctx* c, const uint64_t* x, int nlimbs, uint32_t k, uint64_t* out);
uint32_t orc_rot_index(const orc_ctx* c, int r); /* 5^r mod 2N (r may be negative) */

/* ---- ciphertext-level ops (ct = 2 or 3 components over Q_l) ---- */
/* ks = ModDown(<ModUp(c2), evk>); out = (c0 + ks0, c1 + ks1) */
void orc_relinearize(const orc_ctx* c, const uint64_t* ct3, int l, const uint64_t* key, uint64_t* out2);
void orc_eval_mult(const orc_ctx* c, const uint64_t* ct1, const uint64_t* ct2, int l, const uint64_t* key, uint64_t* out2);
void orc_rescale_ct(const orc_ctx* c, const uint64_t* ct, int l, int ncomp, uint64_t* out);
/* bootstrapping ModRaise: tower 0 of ncomp components (l_in + 1 limbs each) lifted to l + 1 limbs */
void orc_mod_raise(const orc_ctx* c, const uint64_t* ct, int l_in, int ncomp, int l, uint64_t* out);
/* hoisted-formulation rotation: ext key switch of c1, fold P*c0, automorph, ModDown */
void orc_eval_rotate(const orc_ctx* c, const uint64_t* ct, int l, const uint64_t* key, uint32_t k, uint64_t* out);
/* rotation that stays in Q_l∪P (no ModDown): out is 2 x (l+1+K) limbs */
void orc_eval_rotate_ext(const orc_ctx* c, const uint64_t* ct, int l, const uint64_t* key, uint32_t k, int add_c0, uint64_t* out_ext);

/* BSGS linear transform in the reference's extended-basis (hoisted) formulation; see the .c file */
/* Rotation-fold sum from a start factor (the cascade: both components take a ModDown at every fold
 * level); keys / ks one per rotation in fold order. */
void orc_accumulate_sum(const orc_ctx* c, const uint64_t* ct, int l, int radix, int stride, int size, int start,
                        const uint64_t* keys, const uint32_t* ks, uint64_t* out);
/* The lazy rotation-fold sum (c0 stays in Q_l∪P across the levels, one final ModDown of c0). */
void orc_accumulate_lazy(const orc_ctx* c, const uint64_t* ct, int l, int radix, int size, const uint64_t* keys,
                         const uint32_t* ks, uint64_t* out);
void orc_linear_transform(const orc_ctx* c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                          const uint64_t* diags, const uint64_t* baby_keys, const uint32_t* baby_k, const uint64_t* giant_key,
                          uint32_t giant_k, const uint64_t* off_key, uint32_t off_k, uint64_t* out);
/* the same linear map with independent giant-step rotations (giant_keys: gStep-1 keys, entry j-1
 * for j*bStep*stride): this build's own sequence for its bootstrap's transforms */
void orc_linear_transform_flat(const orc_ctx* c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                               const uint64_t* diags, const uint64_t* baby_keys, const uint32_t* baby_k, const uint64_t* giant_keys,
                               const uint32_t* giant_k, const uint64_t* off_key, uint32_t off_k, uint64_t* out);

/* ---- key generation / encryption (seeded, mt19937_64) ---- */
/* ternary secret, EVAL over Q∪P (L1+K limbs); hamming weight h (0 = uniform ternary) */
void orc_gen_secret(const orc_ctx* c, uint64_t seed, int h, uint64_t* s);
/* switching key from s_old (EVAL over Q, L1 limbs) to s_new (EVAL over Q∪P) */
void orc_gen_switch_key(const orc_ctx* c, const uint64_t* s_old, const uint64_t* s_new, uint64_t seed, uint64_t* key);
void orc_gen_relin_key(const orc_ctx* c, const uint64_t* s, uint64_t seed, uint64_t* key);
void orc_gen_rot_key(const orc_ctx* c, const uint64_t* s, uint32_t k, uint64_t seed, uint64_t* key);
/* CKKS encode of n complex slots (re,im interleaved) at scale, over Q_l (EVAL) */
void orc_encode(const orc_ctx* c, const double* slots, int n, double scale, int l, uint64_t* out);
/* the same polynomial over Q_l ∪ P (the extended-basis plaintext form of the linear transform) */
void orc_encode_ext(const orc_ctx* c, const double* slots, int n, double scale, int l, uint64_t* out);
/* decode a plaintext poly over Q_l (EVAL) into n complex slots */
void orc_decode(const orc_ctx* c, const uint64_t* m, int l, int n, double scale, double* slots);
void orc_encrypt(const orc_ctx* c, const uint64_t* m, int l, const uint64_t* s, uint64_t seed, uint64_t* ct);
void orc_decrypt(const orc_ctx* c, const uint64_t* ct, int l, int ncomp, const uint64_t* s, uint64_t* m);
/* public-key encryption with pk = (b, a) over Q (L1 limbs each, b = -a s + e):
 * v ternary, c0 = b v + e0 + m, c1 = a v + e1 (draw order v, e0, e1) */
void orc_encrypt_pk(const orc_ctx* c, const uint64_t* m, int l, const uint64_t* pk_b, const uint64_t* pk_a, uint64_t seed, uint64_t* ct);

/* ---- random data (uniform limbs) ---- */
void orc_random_poly(const orc_ctx* c, uint64_t seed, const int* pids, int nlimbs, uint64_t* out);

/* ---- mt19937_64 reference stream (to check the generator itself) ---- */
uint64_t orc_mt_first(uint64_t seed, int skip);

/* ---- threads used by the parallel loops (OpenMP) ---- */
int orc_num_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
