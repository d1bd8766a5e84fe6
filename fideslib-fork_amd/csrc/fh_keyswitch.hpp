// Rescale, hybrid key switching, automorphism and the ciphertext-level
// compositions (fh_keyswitch.hip). Buffers are device pointers in the layout of
// include/fideslib_hip.h; keys are device copies of [dnum][b, a][L1+K][N].
#pragma once
#include "fh_ops.hpp"

namespace fh {

// digits: beta x (l+1+K) limbs (EVAL)
void modup(Context& c, const uint64_t* x, int l, uint64_t* digits, hipStream_t st);
// npolys inputs x_stride words apart (0 = packed) -> npolys packed digit blocks. copy_own false
// leaves each digit's own limbs unwritten: the key inner product then reads them from x (own).
void modup_batch(Context& c, const uint64_t* x, int l, int npolys, uint64_t x_stride, uint64_t* digits, hipStream_t st,
                 bool copy_own = true, bool rows = true);
// rows false above leaves the ModUp targets after the column pass of their forward NTT; rows_ksk
// then runs their row stages fused with the key inner product of nb rotations (item b: digits at
// digits + b dig_stride, own limbs at own + b own_stride, c0 at c0s + b c0_stride folded as P c0
// when add_c0, key keys[b], automorphism galois[b]) and writes item b's extended result (2 x
// (l+1+K) limbs) at out + b out_stride. rows_ksk_ok: the context's split and tunables allow it.
bool rows_ksk_ok(const Context& c);
void rows_ksk(Context& c, int nb, const uint64_t* digits, uint64_t dig_stride, const uint64_t* c0s, uint64_t c0_stride, bool add_c0,
              const uint64_t* own, uint64_t own_stride, int l, const uint64_t* const* keys, const uint32_t* galois, uint64_t* out,
              uint64_t out_stride, hipStream_t st);
// B rotations sharing nothing but the level; ciphertexts with equal (key, galois) share the key stream
void rotate_batch(Context& c, int B, const uint64_t* cts, uint64_t ct_stride, int l, const uint64_t* const* keys,
                  const uint32_t* galois, bool add_c0, bool do_moddown, uint64_t* out, hipStream_t st);
// n rotations of one ciphertext sharing a single ModUp (keys[i] / galois[i] per rotation)
void rotate_hoisted(Context& c, const uint64_t* ct, int l, int n, const uint64_t* const* keys, const uint32_t* galois, bool add_c0,
                    bool do_moddown, uint64_t* out, hipStream_t st);
// out0/out1: (l+1+K) limbs. c0 (optional, l+1 limbs) is folded as P*c0 into out0's Q limbs;
// galois != 1 writes both outputs through the automorphism X -> X^galois.
// c0_ext: c0 is already P-scaled over Q_l ∪ P (l+1+K limbs) and is added on every limb.
// own (optional): the ModUp input (EVAL, l+1 limbs), read for each digit's own limbs.
// accumulate: out0/out1 += the result instead of out0/out1 = the result.
void ksk_inner(Context& c, const uint64_t* digits, int l, const uint64_t* key, uint64_t* out0, uint64_t* out1,
               const uint64_t* c0, uint32_t galois, hipStream_t st, bool c0_ext = false, const uint64_t* own = nullptr,
               bool accumulate = false);
// BSGS linear transform in the extended-basis formulation (fh_lt.hip); keys[i] for the baby step
// i*stride (null where i*stride = 0 mod slots), giant/offset keys may be null when unused.
void linear_transform(Context& c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                      const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal, const uint64_t* giant_key,
                      uint32_t giant_gal, const uint64_t* off_key, uint32_t off_gal, uint64_t* out, hipStream_t st);
// The same linear map with the giant steps as independent rotations (giant_keys[j-1] /
// giant_gal[j-1] rotate the j-th partial sum by j*bStep*stride; entries whose multiple is 0 mod
// slots are unused): one batched key switch for all of them instead of a sequential Horner chain.
// Its own sequence (the same map, different rounding than linear_transform).
void linear_transform_flat(Context& c, const uint64_t* ct, int l, int rowSize, int bStep, int stride, int offset, int slots,
                           const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal,
                           const uint64_t* const* giant_keys, const uint32_t* giant_gal, const uint64_t* off_key, uint32_t off_gal,
                           uint64_t* out, hipStream_t st);
// Rotation-fold sums at a power-of-two radix (fh_lt.hip); keys / galois one per fold rotation in
// order; out may alias ct. accumulate_sum: the lazy fold (c0 stays in Q_l ∪ P across the levels and
// takes one ModDown at the end); accumulate_cascade: the fold from factor start with a ModDown of
// both components at every level.
void accumulate_sum(Context& c, const uint64_t* ct, int l, int radix, int size, const uint64_t* const* keys, const uint32_t* galois,
                    uint64_t* out, hipStream_t st);
void accumulate_cascade(Context& c, const uint64_t* ct, int l, int radix, int size, int start, const uint64_t* const* keys,
                        const uint32_t* galois, uint64_t* out, hipStream_t st);
// Key inner products of nb items grouped by (key, galois): item b uses the digits (dig_stride
// apart) and c0 (c0_stride apart) of source b / per_src and writes 2 x (l+1+K) limbs at
// eo + b * out_stride. add_c0 folds c0 as P c0 over Q_l, or (c0_ext) as given over Q_l ∪ P.
// own (optional): source s's ModUp input (EVAL, l+1 limbs) at own + s * own_stride, read for each
// digit's own limbs (digits built with copy_own = false). accumulate: each output += its result.
void ksk_items(Context& c, int nb, int per_src, const uint64_t* digits, uint64_t dig_stride, const uint64_t* c0s, uint64_t c0_stride,
               bool add_c0, bool c0_ext, int l, const uint64_t* const* keys, const uint32_t* galois, uint64_t* eo, uint64_t out_stride,
               hipStream_t st, const uint64_t* own = nullptr, uint64_t own_stride = 0, bool accumulate = false);
// count linear transforms of ciphertexts at level l (ct_stride words apart, 0 = packed) sharing one
// diagonal set and key set; outputs packed (2 x (l+1) limbs each). Keys and diagonals are read once
// per batch instead of once per ciphertext.
void linear_transform_batch(Context& c, int count, const uint64_t* cts, uint64_t ct_stride, int l, int rowSize, int bStep, int stride,
                            int offset, int slots, const uint64_t* diags, const uint64_t* const* baby_keys, const uint32_t* baby_gal,
                            const uint64_t* giant_key, uint32_t giant_gal, const uint64_t* off_key, uint32_t off_gal, uint64_t* out,
                            hipStream_t st);
// npolys polynomials over Q_l ∪ P (x_stride words apart, 0 = packed) -> over Q_l (out_stride apart)
void moddown(Context& c, const uint64_t* x, int l, uint64_t* out, int npolys, uint64_t x_stride, uint64_t out_stride, hipStream_t st);
// ncomp components over Q_l -> over Q_{l-1}
void rescale(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, hipStream_t st);
// Rescale with its work buffer supplied by the caller (rescale_work_words words: the INTT of the top
// limbs, then the correction over l limbs per component); x and out must not overlap. Allocates nothing.
inline size_t rescale_work_words(const Context& c, int l, int ncomp) { return (size_t)ncomp * (l + 1) * c.N; }
void rescale_core(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, uint64_t* work, hipStream_t st);
// fh_noalloc.hip: Rescale through a fixed per-device array (no allocation); noalloc_init looks the
// array's device address up for the current device and records the device in the context.
void noalloc_init(Context& c);
size_t noalloc_scratch_words();
void rescale_noalloc(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, hipStream_t st);
// CKKS bootstrapping ModRaise: tower 0 of ncomp components (of l_in + 1 limbs each) to
// coefficient form, lifted centred to q_0..q_l (SwitchModulus) and transformed back: out holds
// ncomp components of l + 1 limbs.
void mod_raise(Context& c, const uint64_t* x, int l_in, int ncomp, int l, uint64_t* out, hipStream_t st);
void automorph(Context& c, const uint64_t* x, int nlimbs, uint32_t k, uint64_t* out, hipStream_t st);

void relinearize(Context& c, const uint64_t* ct3, int l, const uint64_t* key, uint64_t* out2, hipStream_t st);
void eval_mult(Context& c, const uint64_t* ct1, const uint64_t* ct2, int l, const uint64_t* key, uint64_t* out2, hipStream_t st);
// n products at level l of ciphertexts back to back (2 x limbs1 / limbs2 >= l+1 limbs each, of
// which the first l+1 per component count; ct1 == ct2: squares) with one relinearization key: one
// tensor launch, one ModUp / key inner product / ModDown for all
void eval_mult_many(Context& c, int n, const uint64_t* ct1, int limbs1, const uint64_t* ct2, int limbs2, int l, const uint64_t* key, uint64_t* out2,
                    hipStream_t st);
void rotate_ext(Context& c, const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, bool add_c0, uint64_t* out_ext, hipStream_t st);
void eval_rotate(Context& c, const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, uint64_t* out2, hipStream_t st);

}  // namespace fh
