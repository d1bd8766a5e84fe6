// Host launchers of the evaluation kernels (element-wise, rescale, key
// switching, automorphism) and the ciphertext-level compositions built on them.
#pragma once
#include "fh_kernels.hpp"

namespace fh {

// ---- element-wise (fh_eltwise.hip) ----
// op: 0 add, 1 sub, 2 mul
void ew_binary(const Context& c, int op, const LimbSet& a, const LimbSet& b, const LimbSet& o, hipStream_t st);
// op: 0 add scalar, 1 mul scalar, 2 negate (scalars per limb, may be null for negate)
void ew_unary(const Context& c, int op, const LimbSet& a, const LimbSet& o, const uint64_t* scalars, hipStream_t st);
// out (ncomp x towers limbs, packed) = sum_k weights[k] * in[k] over the first towers limbs of each
// component (in[k]: ncomp components of in_limbs[k] limbs), plus add[] on the components whose bit
// is set in add_mask; weights: [n][towers] residues (host), add: [towers] residues (host) or null.
// n <= kMaxWSum, ncomp <= 8.
constexpr int kMaxWSum = 16;
void ew_wsum(Context& c, int n, const uint64_t* const* in, const uint32_t* in_limbs, const uint64_t* weights, int towers, int ncomp,
             const uint64_t* add, uint32_t add_mask, uint64_t* out, hipStream_t st);
// ct1, ct2: 2 x nlimbs over Q_l; out3: 3 x nlimbs. ct1 == ct2 selects the square form.
// CKKS encoding of `count` plaintexts into the output set (npolys == count; EVALUATION form):
// from slot values (count x slots complex, interleaved re/im, host) with the special inverse FFT on
// the device, or from integer coefficients (count x N, host)
void encode_slots(Context& c, const double* values, int count, uint32_t slots, double scale, const LimbSet& out, hipStream_t st);
void encode_coeffs(Context& c, const int64_t* coeffs, int count, const LimbSet& out, hipStream_t st);
// n tensor products over the first nlimbs limbs of ciphertexts back to back (2 x limbs1 / limbs2
// limbs each; 0 = nlimbs) into n x 3 x nlimbs limbs
void ew_tensor(const Context& c, const uint64_t* ct1, const uint64_t* ct2, uint64_t* out3, int nlimbs, hipStream_t st, int n = 1, int limbs1 = 0,
               int limbs2 = 0);
// in place x -> x 2^64 mod q (resident key-switching keys are kept in Montgomery form)
void to_montgomery(const Context& c, const LimbSet& s, hipStream_t st);
// STREAM copy kernel (bandwidth reference; 16-byte aligned)
void stream_copy(const Context& c, void* dst, const void* src, size_t bytes, hipStream_t st);
// uniform residues mod each limb's prime (synthetic inputs)
void fill_uniform(const Context& c, const LimbSet& s, uint64_t seed, hipStream_t st);

}  // namespace fh
