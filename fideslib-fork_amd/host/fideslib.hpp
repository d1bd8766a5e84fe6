// fideslib — host C++ facade of the MI355X RNS-CKKS engine.
//
// Mirrors the reference's public surface for the evaluation path
// (api/CryptoContext.hpp:32-691, api/Ciphertext.hpp, api/CCParams.hpp:37-62): a CryptoContextImpl
// built from CCParams, KeyGen / EvalMultKeyGen / EvalRotateKeyGen, CKKS packed encoding,
// Encrypt / Decrypt, and the operations EvalAdd / EvalSub / EvalMult (ciphertext, plaintext and
// scalar operands), Relinearize, Rescale, EvalRotate, EvalFastRotation (hoisted), EvalRotateMany
// and LinearTransform. Every device operation goes through the C ABI of include/fideslib_hip.h;
// this layer owns the CKKS metadata (level, scale, noise degree, slots) the way the reference's
// CiphertextImpl does (src/CKKS/Ciphertext.cpp:163-237, 569-607) and the scalar CRT encodings of
// ElemForEvalMult / ElemForEvalAddOrSub (src/CKKS/Context.cu:669-863).
//
// Scope: the FIXEDMANUAL (the caller rescales) and FLEXIBLEAUTO (per-level scaling factors,
// rescaling and level / depth alignment done by the operations) scaling techniques, HYBRID key
// switching, uniform-ternary or sparse-ternary secrets. Errors are thrown as std::runtime_error,
// the reference facade's contract.
#pragma once
#include <complex>
#include <cstdint>
#include <iosfwd>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fideslib_hip.h"

namespace fideslib {

enum class ScalingTechnique { FIXEDMANUAL, FLEXIBLEAUTO };
enum class SecretKeyDist { UNIFORM_TERNARY, SPARSE_TERNARY };

// api/CCParams.hpp:37-62 (the setters this path reads)
class CCParams {
 public:
  void SetMultiplicativeDepth(uint32_t d) { depth_ = d; }
  void SetScalingModSize(uint32_t b) { scale_bits_ = b; }
  void SetFirstModSize(uint32_t b) { first_bits_ = b; }
  void SetRingDim(uint32_t n) { ring_dim_ = n; }
  void SetNumLargeDigits(uint32_t d) { dnum_ = d; }
  void SetScalingTechnique(ScalingTechnique t) { tech_ = t; }
  void SetBatchSize(uint32_t s) { batch_ = s; }
  void SetSecretKeyDist(SecretKeyDist d) { dist_ = d; }
  void SetHammingWeight(uint32_t h) { hw_ = h; }
  // Not in OpenFHE's CCParams: the number of special primes (0 = OpenFHE's rule) and the seed
  // of the deterministic sampler (keys, encryption randomness) used by this build.
  void SetNumSpecialPrimes(uint32_t k) { num_special_ = k; }
  void SetSeed(uint64_t s) { seed_ = s; }

  uint32_t depth_ = 1, scale_bits_ = 50, first_bits_ = 60, ring_dim_ = 1 << 13, dnum_ = 1, batch_ = 0, hw_ = 32;
  uint32_t num_special_ = 0;
  uint64_t seed_ = 1;
  ScalingTechnique tech_ = ScalingTechnique::FIXEDMANUAL;
  SecretKeyDist dist_ = SecretKeyDist::UNIFORM_TERNARY;
};

class CryptoContextImpl;

// Owner of the native context: destroyed when the CryptoContextImpl and every buffer allocated
// from it are gone, so ciphertexts may outlive the CryptoContext object.
struct NativeContext {
  fh_context* ctx = nullptr;
  explicit NativeContext(fh_context* c) : ctx(c) {}
  NativeContext(const NativeContext&) = delete;
  NativeContext& operator=(const NativeContext&) = delete;
  ~NativeContext();
};

// A device-resident buffer from the engine's pool.
struct DeviceBuffer {
  std::shared_ptr<NativeContext> owner;
  uint64_t* ptr = nullptr;
  size_t words = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const std::shared_ptr<NativeContext>& c, size_t w);
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer();
};

// Ciphertext: ncomp components over Q_l (l+1 limbs each), device resident.
class CiphertextImpl {
 public:
  std::shared_ptr<DeviceBuffer> data;
  uint32_t towers = 0;         // l + 1
  uint32_t ncomp = 2;
  uint32_t noiseScaleDeg = 1;
  double scalingFactor = 1.0;
  uint32_t slots = 0;
  // extended form (EvalFastRotationExt): each component over Q_l ∪ P (towers + K limbs), for
  // KeySwitchDown only
  bool extended = false;
  // OpenFHE level: towers consumed from the top of the chain
  uint32_t GetLevel(uint32_t L1) const { return L1 - towers; }
  uint32_t GetNoiseScaleDeg() const { return noiseScaleDeg; }
  double GetScalingFactor() const { return scalingFactor; }
  uint32_t GetSlots() const { return slots; }
  size_t NumberCiphertextElements() const { return ncomp; }
};

// Plaintext: one polynomial over Q_l (or Q_l ∪ P when extended), plus the encoded values.
class PlaintextImpl {
 public:
  std::shared_ptr<DeviceBuffer> data;
  uint32_t towers = 0;
  bool extended = false;
  uint32_t noiseScaleDeg = 1;
  double scalingFactor = 1.0;
  uint32_t slots = 0;
  std::vector<std::complex<double>> values;
  const std::vector<std::complex<double>>& GetCKKSPackedValue() const { return values; }
  void SetLength(size_t n) { values.resize(n); }
};

class PrivateKeyImpl {
 public:
  std::shared_ptr<DeviceBuffer> s;   // EVAL over Q ∪ P
};
class PublicKeyImpl {
 public:
  std::shared_ptr<DeviceBuffer> b, a;  // over Q: b = -a s + e
};

using Ciphertext = std::shared_ptr<CiphertextImpl>;
using Plaintext = std::shared_ptr<PlaintextImpl>;
using PrivateKey = std::shared_ptr<PrivateKeyImpl>;
using PublicKey = std::shared_ptr<PublicKeyImpl>;
struct KeyPair {
  PublicKey publicKey;
  PrivateKey secretKey;
};

// one LinearTransform of the factored CoeffsToSlots / SlotsToCoeffs: diagonal k multiplies the
// slots rotated by offset + k stride (values before the BSGS pre-rotation)
struct BootstrapGroup {
  uint32_t rowSize = 0, stride = 1, offset = 0;
  std::vector<std::vector<std::complex<double>>> diags;
};

// the precomputation of EvalBootstrapSetup
struct BootstrapConstants {
  uint32_t slots;
  uint32_t correction;        // input scaled by 2^-correction, output by 2^correction
  double K;                   // EvalMod input range
  int r;                      // double-angle steps
  double pre;                 // power of two applied after the raise
  double scale_fix;           // prod(q_r / scale) over the raise's rescales, folded into the first CoeffsToSlots group
  std::vector<double> cheb;   // Chebyshev coefficients of cos((2 pi K y - pi/2) / 2^r) on [-1, 1]
  std::vector<BootstrapGroup> cts, stc;  // in the order they are applied
};

class CryptoContextImpl;
using CryptoContext = std::shared_ptr<CryptoContextImpl>;
CryptoContext GenCryptoContext(const CCParams& params);

// api/Serialize.hpp:10-15
enum SerType { BINARY = 0, JSON = 1 };

class CryptoContextImpl {
 public:
  explicit CryptoContextImpl(const CCParams& p);
  ~CryptoContextImpl();

  uint32_t GetRingDimension() const { return N_; }
  uint32_t GetCyclotomicOrder() const { return 2 * N_; }
  uint32_t GetNumQ() const { return L1_; }
  const std::vector<uint64_t>& GetModuli() const { return moduli_; }
  fh_context* Native() const { return ctx_; }

  // ---- keys (api/CryptoContext.hpp:84-88)
  KeyPair KeyGen();
  void EvalMultKeyGen(const PrivateKey& sk);
  void EvalRotateKeyGen(const PrivateKey& sk, const std::vector<int32_t>& steps);
  void EvalConjugateKeyGen(const PrivateKey& sk);
  uint32_t FindAutomorphismIndex(int32_t index) const;

  // ---- encoding (api/CryptoContext.hpp:140-172)
  Plaintext MakeCKKSPackedPlaintext(const std::vector<std::complex<double>>& v, size_t noiseScaleDeg = 1, uint32_t level = 0,
                                    uint32_t slots = 0);
  Plaintext MakeCKKSPackedPlaintext(const std::vector<double>& v, size_t noiseScaleDeg = 1, uint32_t level = 0, uint32_t slots = 0);
  Plaintext MakeCKKSPackedPlaintextExtended(const std::vector<std::complex<double>>& v, size_t noiseScaleDeg = 1,
                                            uint32_t level = 0, uint32_t slots = 0);

  // ---- encryption (api/CryptoContext.hpp:285-288)
  Ciphertext Encrypt(const PrivateKey& sk, const Plaintext& pt);
  Ciphertext Encrypt(const PublicKey& pk, const Plaintext& pt);
  Plaintext Decrypt(const PrivateKey& sk, const Ciphertext& ct);

  // ---- arithmetic (api/CryptoContext.hpp:306-365)
  Ciphertext EvalNegate(const Ciphertext& ct);
  Ciphertext EvalAdd(const Ciphertext& a, const Ciphertext& b);
  Ciphertext EvalAdd(const Ciphertext& a, const Plaintext& pt);
  Ciphertext EvalAdd(const Ciphertext& a, double scalar);
  Ciphertext EvalSub(const Ciphertext& a, const Ciphertext& b);
  Ciphertext EvalSub(const Ciphertext& a, const Plaintext& pt);
  Ciphertext EvalSub(const Ciphertext& a, double scalar);
  Ciphertext EvalMult(const Ciphertext& a, const Ciphertext& b);
  Ciphertext EvalMult(const Ciphertext& a, const Plaintext& pt);
  Ciphertext EvalMult(const Ciphertext& a, double scalar);
  Ciphertext EvalMultNoRelin(const Ciphertext& a, const Ciphertext& b);
  Ciphertext Relinearize(const Ciphertext& ct);
  Ciphertext EvalSquare(const Ciphertext& ct);
  Ciphertext EvalAddMany(const std::vector<Ciphertext>& cts);
  // sum_k weights[k] * cts[k] at the lowest input level, scale degree + 1, not rescaled (OpenFHE's
  // EvalLinearWSum; the reference's evalLinearWSumMutable, src/CKKS/Ciphertext.cuh:573): one fused
  // kernel per 16 inputs, the residues those of EvalMult(ct, w) summed
  Ciphertext EvalLinearWSum(const std::vector<Ciphertext>& cts, const std::vector<double>& weights);
  // sum_k w[k] cts[k] + constant at the lowest input level in one kernel: the residues of the
  // additions, negations and integer products it replaces (equal scale degrees, no scale change)
  Ciphertext IntegerCombination(const std::vector<Ciphertext>& cts, const std::vector<int64_t>& w, double constant);
  Ciphertext Rescale(const Ciphertext& ct);
  Ciphertext LevelReduce(const Ciphertext& ct, uint32_t levels);

  // ---- rotations (api/CryptoContext.hpp:367-404)
  Ciphertext EvalRotate(const Ciphertext& ct, int32_t index);
  Ciphertext EvalConjugate(const Ciphertext& ct);
  std::shared_ptr<void> EvalFastRotationPrecompute(const Ciphertext& ct);
  Ciphertext EvalFastRotation(const Ciphertext& ct, int32_t index, uint32_t m, const std::shared_ptr<void>& precomp);
  std::vector<Ciphertext> EvalFastRotation(const Ciphertext& ct, const std::vector<int32_t>& indices, uint32_t m,
                                           const std::shared_ptr<void>& precomp);
  std::vector<Ciphertext> EvalRotateMany(const std::vector<Ciphertext>& cts, int32_t index);
  // rotation from the precomputed digits stopping before ModDown: components over Q_l ∪ P, P*c0
  // folded when addFirst (api/CryptoContext.hpp:372-375); KeySwitchDown finishes it
  Ciphertext EvalFastRotationExt(const Ciphertext& ct, int32_t index, const std::shared_ptr<void>& precomp, bool addFirst);
  std::vector<Ciphertext> EvalFastRotationExt(const Ciphertext& ct, const std::vector<int32_t>& indices, const std::shared_ptr<void>& precomp,
                                              bool addFirst);
  Ciphertext KeySwitchDown(const Ciphertext& ctExt);

  // ---- linear transform (api/CryptoContext.hpp:460-517): diagonals built with
  // MakeCKKSPackedPlaintextExtended at the ciphertext's level, D_k pre-rotated by
  // -(bStep*(k/bStep)*stride + offset); result at the input level, noise degree + 1.
  Ciphertext LinearTransform(const Ciphertext& ct, int rowSize, int bStep, const std::vector<Plaintext>& diagonals, int stride = 1,
                             int offset = 0);
  // the rowSize diagonals back to back in one device buffer (the layout the transform reads), for
  // a transform applied repeatedly: the overload below skips the per-call packing
  struct PackedDiagonals {
    std::shared_ptr<DeviceBuffer> data;
    uint32_t rowSize = 0, towers = 0, noiseScaleDeg = 1;
    double scalingFactor = 1.0;
  };
  PackedDiagonals PackDiagonals(const std::vector<Plaintext>& diagonals, int rowSize) const;
  Ciphertext LinearTransform(const Ciphertext& ct, int bStep, const PackedDiagonals& diagonals, int stride = 1, int offset = 0);
  // a transform whose diagonals are resident, prepared once and reused (the reference's
  // PrepareLinearTransform / LinearTransformInPlace, api/CryptoContext.hpp:576-601,
  // api/PreparedLinearTransform.hpp:33-66): the diagonals are packed into the transform's layout at
  // preparation; each call reads them in place
  struct PreparedLinearTransformImpl {
    int rowSize = 0, bStep = 0, stride = 1, offset = 0;
    PackedDiagonals diagonals;                 // the first (or only) set
    std::vector<PackedDiagonals> more;         // further sets of PrepareLinearTransformMany
  };
  using PreparedLinearTransform = std::shared_ptr<const PreparedLinearTransformImpl>;
  PreparedLinearTransform PrepareLinearTransform(const std::vector<Plaintext>& diagonals, int rowSize, int bStep, int stride = 1,
                                                 int offset = 0) const;
  Ciphertext LinearTransform(const Ciphertext& ct, const PreparedLinearTransform& prepared);
  void LinearTransformInPlace(Ciphertext& ct, const PreparedLinearTransform& prepared);

  // ---- rotation-fold sum (AccumulateSum, api/CryptoContext.hpp:456-458; CudaEngine::accumulateSum,
  // api/engine/cuda/CudaEngine.cpp:755-784): slot j of the result is the sum over i < slots/start
  // of slot j + i*start*stride, folded at radix kAccumulateSumRadix. The two- and three-argument
  // forms run the lazy fold (c0 stays in Q_l ∪ P across the fold levels and takes one ModDown at
  // the end, src/CKKS/AccumulateBroadcast.cu:122-160) and keep the slot count; the start form runs
  // the cascade (a ModDown of both components per level, :42-84) and sets the slot count to
  // stride*start when slots*stride equals it (:82-83). Needs the rotation keys
  // GetAccumulateSumRotationIndices lists.
  static constexpr int kAccumulateSumRadix = 4;
  Ciphertext AccumulateSum(const Ciphertext& ct, int slots, int stride = 1);
  void AccumulateSumInPlace(Ciphertext& ct, int slots, int stride = 1);
  void AccumulateSumInPlace(Ciphertext& ct, int slots, int stride, int start);
  static std::vector<int32_t> GetAccumulateSumRotationIndices(int slots, int stride = 1, int start = 1);

  // ---- polynomial evaluation (OpenFHE EvalChebyshevSeries / EvalChebyshevCoefficients):
  // c_0 / 2 + sum_{k>=1} c_k T_k(y), y = (2x - a - b) / (b - a); the coefficients of f on [a, b]
  // at degree d are c_k = 2/(d+1) sum_j f(x_j) cos(pi k (j + 1/2) / (d+1)).
  Ciphertext EvalChebyshevSeries(const Ciphertext& ct, const std::vector<double>& coeffs, double a, double b);
  static std::vector<double> EvalChebyshevCoefficients(double (*f)(double, const void*), const void* arg, double a, double b,
                                                       uint32_t degree);

  // ---- bootstrapping (api/CryptoContext.hpp:90-107, 423-424; src/CKKS/Bootstrap.cu:437-648).
  // levelBudget {b1, b2}: CoeffsToSlots and SlotsToCoeffs are b1 and b2 BSGS linear transforms
  // over the `slots` slots (groups of the special FFT's butterfly stages). EvalBootstrapKeyGen
  // needs EvalMultKeyGen to have run.
  void EvalBootstrapSetup(const std::vector<uint32_t>& levelBudget = {1, 1}, const std::vector<uint32_t>& dim1 = {0, 0},
                          uint32_t slots = 0, uint32_t correctionFactor = 0);
  void EvalBootstrapKeyGen(const PrivateKey& sk, uint32_t slots);
  Ciphertext EvalBootstrap(const Ciphertext& ct, uint32_t numIterations = 1, uint32_t precision = 0);
  // EvalBootstrap, then the level reduced to outputLevel (OpenFHE level: towers consumed)
  Ciphertext EvalBootstrapToLevel(const Ciphertext& ct, uint32_t outputLevel, uint32_t numIterations = 1, uint32_t precision = 0);

  // ---- in-place and convenience forms of the operations above (api/CryptoContext.hpp:300-456):
  // each replaces its first argument with the result of the corresponding operation
  void EvalAddInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalAdd(a, b); }
  void EvalAddInPlace(Ciphertext& a, const Plaintext& pt) { a = EvalAdd(a, pt); }
  void EvalAddInPlace(Ciphertext& a, double scalar) { a = EvalAdd(a, scalar); }
  void EvalSubInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalSub(a, b); }
  void EvalSubInPlace(Ciphertext& a, const Plaintext& pt) { a = EvalSub(a, pt); }
  void EvalSubInPlace(Ciphertext& a, double scalar) { a = EvalSub(a, scalar); }
  void EvalMultInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalMult(a, b); }
  void EvalMultInPlace(Ciphertext& a, const Plaintext& pt) { a = EvalMult(a, pt); }
  void EvalMultInPlace(Ciphertext& a, double scalar) { a = EvalMult(a, scalar); }
  void EvalNegateInPlace(Ciphertext& a) { a = EvalNegate(a); }
  void EvalSquareInPlace(Ciphertext& a) { a = EvalSquare(a); }
  void RelinearizeInPlace(Ciphertext& a) { a = Relinearize(a); }
  void RescaleInPlace(Ciphertext& a) { a = Rescale(a); }
  void EvalRotateInPlace(Ciphertext& a, int32_t index) { a = EvalRotate(a, index); }
  void EvalChebyshevSeriesInPlace(Ciphertext& a, const std::vector<double>& coeffs, double lo, double hi) {
    a = EvalChebyshevSeries(a, coeffs, lo, hi);
  }
  void EvalBootstrapInPlace(Ciphertext& a, uint32_t numIterations = 1, uint32_t precision = 0) { a = EvalBootstrap(a, numIterations, precision); }
  void EvalBootstrapToLevelInPlace(Ciphertext& a, uint32_t outputLevel, uint32_t numIterations = 1, uint32_t precision = 0) {
    a = EvalBootstrapToLevel(a, outputLevel, numIterations, precision);
  }
  // OpenFHE Compress: rescale down to noise degree noiseScaleDeg, then keep towersLeft towers
  Ciphertext Compress(const Ciphertext& ct, uint32_t towersLeft = 1, size_t noiseScaleDeg = 1);
  // the plaintexts are encoded on the device in any case; the Many form encodes the group in one
  // batched call (api/CryptoContext.hpp:240-281)
  Plaintext MakeCKKSPackedPlaintextDevice(const std::vector<double>& v, size_t noiseScaleDeg = 1, uint32_t level = 0, uint32_t slots = 0) {
    return MakeCKKSPackedPlaintext(v, noiseScaleDeg, level, slots);
  }
  std::vector<Plaintext> MakeCKKSPackedPlaintextDeviceMany(const std::vector<std::vector<double>>& values, size_t noiseScaleDeg = 1,
                                                           uint32_t level = 0, uint32_t slots = 0);
  // rotation indices a BSGS transform needs keys for ({i stride : 1 <= i <= bStep}, and offset
  // when nonzero) and the rotation each diagonal k is stored with (-bStep floor(k / bStep) stride
  // - offset); api/CryptoContext.hpp:478-500
  static std::vector<int> GetLinearTransformRotationIndices(int bStep, int stride = 1, int offset = 0);
  static std::vector<int> GetLinearTransformPlaintextRotationIndices(int rowSize, int bStep, int stride = 1, int offset = 0);
  // several diagonal sets applied to one ciphertext (api/CryptoContext.hpp:544-608); one transform
  // per set here
  PreparedLinearTransform PrepareLinearTransformMany(const std::vector<std::vector<Plaintext>>& diagonalSets, int rowSize, int bStep,
                                                     int stride = 1, int offset = 0) const;
  std::vector<Ciphertext> LinearTransformMany(const Ciphertext& ct, const PreparedLinearTransform& prepared);
  // levels one bootstrap consumes after the raise (prescale, CoeffsToSlots, EvalMod, SlotsToCoeffs)
  uint32_t GetBootstrapDepth() const;
  BootstrapConstants GetBootstrapConstants() const;

  // ---- raw device residues (Ciphertext::load / store, src/CKKS/Ciphertext.cpp:415-465): ncomp
  // components of `towers` limbs in EVALUATION order, copied device to device
  Ciphertext LoadCiphertext(const uint64_t* dev, uint32_t towers, uint32_t ncomp, uint32_t noiseScaleDeg, double scale, uint32_t slots);
  void StoreCiphertext(const Ciphertext& ct, uint64_t* dev) const;
  // key handles for C ABI entry points that take them (the serving loop, batches)
  const fh_key* RotationKeyHandle(int32_t index) const { return rotKey(galoisFor(index)); }
  const fh_key* RelinKeyHandle() const { return relin_; }

  // ---- serialization (api/CryptoContext.hpp:110-127, api/Serialize.hpp; this build's own binary
  // format, since OpenFHE's cereal archives are not available): evaluation keys of this context
  // to / from a stream. BINARY only; JSON returns false.
  bool SerializeEvalMultKey(std::ostream& ser, const SerType& sertype) const;
  bool DeserializeEvalMultKey(std::istream& ser, const SerType& sertype);
  bool SerializeEvalAutomorphismKey(std::ostream& ser, const SerType& sertype) const;
  bool DeserializeEvalAutomorphismKey(std::istream& ser, const SerType& sertype);
  const CCParams& GetParams() const { return params_; }
  // a device buffer owned like the context's own (for objects rebuilt from host data)
  std::shared_ptr<DeviceBuffer> AllocateDevice(size_t words) const;

  // ---- scalar CRT encodings (src/CKKS/Context.cu:669-863), exposed for tests
  std::vector<uint64_t> ElemForEvalMult(uint32_t towers, double operand) const;
  std::vector<uint64_t> ElemForEvalAddOrSub(uint32_t towers, double operand, uint32_t noiseDeg) const;
  // the same encodings at an explicit scaling factor (the bootstrap's exact-scale EvalMod)
  std::vector<uint64_t> ElemForEvalMultAt(uint32_t towers, double operand, double scFactor) const;
  std::vector<uint64_t> ElemForEvalAddOrSubAt(uint32_t towers, double operand, uint32_t noiseDeg, double scFactor) const;

  void Synchronize() const;
  double ScalingFactor() const { return scale_; }
  ScalingTechnique GetScalingTechnique() const { return params_.tech_; }
  // OpenFHE's per-level tables (level = towers consumed, OpenFHE's numbering): the scaling factor
  // of a degree-1 ciphertext, its square for degree 2, and the factor a rescale divides by. Under
  // FIXEDMANUAL all three are the nominal 2^scaleModSize; under FLEXIBLEAUTO the top level's factor
  // is the top Q prime, each further level's the previous one squared over the prime dropped, and
  // a rescale divides by that prime (CryptoParametersCKKSRNS, restated by the reference at
  // src/CKKS/openfhe-interface/RawCiphertext.cu:305-317).
  double GetScalingFactorReal(uint32_t level) const;
  double GetScalingFactorRealBig(uint32_t level) const;
  double GetModReduceFactor(uint32_t level) const;

 private:
  void check(int rc) const;
  Ciphertext newCt(uint32_t towers, uint32_t ncomp) const;
  Ciphertext adjusted(const Ciphertext& ct, uint32_t towers) const;
  void encodeInto(const std::vector<std::complex<double>>& v, uint32_t slots, double scale, const std::vector<int32_t>& pids,
                  uint64_t* dev) const;
  std::shared_ptr<DeviceBuffer> smallToEval(const std::vector<int64_t>& coeffs, const std::vector<int32_t>& pids) const;
  std::shared_ptr<DeviceBuffer> uniform(const std::vector<int32_t>& pids);
  std::vector<int64_t> gaussian();
  fh_key* switchKey(const DeviceBuffer& s_old, const DeviceBuffer& s_new);
  const fh_key* rotKey(uint32_t galois) const;
  // the transform with every giant-step multiple rotated by its own key (fh_linear_transform_flat):
  // the bootstrap's CtS / StC transforms, whose keys EvalBootstrapKeyGen generates
  Ciphertext LinearTransformFlat(const Ciphertext& ct, int bStep, const PackedDiagonals& diagonals, int stride, int offset);
  Ciphertext accumulate(const Ciphertext& ct, int slots, int stride, int start);
  uint32_t galoisFor(int32_t index) const;
  void automorphismKey(const PrivateKey& sk, uint32_t galois);

  // FLEXIBLEAUTO metadata algebra (OpenFHE's LeveledSHECKKSRNS, as the reference restates it in
  // src/CKKS/Ciphertext.cpp:624-658, 1917-2070 and src/CKKS/Context.cu:669-863)
  bool flex() const { return params_.tech_ == ScalingTechnique::FLEXIBLEAUTO && manual_ == 0; }
  double sfReal(uint32_t towers) const { return flex() ? sfReal_[towers - 1] : scale_; }
  double ptScale(uint32_t towers, size_t noiseScaleDeg) const;      // the scale a plaintext is encoded at
  Ciphertext modReduce(const Ciphertext& ct) const;                 // rescale with the technique's metadata
  Ciphertext multCore(const Ciphertext& ct, double operand) const;  // EvalMultCore(ct, double): degree + 1
  Ciphertext adjustDown(const Ciphertext& c1, uint32_t towers2, uint32_t deg2, double scale2) const;
  void adjustLevelsAndDepth(Ciphertext& a, Ciphertext& b) const;
  void adjustLevelsAndDepthToOne(Ciphertext& a, Ciphertext& b) const;
  Ciphertext morph(const Plaintext& pt) const;                      // a plaintext as a 1-component ciphertext
  Ciphertext mulPlain(const Ciphertext& a, const Ciphertext& p) const;
  Ciphertext addPlain(const Ciphertext& a, const Ciphertext& p, int op) const;
  // FIXEDMANUAL semantics for the duration of a scope (the bootstrap and Chebyshev internals manage
  // their rescales explicitly, with the scales they track)
  struct ManualScope {
    const CryptoContextImpl& cc;
    explicit ManualScope(const CryptoContextImpl& c) : cc(c) { ++cc.manual_; }
    ~ManualScope() { --cc.manual_; }
  };
  mutable int manual_ = 0;
  std::vector<double> sfReal_, sfBig_, mrf_;   // by towers - 1

  std::shared_ptr<NativeContext> native_;
  fh_context* ctx_ = nullptr;
  uint32_t N_ = 0, L1_ = 0, K_ = 0, dnum_ = 0, alpha_ = 0;
  double scale_ = 0;
  std::vector<uint64_t> moduli_;
  std::vector<uint64_t> pmodq_;
  CCParams params_;
  struct Rng;
  std::unique_ptr<Rng> rng_;
  fh_key* relin_ = nullptr;
  std::map<uint32_t, fh_key*> rot_;

  // bootstrapping state (bootstrap.cpp)
  friend struct BootstrapOps;
  fh_batch* batch_ = nullptr;  // the bootstrap's batched products (created on first use)
  Ciphertext MultByMonomial(const Ciphertext& ct, uint32_t power);
  Ciphertext MultByInteger(const Ciphertext& ct, uint64_t v);
  PackedDiagonals groupDiagonals(const BootstrapGroup& g, uint32_t n, uint32_t level);
  std::shared_ptr<struct BootstrapState> boot_;
  std::map<std::pair<uint32_t, uint32_t>, std::shared_ptr<DeviceBuffer>> monomials_;
};

namespace Serial {
// api/Serialize.hpp:23-29. Keys and ciphertexts are read back into an existing context (the
// reference resolves it through OpenFHE's context registry): pass the context deserialized first.
bool SerializeToFile(const std::string& filename, const CryptoContext& obj, const SerType& sertype);
bool DeserializeFromFile(const std::string& filename, CryptoContext& obj, const SerType& sertype);
bool SerializeToFile(const std::string& filename, const PublicKey& obj, const SerType& sertype);
bool DeserializeFromFile(const std::string& filename, PublicKey& obj, const SerType& sertype, const CryptoContext& cc);
bool SerializeToFile(const std::string& filename, const PrivateKey& obj, const SerType& sertype);
bool DeserializeFromFile(const std::string& filename, PrivateKey& obj, const SerType& sertype, const CryptoContext& cc);
bool SerializeToFile(const std::string& filename, const Ciphertext& obj, const SerType& sertype);
bool DeserializeFromFile(const std::string& filename, Ciphertext& obj, const SerType& sertype, const CryptoContext& cc);
}  // namespace Serial

}  // namespace fideslib
