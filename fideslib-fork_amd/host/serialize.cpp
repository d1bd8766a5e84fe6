// Serialization of the facade's objects (contexts, public and secret keys, ciphertexts, the
// relinearization and rotation keys) in this build's own little-endian binary format. The
// reference delegates to OpenFHE's cereal archives plus a device sidecar (api/Serialize.cpp:17-312,
// api/CryptoContext.hpp:110-127); OpenFHE is not available here, so the archive is a plain record:
//   8-byte magic "FHSER001", 8-byte object tag, then the object's fields.
// A context is stored as its parameters and moduli and rebuilt with GenCryptoContext (the moduli
// must come out equal); keys and ciphertexts store their canonical residues in the EVALUATION
// layout of the C ABI.
#include <cmath>
#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>

#include "fideslib.hpp"

namespace fideslib {

namespace {

constexpr char kMagic[9] = "FHSER001";

void put_raw(std::ostream& os, const void* p, size_t n) { os.write(static_cast<const char*>(p), (std::streamsize)n); }
void put_u32(std::ostream& os, uint32_t v) { put_raw(os, &v, 4); }
void put_u64(std::ostream& os, uint64_t v) { put_raw(os, &v, 8); }
void put_f64(std::ostream& os, double v) { put_raw(os, &v, 8); }
void put_header(std::ostream& os, const char* tag) {
  put_raw(os, kMagic, 8);
  char t[8] = {' ', ' ', ' ', ' ', ' ', ' ', ' ', ' '};
  std::memcpy(t, tag, std::min<size_t>(8, std::strlen(tag)));
  put_raw(os, t, 8);
}

bool get_raw(std::istream& is, void* p, size_t n) { return (bool)is.read(static_cast<char*>(p), (std::streamsize)n); }
bool get_u32(std::istream& is, uint32_t& v) { return get_raw(is, &v, 4); }
bool get_u64(std::istream& is, uint64_t& v) { return get_raw(is, &v, 8); }
bool get_f64(std::istream& is, double& v) { return get_raw(is, &v, 8); }
bool get_header(std::istream& is, const char* tag) {
  char m[8], t[8];
  if (!get_raw(is, m, 8) || !get_raw(is, t, 8) || std::memcmp(m, kMagic, 8) != 0) return false;
  char want[8] = {' ', ' ', ' ', ' ', ' ', ' ', ' ', ' '};
  std::memcpy(want, tag, std::min<size_t>(8, std::strlen(tag)));
  return std::memcmp(t, want, 8) == 0;
}

// device words -> stream, stream -> device words
bool put_device(std::ostream& os, fh_context* ctx, const uint64_t* dev, size_t words) {
  std::vector<uint64_t> h(words);
  if (fh_copy_d2h(ctx, h.data(), dev, words * 8) != FH_OK) return false;
  put_u64(os, words);
  put_raw(os, h.data(), words * 8);
  return (bool)os;
}
// Every residue must be canonical for its limb's prime: the data is blocks of `block` limbs whose
// limb j is over moduli[j] (a component of a ciphertext, a key's Q ∪ P block). Non-canonical words
// would feed the lazy Shoup / Montgomery kernels values outside their input ranges.
bool canonical(const std::vector<uint64_t>& h, const std::vector<uint64_t>& moduli, size_t block, size_t N) {
  if (block == 0 || block > moduli.size()) return false;
  for (size_t w = 0; w < h.size(); ++w)
    if (h[w] >= moduli[(w / N) % block]) return false;
  return true;
}

std::shared_ptr<DeviceBuffer> get_device(std::istream& is, const CryptoContextImpl& cc, size_t expect, size_t block) {
  uint64_t words = 0;
  if (!get_u64(is, words) || words != expect) return nullptr;
  std::vector<uint64_t> h(words);
  if (!get_raw(is, h.data(), words * 8)) return nullptr;
  if (!canonical(h, cc.GetModuli(), block, cc.GetRingDimension())) return nullptr;
  auto d = cc.AllocateDevice(words);
  if (fh_copy_h2d(cc.Native(), d->ptr, h.data(), words * 8) != FH_OK) return nullptr;
  return d;
}

size_t key_words(const CryptoContextImpl& cc) {
  return (size_t)cc.GetParams().dnum_ * 2 * cc.GetModuli().size() * cc.GetRingDimension();
}

bool put_key(std::ostream& os, fh_context* ctx, const fh_key* k, size_t words) {
  std::vector<uint64_t> h(words);
  if (fh_key_export(ctx, k, h.data()) != FH_OK) return false;
  put_u64(os, words);
  put_raw(os, h.data(), words * 8);
  return (bool)os;
}
fh_key* get_key(std::istream& is, const CryptoContextImpl& cc, fh_context* ctx, size_t expect) {
  uint64_t words = 0;
  if (!get_u64(is, words) || words != expect) return nullptr;
  std::vector<uint64_t> h(words);
  if (!get_raw(is, h.data(), words * 8)) return nullptr;
  if (!canonical(h, cc.GetModuli(), cc.GetModuli().size(), cc.GetRingDimension())) return nullptr;
  fh_key* k = nullptr;
  if (fh_key_create(ctx, h.data(), &k) != FH_OK) return nullptr;
  return k;
}

}  // namespace

std::shared_ptr<DeviceBuffer> CryptoContextImpl::AllocateDevice(size_t words) const { return std::make_shared<DeviceBuffer>(native_, words); }

bool CryptoContextImpl::SerializeEvalMultKey(std::ostream& ser, const SerType& sertype) const {
  if (sertype != SerType::BINARY || !relin_) return false;
  put_header(ser, "MULTKEY");
  return put_key(ser, ctx_, relin_, key_words(*this));
}

bool CryptoContextImpl::DeserializeEvalMultKey(std::istream& ser, const SerType& sertype) {
  if (sertype != SerType::BINARY || !get_header(ser, "MULTKEY")) return false;
  fh_key* k = get_key(ser, *this, ctx_, key_words(*this));
  if (!k) return false;
  if (relin_) fh_key_destroy(ctx_, relin_);
  relin_ = k;
  return true;
}

bool CryptoContextImpl::SerializeEvalAutomorphismKey(std::ostream& ser, const SerType& sertype) const {
  if (sertype != SerType::BINARY) return false;
  put_header(ser, "AUTOKEY");
  put_u32(ser, (uint32_t)rot_.size());
  for (const auto& kv : rot_) {
    put_u32(ser, kv.first);
    if (!put_key(ser, ctx_, kv.second, key_words(*this))) return false;
  }
  return (bool)ser;
}

bool CryptoContextImpl::DeserializeEvalAutomorphismKey(std::istream& ser, const SerType& sertype) {
  uint32_t n = 0;
  if (sertype != SerType::BINARY || !get_header(ser, "AUTOKEY") || !get_u32(ser, n)) return false;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t g = 0;
    if (!get_u32(ser, g) || (g & 1) == 0 || g >= 2 * N_) return false;
    fh_key* k = get_key(ser, *this, ctx_, key_words(*this));
    if (!k) return false;
    auto it = rot_.find(g);
    if (it != rot_.end()) fh_key_destroy(ctx_, it->second);
    rot_[g] = k;
  }
  return true;
}

namespace Serial {

bool SerializeToFile(const std::string& filename, const CryptoContext& obj, const SerType& sertype) {
  if (sertype != SerType::BINARY || !obj) return false;
  std::ofstream f(filename, std::ios::binary);
  if (!f) return false;
  const CCParams& p = obj->GetParams();
  put_header(f, "CONTEXT");
  for (uint32_t v : {p.depth_, p.scale_bits_, p.first_bits_, p.ring_dim_, p.dnum_, p.batch_, p.hw_, p.num_special_,
                     (uint32_t)p.tech_, (uint32_t)p.dist_})
    put_u32(f, v);
  put_u64(f, p.seed_);
  const auto& m = obj->GetModuli();
  put_u32(f, (uint32_t)m.size());
  put_u32(f, obj->GetNumQ());
  for (uint64_t q : m) put_u64(f, q);
  return (bool)f;
}

bool DeserializeFromFile(const std::string& filename, CryptoContext& obj, const SerType& sertype) {
  if (sertype != SerType::BINARY) return false;
  std::ifstream f(filename, std::ios::binary);
  if (!f || !get_header(f, "CONTEXT")) return false;
  CCParams p;
  uint32_t v[10];
  for (auto& x : v)
    if (!get_u32(f, x)) return false;
  if (v[8] > (uint32_t)ScalingTechnique::FLEXIBLEAUTO || v[9] > (uint32_t)SecretKeyDist::SPARSE_TERNARY) return false;
  p.depth_ = v[0], p.scale_bits_ = v[1], p.first_bits_ = v[2], p.ring_dim_ = v[3], p.dnum_ = v[4], p.batch_ = v[5], p.hw_ = v[6];
  p.num_special_ = v[7];
  p.tech_ = (ScalingTechnique)v[8];
  p.dist_ = (SecretKeyDist)v[9];
  uint32_t nm = 0, nq = 0;
  if (!get_u64(f, p.seed_) || !get_u32(f, nm) || !get_u32(f, nq) || nm > 256) return false;
  std::vector<uint64_t> m(nm);
  for (auto& q : m)
    if (!get_u64(f, q)) return false;
  CryptoContext cc = GenCryptoContext(p);
  if (cc->GetModuli() != m || cc->GetNumQ() != nq) return false;  // the same chain must come out
  obj = cc;
  return true;
}

bool SerializeToFile(const std::string& filename, const PublicKey& obj, const SerType& sertype) {
  if (sertype != SerType::BINARY || !obj || !obj->b || !obj->a) return false;
  std::ofstream f(filename, std::ios::binary);
  if (!f) return false;
  put_header(f, "PUBKEY");
  fh_context* ctx = obj->b->owner->ctx;
  return put_device(f, ctx, obj->b->ptr, obj->b->words) && put_device(f, ctx, obj->a->ptr, obj->a->words);
}

bool DeserializeFromFile(const std::string& filename, PublicKey& obj, const SerType& sertype, const CryptoContext& cc) {
  if (sertype != SerType::BINARY || !cc) return false;
  std::ifstream f(filename, std::ios::binary);
  if (!f || !get_header(f, "PUBKEY")) return false;
  const size_t words = (size_t)cc->GetNumQ() * cc->GetRingDimension();
  auto pk = std::make_shared<PublicKeyImpl>();
  pk->b = get_device(f, *cc, words, cc->GetNumQ());
  pk->a = pk->b ? get_device(f, *cc, words, cc->GetNumQ()) : nullptr;
  if (!pk->a) return false;
  obj = pk;
  return true;
}

bool SerializeToFile(const std::string& filename, const PrivateKey& obj, const SerType& sertype) {
  if (sertype != SerType::BINARY || !obj || !obj->s) return false;
  std::ofstream f(filename, std::ios::binary);
  if (!f) return false;
  put_header(f, "SECKEY");
  return put_device(f, obj->s->owner->ctx, obj->s->ptr, obj->s->words);
}

bool DeserializeFromFile(const std::string& filename, PrivateKey& obj, const SerType& sertype, const CryptoContext& cc) {
  if (sertype != SerType::BINARY || !cc) return false;
  std::ifstream f(filename, std::ios::binary);
  if (!f || !get_header(f, "SECKEY")) return false;
  auto sk = std::make_shared<PrivateKeyImpl>();
  sk->s = get_device(f, *cc, cc->GetModuli().size() * (size_t)cc->GetRingDimension(), cc->GetModuli().size());
  if (!sk->s) return false;
  obj = sk;
  return true;
}

bool SerializeToFile(const std::string& filename, const Ciphertext& obj, const SerType& sertype) {
  if (sertype != SerType::BINARY || !obj || !obj->data) return false;
  std::ofstream f(filename, std::ios::binary);
  if (!f) return false;
  put_header(f, "CIPHER");
  put_u32(f, obj->towers);
  put_u32(f, obj->ncomp);
  put_u32(f, obj->noiseScaleDeg);
  put_u32(f, obj->slots);
  put_f64(f, obj->scalingFactor);
  return put_device(f, obj->data->owner->ctx, obj->data->ptr, obj->data->words);
}

bool DeserializeFromFile(const std::string& filename, Ciphertext& obj, const SerType& sertype, const CryptoContext& cc) {
  if (sertype != SerType::BINARY || !cc) return false;
  std::ifstream f(filename, std::ios::binary);
  if (!f || !get_header(f, "CIPHER")) return false;
  auto ct = std::make_shared<CiphertextImpl>();
  if (!get_u32(f, ct->towers) || !get_u32(f, ct->ncomp) || !get_u32(f, ct->noiseScaleDeg) || !get_u32(f, ct->slots) ||
      !get_f64(f, ct->scalingFactor))
    return false;
  if (ct->towers < 1 || ct->towers > cc->GetNumQ() || ct->ncomp < 2 || ct->ncomp > 3 || ct->noiseScaleDeg < 1 || ct->noiseScaleDeg > 2)
    return false;
  // slots: a power of two in [1, N/2]; the scale a finite positive number
  if (ct->slots < 1 || ct->slots > cc->GetRingDimension() / 2 || (ct->slots & (ct->slots - 1)) != 0) return false;
  if (!(ct->scalingFactor > 0) || !std::isfinite(ct->scalingFactor)) return false;
  ct->data = get_device(f, *cc, (size_t)ct->ncomp * ct->towers * cc->GetRingDimension(), ct->towers);
  if (!ct->data) return false;
  obj = ct;
  return true;
}

}  // namespace Serial
}  // namespace fideslib
