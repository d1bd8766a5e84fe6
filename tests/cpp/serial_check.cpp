// Serialization round trip through the fideslib facade: a context, its public and secret keys,
// relinearization and rotation keys and a ciphertext are written to <dir>, read back into a second
// context rebuilt from the file, and used there. Prints one JSON line; tests/test_gpu_serial.py
// checks that every device word and every decryption agrees with the original context's.
//
// usage: serial_check <dir>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../fideslib-fork_amd/host/fideslib.hpp"

using namespace fideslib;

static std::vector<uint64_t> words(CryptoContextImpl& cc, const Ciphertext& ct) {
  std::vector<uint64_t> h((size_t)ct->ncomp * ct->towers * cc.GetRingDimension());
  if (fh_copy_d2h(cc.Native(), h.data(), ct->data->ptr, h.size() * 8) != FH_OK) throw std::runtime_error(fh_last_error());
  return h;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: serial_check <dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  try {
    CCParams p;
    p.SetRingDim(1 << 12);
    p.SetMultiplicativeDepth(5);
    p.SetScalingModSize(50);
    p.SetFirstModSize(60);
    p.SetNumLargeDigits(3);
    p.SetSeed(17);
    auto a = GenCryptoContext(p);
    auto kp = a->KeyGen();
    a->EvalMultKeyGen(kp.secretKey);
    a->EvalRotateKeyGen(kp.secretKey, {1, -3});
    const uint32_t n = a->GetRingDimension() / 2;
    std::mt19937_64 rng(5);
    std::uniform_real_distribution<double> U(-1, 1);
    std::vector<std::complex<double>> z(n);
    for (auto& v : z) v = {U(rng), U(rng)};
    auto ct = a->Encrypt(kp.publicKey, a->MakeCKKSPackedPlaintext(z));

    bool ok = Serial::SerializeToFile(dir + "/cc.bin", a, SerType::BINARY) && Serial::SerializeToFile(dir + "/pk.bin", kp.publicKey, SerType::BINARY) &&
              Serial::SerializeToFile(dir + "/sk.bin", kp.secretKey, SerType::BINARY) && Serial::SerializeToFile(dir + "/ct.bin", ct, SerType::BINARY);
    {
      std::ofstream f(dir + "/mult.bin", std::ios::binary);
      ok = ok && a->SerializeEvalMultKey(f, SerType::BINARY);
      std::ofstream g(dir + "/rot.bin", std::ios::binary);
      ok = ok && a->SerializeEvalAutomorphismKey(g, SerType::BINARY);
    }
    const bool json_refused = !Serial::SerializeToFile(dir + "/cc.json", a, SerType::JSON);

    CryptoContext b;
    PublicKey pk;
    PrivateKey sk;
    Ciphertext ct2;
    ok = ok && Serial::DeserializeFromFile(dir + "/cc.bin", b, SerType::BINARY) &&
         Serial::DeserializeFromFile(dir + "/pk.bin", pk, SerType::BINARY, b) && Serial::DeserializeFromFile(dir + "/sk.bin", sk, SerType::BINARY, b) &&
         Serial::DeserializeFromFile(dir + "/ct.bin", ct2, SerType::BINARY, b);
    {
      std::ifstream f(dir + "/mult.bin", std::ios::binary);
      ok = ok && b->DeserializeEvalMultKey(f, SerType::BINARY);
      std::ifstream g(dir + "/rot.bin", std::ios::binary);
      ok = ok && b->DeserializeEvalAutomorphismKey(g, SerType::BINARY);
    }
    // a file of another kind is refused
    const bool wrong_kind_refused = !Serial::DeserializeFromFile(dir + "/pk.bin", ct2, SerType::BINARY, b);
    if (!ok) {
      std::printf("{\"ok\": false}\n");
      return 1;
    }
    // corrupt copies are refused: a residue equal to its prime, a slot count that is not a power of
    // two, a key word above its prime
    auto patched = [&](const std::string& src, const std::string& dst, size_t off, const void* v, size_t n) {
      std::ifstream in(src, std::ios::binary);
      std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      std::memcpy(&bytes[off], v, n);
      std::ofstream out(dst, std::ios::binary);
      out.write(bytes.data(), (std::streamsize)bytes.size());
    };
    const size_t N = b->GetRingDimension();
    const uint64_t q0 = b->GetModuli()[0], q1 = b->GetModuli()[1];
    const uint32_t three = 3;
    patched(dir + "/ct.bin", dir + "/ct_noncanon.bin", 16 + 24 + 8 + 8 * 5, &q0, 8);   // header, metadata, word count
    patched(dir + "/ct.bin", dir + "/ct_slots.bin", 16 + 12, &three, 4);
    patched(dir + "/mult.bin", dir + "/mult_noncanon.bin", 16 + 8 + 8 * (N + 7), &q1, 8);
    Ciphertext bad;
    bool corrupt_refused = !Serial::DeserializeFromFile(dir + "/ct_noncanon.bin", bad, SerType::BINARY, b) &&
                           !Serial::DeserializeFromFile(dir + "/ct_slots.bin", bad, SerType::BINARY, b);
    {
      std::ifstream f(dir + "/mult_noncanon.bin", std::ios::binary);
      corrupt_refused = corrupt_refused && !b->DeserializeEvalMultKey(f, SerType::BINARY);
    }
    // the untouched copy still loads
    corrupt_refused = corrupt_refused && Serial::DeserializeFromFile(dir + "/ct.bin", bad, SerType::BINARY, b);
    Ciphertext ct3;
    Serial::DeserializeFromFile(dir + "/ct.bin", ct3, SerType::BINARY, b);

    // the same computation in both contexts
    auto run = [&](CryptoContextImpl& cc, const Ciphertext& c) { return cc.EvalRotate(cc.Rescale(cc.EvalMult(c, c)), -3); };
    auto ra = run(*a, ct), rb = run(*b, ct3);
    const bool same_ct = words(*a, ct) == words(*b, ct3);
    const bool same_result = words(*a, ra) == words(*b, rb);
    auto da = a->Decrypt(kp.secretKey, ra), db = b->Decrypt(sk, rb);
    bool same_dec = da->values.size() == db->values.size();
    for (size_t i = 0; same_dec && i < da->values.size(); ++i) same_dec = da->values[i] == db->values[i];
    // a fresh encryption under the deserialized public key decrypts under the deserialized secret key
    auto fresh = b->Encrypt(pk, b->MakeCKKSPackedPlaintext(z));
    auto df = b->Decrypt(sk, fresh);
    double err = 0;
    for (uint32_t i = 0; i < n; ++i) err = std::max(err, std::abs(df->values[i] - z[i]));
    std::printf("{\"ok\": true, \"same_moduli\": %s, \"same_ciphertext\": %s, \"same_result\": %s, \"same_decryption\": %s, "
                "\"fresh_encryption_err\": %.3g, \"json_refused\": %s, \"wrong_kind_refused\": %s, \"corrupt_refused\": %s}\n",
                a->GetModuli() == b->GetModuli() ? "true" : "false", same_ct ? "true" : "false", same_result ? "true" : "false",
                same_dec ? "true" : "false", err, json_refused ? "true" : "false", wrong_kind_refused ? "true" : "false",
                corrupt_refused ? "true" : "false");
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
