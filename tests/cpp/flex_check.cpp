// A FLEXIBLEAUTO session through the fideslib facade on the GPU: per-level scaling factors,
// products left at depth 2 and rescaled by the next operation, and every branch of the level /
// depth alignment (OpenFHE's AdjustLevelsAndDepthInPlace as the reference restates it,
// src/CKKS/Ciphertext.cpp:1917-2070) reached through EvalAdd, EvalMult, plaintext operands, scalar
// operands and EvalLinearWSum. Writes every result (residues, metadata, decryption error against
// the expected slots) for tests/test_gpu_flex.py, which replays the session in the oracle.
//
// usage: flex_check <out_dir> [seed]
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../fideslib-fork_amd/host/fideslib.hpp"

using namespace fideslib;
using cvec = std::vector<std::complex<double>>;

static std::string g_dir;
static FILE* g_manifest = nullptr;
static bool g_first = true;

static void write_raw(const std::string& name, const void* p, size_t bytes) {
  FILE* f = std::fopen((g_dir + "/" + name).c_str(), "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
    std::fprintf(stderr, "cannot write %s\n", name.c_str());
    std::exit(1);
  }
  std::fclose(f);
}

static void sep() {
  std::fprintf(g_manifest, g_first ? "\n" : ",\n");
  g_first = false;
}

static void dump_values(const std::string& name, const cvec& v) {
  std::vector<double> flat(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    flat[2 * i] = v[i].real();
    flat[2 * i + 1] = v[i].imag();
  }
  write_raw(name + ".f64", flat.data(), flat.size() * 8);
}

static void dump_ct(CryptoContextImpl& cc, const PrivateKey& sk, const std::string& name, const Ciphertext& ct, const cvec& expect) {
  const size_t words = (size_t)ct->ncomp * ct->towers * cc.GetRingDimension();
  std::vector<uint64_t> h(words);
  if (fh_copy_d2h(cc.Native(), h.data(), ct->data->ptr, words * 8) != FH_OK) {
    std::fprintf(stderr, "copy failed: %s\n", fh_last_error());
    std::exit(1);
  }
  write_raw(name + ".u64", h.data(), words * 8);
  auto pt = cc.Decrypt(sk, ct);
  double err = 0;
  for (size_t i = 0; i < expect.size(); ++i) err = std::max(err, std::abs(pt->values[i] - expect[i]));
  sep();
  std::fprintf(g_manifest, "  \"%s\": {\"towers\": %u, \"ncomp\": %u, \"deg\": %u, \"scale\": %.17g, \"err\": %.6g}", name.c_str(),
               ct->towers, ct->ncomp, ct->noiseScaleDeg, ct->scalingFactor, err);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: flex_check <out_dir> [seed]\n");
    return 2;
  }
  g_dir = argv[1];
  const uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 77;
  g_manifest = std::fopen((g_dir + "/manifest.json").c_str(), "w");
  if (!g_manifest) return 1;
  std::fprintf(g_manifest, "{");
  try {
    CCParams p;
    p.SetMultiplicativeDepth(5);
    p.SetScalingModSize(50);
    p.SetFirstModSize(60);
    p.SetRingDim(1 << 13);
    p.SetNumLargeDigits(2);
    p.SetScalingTechnique(ScalingTechnique::FLEXIBLEAUTO);
    p.SetSeed(seed);
    auto cc = GenCryptoContext(p);
    const uint32_t N = cc->GetRingDimension(), slots = N / 2, L1 = cc->GetNumQ();
    auto kp = cc->KeyGen();                        // seeds: secret +0, public key +1
    cc->EvalMultKeyGen(kp.secretKey);              // +2
    cc->EvalRotateKeyGen(kp.secretKey, {1});       // +3
    cvec x(slots), y(slots), e(slots);
    for (uint32_t i = 0; i < slots; ++i) {
      x[i] = {0.6 * std::sin(0.011 * i), 0.3 * std::cos(0.02 * i)};
      y[i] = {0.7 * std::cos(0.005 * i), -0.2 * std::sin(0.013 * i)};
    }
    dump_values("x", x);
    dump_values("y", y);
    auto ptx = cc->MakeCKKSPackedPlaintext(x), pty = cc->MakeCKKSPackedPlaintext(y);
    auto ct1 = cc->Encrypt(kp.secretKey, ptx);     // +4
    auto ct2 = cc->Encrypt(kp.secretKey, pty);     // +5

    sep();
    std::fprintf(g_manifest, "  \"params\": {\"logN\": 13, \"L1\": %u, \"K\": %u, \"dnum\": 2, \"seed\": %llu}", L1,
                 (unsigned)(cc->GetModuli().size() - L1), (unsigned long long)seed);
    sep();
    std::fprintf(g_manifest, "  \"moduli\": [");
    for (size_t i = 0; i < cc->GetModuli().size(); ++i) std::fprintf(g_manifest, "%s\"%llu\"", i ? ", " : "", (unsigned long long)cc->GetModuli()[i]);
    std::fprintf(g_manifest, "]");
    sep();
    std::fprintf(g_manifest, "  \"tables\": {\"real\": [");
    for (uint32_t l = 0; l < L1; ++l) std::fprintf(g_manifest, "%s%.17g", l ? ", " : "", cc->GetScalingFactorReal(l));
    std::fprintf(g_manifest, "], \"big\": [");
    for (uint32_t l = 0; l < L1; ++l) std::fprintf(g_manifest, "%s%.17g", l ? ", " : "", cc->GetScalingFactorRealBig(l));
    std::fprintf(g_manifest, "], \"mrf\": [");
    for (uint32_t l = 0; l < L1; ++l) std::fprintf(g_manifest, "%s%.17g", l ? ", " : "", cc->GetModReduceFactor(l));
    std::fprintf(g_manifest, "]}");
    sep();
    std::fprintf(g_manifest, "  \"pt_scale\": %.17g", ptx->scalingFactor);

    auto xy = [&](auto f) {
      for (uint32_t i = 0; i < slots; ++i) e[i] = f(x[i], y[i]);
      return e;
    };
    using C = std::complex<double>;
    dump_ct(*cc, kp.secretKey, "ct1", ct1, x);
    dump_ct(*cc, kp.secretKey, "ct2", ct2, y);
    // a product stays at depth 2 (Rescale does nothing under FLEXIBLEAUTO)
    auto m = cc->Rescale(cc->EvalMult(ct1, ct2));
    dump_ct(*cc, kp.secretKey, "m", m, xy([](C a, C b) { return a * b; }));
    // equal levels, depths 2 and 1: the depth-1 operand times 1.0
    dump_ct(*cc, kp.secretKey, "m_add", cc->EvalAdd(m, ct1), xy([](C a, C b) { return a * b + a; }));
    // depth-2 inputs rescaled before the product
    auto m2 = cc->EvalMult(m, ct1);
    dump_ct(*cc, kp.secretKey, "m2", m2, xy([](C a, C b) { return a * b * a; }));
    // a depth-1 operand two levels up brought down to a depth-2 one
    auto m3 = cc->EvalMult(m2, ct2);
    dump_ct(*cc, kp.secretKey, "m3", m3, xy([](C a, C b) { return a * b * a * b; }));
    auto lr1 = cc->LevelReduce(ct2, 1), lr2 = cc->LevelReduce(ct2, 2);
    // depth 2 over depth 1 one level down: one rescale
    dump_ct(*cc, kp.secretKey, "a_d2d1_adj", cc->EvalAdd(m, lr1), xy([](C a, C b) { return a * b + b; }));
    // depth 2 over depth 1 two levels down: scalar product, rescale, drop, rescale
    dump_ct(*cc, kp.secretKey, "a_d2d1", cc->EvalAdd(m, lr2), xy([](C a, C b) { return a * b + b; }));
    // depth 2 over depth 2 two levels down
    dump_ct(*cc, kp.secretKey, "a_d2d2", cc->EvalAdd(m, m3), xy([](C a, C b) { return a * b + a * b * a * b; }));
    // depth 1 over depth 1 two levels down
    dump_ct(*cc, kp.secretKey, "a_d1d1", cc->EvalAdd(ct1, lr2), xy([](C a, C b) { return a + b; }));
    // depth 1 over depth 2 two levels down
    dump_ct(*cc, kp.secretKey, "a_d1d2", cc->EvalSub(ct1, m3), xy([](C a, C b) { return a - a * b * a * b; }));
    // scalar operands at depth 2
    dump_ct(*cc, kp.secretKey, "add_s", cc->EvalAdd(m2, 0.5), xy([](C a, C b) { return a * b * a + 0.5; }));
    dump_ct(*cc, kp.secretKey, "mul_s", cc->EvalMult(m2, -0.75), xy([](C a, C b) { return a * b * a * -0.75; }));
    // plaintext operands (encoded at the top level's factor)
    dump_ct(*cc, kp.secretKey, "mul_p", cc->EvalMult(m, ptx), xy([](C a, C b) { return a * b * a; }));
    dump_ct(*cc, kp.secretKey, "add_p", cc->EvalAdd(m2, pty), xy([](C a, C b) { return a * b * a + b; }));
    dump_ct(*cc, kp.secretKey, "sub_p", cc->EvalSub(ct1, pty), xy([](C a, C b) { return a - b; }));
    // rotation of a depth-2 ciphertext (no metadata change)
    {
      auto r = cc->EvalRotate(m2, 1);
      const cvec base = xy([](C a, C b) { return a * b * a; });
      cvec rot(slots);
      for (uint32_t i = 0; i < slots; ++i) rot[i] = base[(i + 1) % slots];
      dump_ct(*cc, kp.secretKey, "rot", r, rot);
    }
    // weighted sum of inputs at three levels and two depths
    dump_ct(*cc, kp.secretKey, "wsum", cc->EvalLinearWSum({ct1, m2, ct2}, {0.5, -1.25, 2.0}),
            xy([](C a, C b) { return 0.5 * a - 1.25 * a * b * a + 2.0 * b; }));
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    std::fprintf(g_manifest, "}\n");
    std::fclose(g_manifest);
    return 1;
  }
  std::fprintf(g_manifest, "\n}\n");
  std::fclose(g_manifest);
  std::printf("flex_check ok\n");
  return 0;
}
