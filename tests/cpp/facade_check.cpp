// Drives the fideslib C++ facade through one CKKS session on the GPU and writes every result
// (raw residues, metadata, decrypted slots) to a directory, for tests/test_gpu_facade.py to
// recompute with the oracle. The scenario follows the reference's facade tests
// (test/ApiTests.cpp:95-330: keygen with mult and rotation keys, encode, public- and secret-key
// encryption, EvalAdd/EvalSub/EvalMult over ciphertext, plaintext and scalar operands, rotations,
// decrypt).
//
// usage: facade_check <out_dir> [seed]
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

static void entry_sep() {
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

static double max_err(const cvec& a, const cvec& b) {
  double e = 0;
  for (size_t i = 0; i < a.size(); ++i) e = std::max(e, std::abs(a[i] - b[i]));
  return e;
}

// writes the ciphertext's residues and metadata; decrypts and reports the error against `expect`
static void dump_ct(CryptoContextImpl& cc, const PrivateKey& sk, const std::string& name, const Ciphertext& ct, const cvec& expect) {
  const size_t words = (size_t)ct->ncomp * ct->towers * cc.GetRingDimension();
  std::vector<uint64_t> h(words);
  if (fh_copy_d2h(cc.Native(), h.data(), ct->data->ptr, words * 8) != FH_OK) {
    std::fprintf(stderr, "copy failed: %s\n", fh_last_error());
    std::exit(1);
  }
  write_raw(name + ".u64", h.data(), words * 8);
  auto pt = cc.Decrypt(sk, ct);
  dump_values(name + "_dec", pt->values);
  entry_sep();
  std::fprintf(g_manifest, "  \"%s\": {\"towers\": %u, \"ncomp\": %u, \"noise_deg\": %u, \"scale\": %.17g, \"slots\": %u, \"err\": %.6g}",
               name.c_str(), ct->towers, ct->ncomp, ct->noiseScaleDeg, ct->scalingFactor, ct->slots, max_err(pt->values, expect));
  std::printf("%-12s towers=%u ncomp=%u deg=%u err=%.3g\n", name.c_str(), ct->towers, ct->ncomp, ct->noiseScaleDeg, max_err(pt->values, expect));
}

static void dump_u64_list(const std::string& key, const std::vector<uint64_t>& v) {
  entry_sep();
  std::fprintf(g_manifest, "  \"%s\": [", key.c_str());
  for (size_t i = 0; i < v.size(); ++i) std::fprintf(g_manifest, "%s\"%llu\"", i ? ", " : "", (unsigned long long)v[i]);
  std::fprintf(g_manifest, "]");
}

static cvec rotate(const cvec& v, int r) {
  const int n = (int)v.size();
  cvec o(n);
  for (int i = 0; i < n; ++i) o[i] = v[((i + r) % n + n) % n];
  return o;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: facade_check <out_dir> [seed]\n");
    return 2;
  }
  g_dir = argv[1];
  const uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1234;
  g_manifest = std::fopen((g_dir + "/manifest.json").c_str(), "w");
  if (!g_manifest) return 1;
  std::fprintf(g_manifest, "{");

  try {
    CCParams p;
    p.SetMultiplicativeDepth(3);
    p.SetScalingModSize(50);
    p.SetFirstModSize(60);
    p.SetRingDim(1 << 13);
    p.SetNumLargeDigits(2);
    p.SetScalingTechnique(ScalingTechnique::FIXEDMANUAL);
    p.SetSeed(seed);
    auto cc = GenCryptoContext(p);
    const uint32_t N = cc->GetRingDimension(), slots = N / 2;

    auto kp = cc->KeyGen();                       // seeds: secret +0, public key +1
    cc->EvalMultKeyGen(kp.secretKey);             // +2
    const std::vector<int32_t> steps = {1, 2, -3, 7};
    cc->EvalRotateKeyGen(kp.secretKey, steps);    // +3 .. +6
    cc->EvalConjugateKeyGen(kp.secretKey);        // +7

    cvec x(slots), y(slots);
    for (uint32_t i = 0; i < slots; ++i) {
      x[i] = std::complex<double>(0.5 * std::sin(0.01 * i), 0.25 * std::cos(0.03 * i));
      y[i] = std::complex<double>(0.8 * std::cos(0.007 * i), 0.0);
    }
    dump_values("x", x);
    dump_values("y", y);
    auto ptx = cc->MakeCKKSPackedPlaintext(x);
    auto pty = cc->MakeCKKSPackedPlaintext(y);
    auto ct1 = cc->Encrypt(kp.secretKey, ptx);    // +8
    auto ct2 = cc->Encrypt(kp.publicKey, pty);    // +9

    // secret and public key, for the oracle to start from the same keys
    {
      const uint32_t np = (uint32_t)cc->GetModuli().size();
      std::vector<uint64_t> h((size_t)np * N);
      fh_copy_d2h(cc->Native(), h.data(), kp.secretKey->s->ptr, h.size() * 8);
      write_raw("secret.u64", h.data(), h.size() * 8);
    }
    dump_u64_list("moduli", cc->GetModuli());
    std::vector<uint64_t> gal;
    for (int32_t r : steps) gal.push_back(cc->FindAutomorphismIndex(r < 0 ? r + (int32_t)N / 2 : r));
    dump_u64_list("galois", gal);
    entry_sep();
    std::fprintf(g_manifest, "  \"params\": {\"logN\": 13, \"L1\": %u, \"K\": %u, \"dnum\": 2, \"seed\": %llu, \"scale\": %.17g}",
                 cc->GetNumQ(), (unsigned)(cc->GetModuli().size() - cc->GetNumQ()), (unsigned long long)seed, cc->ScalingFactor());

    cvec e(slots);
    dump_ct(*cc, kp.secretKey, "ct1", ct1, x);
    dump_ct(*cc, kp.secretKey, "ct2", ct2, y);

    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] + y[i];
    dump_ct(*cc, kp.secretKey, "add", cc->EvalAdd(ct1, ct2), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] - y[i];
    dump_ct(*cc, kp.secretKey, "sub", cc->EvalSub(ct1, ct2), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] + 0.25;
    dump_ct(*cc, kp.secretKey, "add_s", cc->EvalAdd(ct1, 0.25), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] - 1.5;
    dump_ct(*cc, kp.secretKey, "sub_s", cc->EvalSub(ct1, 1.5), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] + y[i];
    dump_ct(*cc, kp.secretKey, "add_p", cc->EvalAdd(ct1, pty), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = -x[i];
    dump_ct(*cc, kp.secretKey, "neg", cc->EvalNegate(ct1), e);

    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] * y[i];
    auto m12 = cc->EvalMult(ct1, ct2);
    dump_ct(*cc, kp.secretKey, "mult", m12, e);
    dump_ct(*cc, kp.secretKey, "mult_rs", cc->Rescale(m12), e);
    auto nr = cc->EvalMultNoRelin(ct1, ct2);
    dump_ct(*cc, kp.secretKey, "mult_nr", nr, e);
    dump_ct(*cc, kp.secretKey, "relin", cc->Relinearize(nr), e);
    dump_ct(*cc, kp.secretKey, "mult_p", cc->Rescale(cc->EvalMult(ct1, pty)), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = y[i] * y[i];
    dump_ct(*cc, kp.secretKey, "square", cc->Rescale(cc->EvalSquare(ct2)), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] * 2.5;
    dump_ct(*cc, kp.secretKey, "mult_s", cc->Rescale(cc->EvalMult(ct1, 2.5)), e);
    for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] * -0.75;
    dump_ct(*cc, kp.secretKey, "mult_sn", cc->Rescale(cc->EvalMult(ct1, -0.75)), e);
    dump_ct(*cc, kp.secretKey, "lvl_red", cc->LevelReduce(ct1, 1), x);

    dump_ct(*cc, kp.secretKey, "rot1", cc->EvalRotate(ct1, 1), rotate(x, 1));
    dump_ct(*cc, kp.secretKey, "rotm3", cc->EvalRotate(ct2, -3), rotate(y, -3));
    {
      // rotation-fold sum over 3 summands (rotations by 1 and 2 at radix 4)
      const cvec r1 = rotate(x, 1), r2 = rotate(x, 2);
      for (uint32_t i = 0; i < slots; ++i) e[i] = x[i] + r1[i] + r2[i];
      dump_ct(*cc, kp.secretKey, "acc3", cc->AccumulateSum(ct1, 3), e);
      // two fold levels (16 summands: rotations 1, 2, 3, then 4, 8, 12): the lazy fold, and the
      // cascade from start factor 4 (rotations 4, 8, 12 of the input: 4 summands at stride 4)
      cc->EvalRotateKeyGen(kp.secretKey, std::vector<int32_t>{3, 4, 8, 12});  // +10 .. +13
      for (uint32_t i = 0; i < slots; ++i) {
        e[i] = 0;
        for (uint32_t k = 0; k < 16; ++k) e[i] += x[(i + k) % slots];
      }
      dump_ct(*cc, kp.secretKey, "acc16", cc->AccumulateSum(ct1, 16), e);
      for (uint32_t i = 0; i < slots; ++i) {
        e[i] = 0;
        for (uint32_t k = 0; k < 16; k += 4) e[i] += x[(i + k) % slots];
      }
      Ciphertext t4 = ct1;
      cc->AccumulateSumInPlace(t4, 16, 1, 4);
      t4->slots = ct1->slots;
      dump_ct(*cc, kp.secretKey, "acc16s4", t4, e);
    }
    for (uint32_t i = 0; i < slots; ++i) e[i] = std::conj(x[i]);
    dump_ct(*cc, kp.secretKey, "conj", cc->EvalConjugate(ct1), e);
    auto pre = cc->EvalFastRotationPrecompute(ct1);
    dump_ct(*cc, kp.secretKey, "fast7", cc->EvalFastRotation(ct1, 7, 2 * N, pre), rotate(x, 7));
    // the extended form from the same digits, finished by KeySwitchDown
    auto exts = cc->EvalFastRotationExt(ct1, std::vector<int32_t>{7, -3}, pre, true);
    dump_ct(*cc, kp.secretKey, "fastext7", cc->KeySwitchDown(exts[0]), rotate(x, 7));
    dump_ct(*cc, kp.secretKey, "fastextm3", cc->KeySwitchDown(exts[1]), rotate(x, -3));
    auto many = cc->EvalRotateMany({ct1, ct2}, 2);
    dump_ct(*cc, kp.secretKey, "many0", many[0], rotate(x, 2));
    dump_ct(*cc, kp.secretKey, "many1", many[1], rotate(y, 2));

    // BSGS linear transform: rowSize 4, bStep 2, stride 1; D_k stored pre-rotated
    const int rowSize = 4, bStep = 2;
    std::vector<Plaintext> diags;
    std::vector<cvec> D(rowSize, cvec(slots));
    for (int k = 0; k < rowSize; ++k) {
      for (uint32_t i = 0; i < slots; ++i) D[k][i] = std::complex<double>(0.3 * std::cos(0.002 * (i + 17 * k)), 0.1 * (k + 1) * std::sin(0.004 * i));
      dump_values("diag" + std::to_string(k), D[k]);
      diags.push_back(cc->MakeCKKSPackedPlaintextExtended(rotate(D[k], -(bStep * (k / bStep))), 1, 0));
    }
    for (uint32_t i = 0; i < slots; ++i) {
      e[i] = 0;
      for (int k = 0; k < rowSize; ++k) e[i] += x[(i + k) % slots] * D[k][i];
    }
    dump_ct(*cc, kp.secretKey, "lt", cc->LinearTransform(ct1, rowSize, bStep, diags, 1, 0), e);
    // the same transform prepared once (resident packed diagonals), applied in place
    auto prepared = cc->PrepareLinearTransform(diags, rowSize, bStep, 1, 0);
    Ciphertext ltp = ct1;
    cc->LinearTransformInPlace(ltp, prepared);
    dump_ct(*cc, kp.secretKey, "lt_prepared", ltp, e);

    // API forms: in-place operations, the batched device encoder, rotation-index helpers,
    // LinearTransformMany, Compress (each compared with the plain form on the device words)
    auto dev = [&](const uint64_t* p, size_t words) {
      std::vector<uint64_t> h(words);
      if (fh_copy_d2h(cc->Native(), h.data(), p, words * 8) != FH_OK) throw std::runtime_error(fh_last_error());
      return h;
    };
    auto ctw = [&](const Ciphertext& c) { return dev(c->data->ptr, (size_t)c->ncomp * c->towers * N); };
    int api_ok = 0, api_total = 0;
    auto expect = [&](bool v) { ++api_total; api_ok += v ? 1 : 0; };
    {
      Ciphertext t = ct1;
      cc->EvalAddInPlace(t, ct2);
      expect(ctw(t) == ctw(cc->EvalAdd(ct1, ct2)));
      t = ct1;
      cc->EvalMultInPlace(t, ct2);
      cc->RescaleInPlace(t);
      expect(ctw(t) == ctw(cc->Rescale(cc->EvalMult(ct1, ct2))));
      t = ct1;
      cc->EvalRotateInPlace(t, 1);
      expect(ctw(t) == ctw(cc->EvalRotate(ct1, 1)));
      std::vector<double> r0(slots), r1(slots);
      for (uint32_t i = 0; i < slots; ++i) r0[i] = x[i].real(), r1[i] = 0.5 * x[i].imag();
      auto many = cc->MakeCKKSPackedPlaintextDeviceMany({r0, r1}, 1, 1);
      auto one0 = cc->MakeCKKSPackedPlaintext(r0, 1, 1), one1 = cc->MakeCKKSPackedPlaintext(r1, 1, 1);
      const size_t pw = (size_t)one0->towers * N;
      expect(many.size() == 2 && many[0]->towers == one0->towers && dev(many[0]->data->ptr, pw) == dev(one0->data->ptr, pw) &&
             dev(many[1]->data->ptr, pw) == dev(one1->data->ptr, pw));
      expect(CryptoContextImpl::GetLinearTransformRotationIndices(2, 1, 0) == std::vector<int>{1, 2});
      expect(CryptoContextImpl::GetLinearTransformPlaintextRotationIndices(4, 2, 1, 0) == std::vector<int>{0, 0, -2, -2});
      auto pm = cc->PrepareLinearTransformMany({diags, diags}, rowSize, bStep, 1, 0);
      auto lts = cc->LinearTransformMany(ct1, pm);
      const auto ref = ctw(ltp);
      expect(lts.size() == 2 && ctw(lts[0]) == ref && ctw(lts[1]) == ref);
      auto cmp = cc->Compress(cc->EvalMult(ct1, ct2), 1);
      expect(cmp->towers == 1 && cmp->noiseScaleDeg == 1);
    }
    entry_sep();
    std::fprintf(g_manifest, "  \"api_forms\": [%d, %d]", api_ok, api_total);

    // scalar CRT encodings
    const double ops[] = {2.5, -0.75, 1e-9, 3.0e30, -7.25e25, 0.0};
    for (int i = 0; i < 6; ++i) dump_u64_list("em" + std::to_string(i), cc->ElemForEvalMult(cc->GetNumQ(), ops[i]));
    const double aops[] = {0.25, 1.5, 1e12, 3.3e20};
    for (int i = 0; i < 4; ++i)
      for (uint32_t d = 1; d <= 2; ++d) dump_u64_list("ea" + std::to_string(i) + "_" + std::to_string(d), cc->ElemForEvalAddOrSub(cc->GetNumQ(), aops[i], d));

    // error paths the reference's facade reports
    int errors = 0;
    try {
      cc->EvalMult(nr, ct1);
    } catch (const std::runtime_error&) {
      ++errors;
    }
    try {
      cc->Rescale(exts[0]);  // an extended ciphertext only goes to KeySwitchDown
    } catch (const std::runtime_error&) {
      ++errors;
    }
    try {
      cc->EvalRotate(ct1, 5);  // no key for 5
    } catch (const std::runtime_error&) {
      ++errors;
    }
    try {
      cc->EvalAdd(m12, ct1);  // degree 2 + degree 1
    } catch (const std::runtime_error&) {
      ++errors;
    }
    entry_sep();
    std::fprintf(g_manifest, "  \"errors_thrown\": %d", errors);
    cc->Synchronize();
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "facade_check failed: %s\n", ex.what());
    std::fprintf(g_manifest, "\n}\n");
    std::fclose(g_manifest);
    return 1;
  }
  std::fprintf(g_manifest, "\n}\n");
  std::fclose(g_manifest);
  std::printf("facade_check done\n");
  return 0;
}
