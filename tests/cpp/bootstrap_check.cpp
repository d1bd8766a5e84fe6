// Bootstraps fresh ciphertexts through the fideslib C++ facade on the GPU and reports the
// decryption error (tests/test_gpu_bootstrap.py checks it against the stated bound). Also
// evaluates a Chebyshev series and reports its error against the same polynomial in double
// precision. The scenario follows the reference's bootstrap tests (test/zBtsTiming.cu,
// test/ApiTests.cpp: encrypt at the last level, EvalBootstrap, decrypt, compare).
//
// The inputs, outputs and bootstrap constants are written to <out_dir> so that the test can replay
// the same sequence with the oracle and compare residues.
//
// usage: bootstrap_check <out_dir> <logN> <slots> [seed] [cts_budget stc_budget depth scale_bits hw technique meta_precision
//                        first_mod_bits dnum real_input input_level]
// technique: 0 FIXEDMANUAL (default), 1 FLEXIBLEAUTO; meta_precision > 0 also runs the two-iteration
// bootstrap (Meta-BTS) at that precision and reports its error; first_mod_bits (default 60) and
// dnum (default 3) shape the modulus chain; real_input 1 draws real slot values (the reference's
// precision test); input_level > 0 encrypts at that OpenFHE level (levels consumed) instead of the
// lowest one the bootstrap accepts. (prints one JSON line)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../fideslib-fork_amd/host/fideslib.hpp"

using namespace fideslib;
using cvec = std::vector<std::complex<double>>;

static double max_err(const cvec& a, const cvec& b, size_t n) {
  double e = 0;
  for (size_t i = 0; i < n; ++i) e = std::max(e, std::abs(a[i] - b[i]));
  return e;
}

// sum c_k T_k(y) with the constant term halved, by the three-term recurrence
static double cheb_value(const std::vector<double>& c, double y) {
  double t0 = 1.0, t1 = y, s = 0.5 * c[0] + c[1] * y;
  for (size_t k = 2; k < c.size(); ++k) {
    const double t2 = 2.0 * y * t1 - t0;
    s += c[k] * t2;
    t0 = t1;
    t1 = t2;
  }
  return s;
}

static double sigmoid_like(double x, const void*) { return std::tanh(1.5 * x) + 0.25 * x * x; }

static std::string g_dir;
static void write_raw(const std::string& name, const void* p, size_t bytes) {
  FILE* f = std::fopen((g_dir + "/" + name).c_str(), "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
    std::fprintf(stderr, "cannot write %s\n", name.c_str());
    std::exit(1);
  }
  std::fclose(f);
}
static void write_ct(CryptoContextImpl& cc, const std::string& name, const Ciphertext& ct) {
  const size_t words = (size_t)ct->ncomp * ct->towers * cc.GetRingDimension();
  std::vector<uint64_t> h(words);
  if (fh_copy_d2h(cc.Native(), h.data(), ct->data->ptr, words * 8) != FH_OK) {
    std::fprintf(stderr, "copy failed: %s\n", fh_last_error());
    std::exit(1);
  }
  write_raw(name + ".u64", h.data(), words * 8);
}
static void write_cplx(const std::string& name, const std::vector<std::complex<double>>& v) {
  std::vector<double> f(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) f[2 * i] = v[i].real(), f[2 * i + 1] = v[i].imag();
  write_raw(name + ".f64", f.data(), f.size() * 8);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: bootstrap_check <out_dir> <logN> <slots> [seed]\n");
    return 2;
  }
  g_dir = argv[1];
  const int logN = std::atoi(argv[2]);
  const uint32_t slots = (uint32_t)std::atoi(argv[3]);
  const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 7;
  const uint32_t b1 = argc > 5 ? (uint32_t)std::atoi(argv[5]) : 1, b2 = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 1;
  const uint32_t depth = argc > 7 ? (uint32_t)std::atoi(argv[7]) : 15;
  const uint32_t sbits = argc > 8 ? (uint32_t)std::atoi(argv[8]) : 50;
  const uint32_t hw = argc > 9 ? (uint32_t)std::atoi(argv[9]) : 16;
  const bool flex = argc > 10 && std::atoi(argv[10]) == 1;
  const uint32_t meta_prec = argc > 11 ? (uint32_t)std::atoi(argv[11]) : 0;
  const uint32_t first_bits = argc > 12 ? (uint32_t)std::atoi(argv[12]) : 60;
  const uint32_t dnum = argc > 13 ? (uint32_t)std::atoi(argv[13]) : 3;
  const bool real_input = argc > 14 && std::atoi(argv[14]) == 1;
  const uint32_t in_level = argc > 15 ? (uint32_t)std::atoi(argv[15]) : 0;
  try {
    CCParams p;
    p.SetRingDim(1u << logN);
    p.SetMultiplicativeDepth(depth);
    p.SetScalingModSize(sbits);
    p.SetFirstModSize(first_bits);
    p.SetNumLargeDigits(dnum);
    p.SetSecretKeyDist(SecretKeyDist::SPARSE_TERNARY);
    p.SetHammingWeight(hw);
    p.SetBatchSize(slots);
    p.SetSeed(seed);
    p.SetScalingTechnique(flex ? ScalingTechnique::FLEXIBLEAUTO : ScalingTechnique::FIXEDMANUAL);
    auto cc = GenCryptoContext(p);
    auto kp = cc->KeyGen();
    cc->EvalMultKeyGen(kp.secretKey);
    cc->EvalBootstrapSetup({b1, b2}, {0, 0}, slots);
    cc->EvalBootstrapKeyGen(kp.secretKey, slots);
    const uint32_t L1 = cc->GetNumQ();

    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    cvec z(slots);
    for (auto& v : z) v = real_input ? std::complex<double>(U(rng), 0.0) : std::complex<double>(U(rng), U(rng));

    // a ciphertext with the towers the bootstrap needs left (one, two with a correction factor or
    // under FLEXIBLEAUTO, whose scale adjustment takes a rescale)
    const uint32_t tin = cc->GetBootstrapConstants().correction || flex ? 2 : 1;
    if (in_level > L1 - tin) throw std::runtime_error("input level leaves the bootstrap too few towers");
    auto pt = cc->MakeCKKSPackedPlaintext(z, 1, in_level ? in_level : L1 - tin, slots);
    auto ct = cc->Encrypt(kp.secretKey, pt);
    auto bt = cc->EvalBootstrap(ct);
    auto dec = cc->Decrypt(kp.secretKey, bt);
    const double e1 = max_err(dec->values, z, slots);
    // the two-iteration bootstrap (Meta-BTS) of the same ciphertext
    double e_meta = -1, meta_ms = -1;
    if (meta_prec) {
      cc->Synchronize();
      const auto t0 = std::chrono::steady_clock::now();
      auto bm = cc->EvalBootstrap(ct, 2, meta_prec);
      cc->Synchronize();
      meta_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      e_meta = max_err(cc->Decrypt(kp.secretKey, bm)->values, z, slots);
    }
    // determinism: a second bootstrap of the same ciphertext gives the same residues
    auto bt2 = cc->EvalBootstrap(ct);
    const size_t words = (size_t)2 * bt->towers * cc->GetRingDimension();
    std::vector<uint64_t> h1(words), h2(words);
    fh_copy_d2h(cc->Native(), h1.data(), bt->data->ptr, words * 8);
    fh_copy_d2h(cc->Native(), h2.data(), bt2->data->ptr, words * 8);
    const bool same = h1 == h2;
    // latency: median of three bootstraps (the device synchronized around each)
    std::vector<double> ms;
    for (int i = 0; i < 3; ++i) {
      cc->Synchronize();
      const auto t0 = std::chrono::steady_clock::now();
      auto b = cc->EvalBootstrap(ct);
      cc->Synchronize();
      ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(ms.begin(), ms.end());
    // the bootstrapped ciphertext computes: one more product
    auto sq = cc->Rescale(cc->EvalMult(bt, bt));   // (Rescale does nothing under FLEXIBLEAUTO; Decrypt rescales)
    auto dsq = cc->Decrypt(kp.secretKey, sq);
    cvec zz(slots);
    for (uint32_t i = 0; i < slots; ++i) zz[i] = z[i] * z[i];
    const double e2 = max_err(dsq->values, zz, slots);

    // Chebyshev series of degree 27 on [-2, 2] of a real input
    std::vector<double> xr(slots);
    cvec xc(slots), want(slots), fwant(slots);
    for (uint32_t i = 0; i < slots; ++i) {
      xr[i] = 2.0 * U(rng);
      xc[i] = xr[i];
    }
    const auto coeffs = CryptoContextImpl::EvalChebyshevCoefficients(sigmoid_like, nullptr, -2.0, 2.0, 27);
    for (uint32_t i = 0; i < slots; ++i) {
      want[i] = cheb_value(coeffs, xr[i] / 2.0);
      fwant[i] = sigmoid_like(xr[i], nullptr);
    }
    auto cx = cc->Encrypt(kp.secretKey, cc->MakeCKKSPackedPlaintext(xc, 1, 0, slots));
    auto cy = cc->EvalChebyshevSeries(cx, coeffs, -2.0, 2.0);
    auto dy = cc->Decrypt(kp.secretKey, cy);
    const double e3 = max_err(dy->values, want, slots), e4 = max_err(dy->values, fwant, slots);

    // everything the oracle replay needs
    const BootstrapConstants bc = cc->GetBootstrapConstants();
    write_cplx("z", z);
    write_cplx("x", xc);
    write_ct(*cc, "ct", ct);
    write_ct(*cc, "bt", bt);
    write_ct(*cc, "cx", cx);
    write_ct(*cc, "cy", cy);
    write_raw("cheb.f64", bc.cheb.data(), bc.cheb.size() * 8);
    write_raw("series.f64", coeffs.data(), coeffs.size() * 8);
    std::string groups;
    for (int t = 0; t < 2; ++t) {
      const auto& G = t ? bc.stc : bc.cts;
      groups += std::string(t ? ", \"stc\": [" : "\"cts\": [");
      for (size_t g = 0; g < G.size(); ++g) {
        std::vector<std::complex<double>> flat;
        for (const auto& d : G[g].diags) flat.insert(flat.end(), d.begin(), d.end());
        write_cplx(std::string(t ? "stc" : "cts") + std::to_string(g), flat);
        groups += (g ? ", " : "") + std::string("[") + std::to_string(G[g].rowSize) + ", " + std::to_string(G[g].stride) + ", " +
                  std::to_string(G[g].offset) + "]";
      }
      groups += "]";
    }
    {
      FILE* f = std::fopen((g_dir + "/meta.json").c_str(), "w");
      std::fprintf(f, "{\"flex\": %d, \"logN\": %d, \"slots\": %u, \"seed\": %llu, \"L1\": %u, \"K\": %u, \"dnum\": %u, \"hw\": %u, "
                      "\"scale\": %.17g, \"Krange\": %.17g, \"r\": %d, \"pre\": %.17g, \"scale_fix\": %.17g, \"correction\": %u, %s, \"moduli\": [",
                   flex ? 1 : 0, logN, slots, (unsigned long long)seed, L1, (unsigned)(cc->GetModuli().size() - L1), dnum, hw, cc->ScalingFactor(), bc.K, bc.r, bc.pre, bc.scale_fix,
                   bc.correction, groups.c_str());
      for (size_t i = 0; i < cc->GetModuli().size(); ++i)
        std::fprintf(f, "%s\"%llu\"", i ? ", " : "", (unsigned long long)cc->GetModuli()[i]);
      std::fprintf(f, "], \"bt_towers\": %u, \"cy_towers\": %u}\n", bt->towers, cy->towers);
      std::fclose(f);
    }

    std::printf("{\"technique\": \"%s\", \"logN\": %d, \"slots\": %u, \"budget\": [%u, %u], \"L1\": %u, \"depth\": %u, \"towers_in\": %u, \"towers_out\": %u, "
                "\"boot_err\": %.6g, \"boot_bits\": %.2f, \"deterministic\": %s, \"square_err\": %.6g, "
                "\"cheb_towers_out\": %u, \"cheb_err_vs_series\": %.6g, \"cheb_err_vs_function\": %.6g, \"boot_ms\": %.3f, "
                "\"meta_precision\": %u, \"meta_err\": %.6g, \"meta_bits\": %.2f, \"meta_ms\": %.3f}\n",
                flex ? "FLEXIBLEAUTO" : "FIXEDMANUAL", logN, slots, b1, b2, L1, cc->GetBootstrapDepth(), ct->towers, bt->towers, e1, -std::log2(e1), same ? "true" : "false", e2,
                cy->towers, e3, e4, ms[1], meta_prec, e_meta, e_meta > 0 ? -std::log2(e_meta) : 0.0, meta_ms);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
