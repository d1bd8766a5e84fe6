// Serving sessions with bootstrap requests: the continuous-batching loop (fh_serve_run) over a
// fideslib facade context whose EvalBootstrap serves the bootstrap requests through the loop's
// callback. Runs the same seeded session batched (mode 0), one call per operation (mode 1) and one
// request at a time (mode 2), after a discarded warm-up session, and prints the results as one JSON
// line with whether every request's result is the same in all three (tests/test_gpu_serve.py).
//
// usage: serve_check <logN> <slots> <cts_budget> <stc_budget> <depth> <scale_bits> <rate_per_s>
//                    <requests> <bootstrap_share> [seed] [inference_share] [hamming_weight] [technique]
//                    [concurrent] [ticks_in_flight]
// technique 1: the facade context (and so the bootstrap requests' EvalBootstrap) is FLEXIBLEAUTO;
// the loop's other request programs are integer programs at the C ABI, the same under either.
// concurrent 1: the bootstrap requests run on the loop's worker thread, through a second facade
// context built from the same parameters and seed (so the same keys) with its own stream, while
// the other requests keep ticking on the first; 2: the same with the worker yielding (one
// bootstrap operation at a time, between the loop's ticks). Every request's whole result is captured in each
// session and compared bit for bit.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../fideslib-fork_amd/host/fideslib.hpp"

using namespace fideslib;

struct BootUser {
  CryptoContextImpl* cc;
  uint32_t slots;
  int calls;
  bool flex;
  bool sync;  // return only once the output is written (the concurrent worker's contract)
};

// the loop's bootstrap callback: the request register as a facade ciphertext, bootstrapped
static int boot_cb(void* user, const uint64_t* in, int level, uint64_t* out, int* out_level) {
  auto* u = static_cast<BootUser*>(user);
  try {
    // a fresh request ciphertext carries its level's scaling factor
    const double scale = u->flex ? u->cc->GetScalingFactorReal(u->cc->GetNumQ() - 1 - (uint32_t)level) : u->cc->ScalingFactor();
    auto ct = u->cc->LoadCiphertext(in, (uint32_t)level + 1, 2, 1, scale, u->slots);
    auto bt = u->cc->EvalBootstrap(ct);
    u->cc->StoreCiphertext(bt, out);
    if (u->sync) u->cc->Synchronize();
    *out_level = (int)bt->towers - 1;
    ++u->calls;
    return 0;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "bootstrap callback: %s\n", ex.what());
    return 1;
  }
}

static std::string result_json(const fh_serve_result& r) {
  char b[512];
  std::snprintf(b, sizeof b,
                "{\"p50_ms\": %.4f, \"p99_ms\": %.4f, \"mean_ms\": %.4f, \"max_ms\": %.4f, \"requests_per_s\": %.3f, \"wall_s\": %.4f, "
                "\"mean_batch_ops\": %.3f, \"ticks\": %d, \"completed\": %d, \"inference\": %d, \"polynomial\": %d, "
                "\"bootstraps\": %d, \"max_active\": %d, \"checksum\": \"%llu\", \"p50_other_ms\": %.4f, \"p99_other_ms\": %.4f}",
                r.p50_ms, r.p99_ms, r.mean_ms, r.max_ms, r.requests_per_s, r.wall_s, r.mean_batch_ops, r.ticks, r.completed, r.inference,
                r.polynomial, r.bootstraps, r.max_active, (unsigned long long)r.checksum, r.p50_other_ms, r.p99_other_ms);
  return b;
}

int main(int argc, char** argv) {
  if (argc < 10) {
    std::fprintf(stderr, "usage: serve_check <logN> <slots> <b1> <b2> <depth> <scale_bits> <rate> <requests> <boot_share> [seed] [inf_share] [hw] [technique] [concurrent] [ticks_in_flight]\n");
    return 2;
  }
  const int logN = std::atoi(argv[1]);
  const uint32_t slots = (uint32_t)std::atoi(argv[2]);
  const uint32_t b1 = (uint32_t)std::atoi(argv[3]), b2 = (uint32_t)std::atoi(argv[4]);
  const uint32_t depth = (uint32_t)std::atoi(argv[5]), sbits = (uint32_t)std::atoi(argv[6]);
  const double rate = std::atof(argv[7]);
  const int requests = std::atoi(argv[8]);
  const double boot_share = std::atof(argv[9]);
  const uint64_t seed = argc > 10 ? std::strtoull(argv[10], nullptr, 10) : 5;
  const double inf_share = argc > 11 ? std::atof(argv[11]) : 0.5;
  const uint32_t hw = argc > 12 ? (uint32_t)std::atoi(argv[12]) : 16;
  const bool flex = argc > 13 && std::atoi(argv[13]) == 1;
  const int concurrent = argc > 14 ? std::atoi(argv[14]) : 0;
  const int ticks_in_flight = argc > 15 ? std::atoi(argv[15]) : 0;
  try {
    CCParams p;
    p.SetRingDim(1u << logN);
    p.SetMultiplicativeDepth(depth);
    p.SetScalingModSize(sbits);
    p.SetFirstModSize(60);
    p.SetNumLargeDigits(3);
    p.SetSecretKeyDist(SecretKeyDist::SPARSE_TERNARY);
    p.SetHammingWeight(hw);
    p.SetBatchSize(slots);
    p.SetSeed(seed);
    p.SetScalingTechnique(flex ? ScalingTechnique::FLEXIBLEAUTO : ScalingTechnique::FIXEDMANUAL);
    std::vector<int32_t> steps;
    for (int k = 0; k < 8; ++k) steps.push_back(1 << k);
    // the same key-generation sequence from the same seed gives the same keys in every context
    auto make = [&] {
      auto cc = GenCryptoContext(p);
      auto kp = cc->KeyGen();
      cc->EvalMultKeyGen(kp.secretKey);
      cc->EvalRotateKeyGen(kp.secretKey, steps);
      cc->EvalBootstrapSetup({b1, b2}, {0, 0}, slots);
      cc->EvalBootstrapKeyGen(kp.secretKey, slots);
      return cc;
    };
    auto cc = make();
    CryptoContext cc_boot = concurrent ? make() : cc;

    std::vector<const fh_key*> rk;
    std::vector<uint32_t> rg;
    for (int32_t s : steps) {
      rk.push_back(cc->RotationKeyHandle(s));
      uint32_t g = 0;
      if (fh_galois_index(cc->Native(), s, &g) != FH_OK) throw std::runtime_error(fh_last_error());
      rg.push_back(g);
    }
    BootUser user{cc_boot.get(), slots, 0, flex, concurrent != 0};
    fh_serve_result res[4];
    std::vector<std::vector<uint64_t>> per(4, std::vector<uint64_t>(requests, 0));
    // whole results of the batched, per-operation and alone sessions
    const size_t cap = 2ull * cc->GetNumQ() * cc->GetRingDimension();
    std::vector<std::vector<std::vector<uint64_t>>> outs(4);
    std::vector<std::vector<uint64_t*>> optr(4);
    std::vector<std::vector<uint64_t>> owords(4, std::vector<uint64_t>(requests, 0));
    for (int run = 1; run < 4; ++run) {
      outs[run].assign(requests, std::vector<uint64_t>(cap));
      for (int i = 0; i < requests; ++i) optr[run].push_back(outs[run][i].data());
    }
    // a warm-up session (batched, discarded) fills the context's buffer pool, the batch's table
    // staging and the bootstrap's first-use state, so that no measured session pays for them;
    // then batched, per operation, and one request at a time (each alone)
    for (int run = 0; run < 4; ++run) {
      const int mode = run == 0 ? 0 : run - 1;
      fh_serve_config cfg{};
      cfg.rate_per_s = rate;
      cfg.requests = requests;
      cfg.start_level = (int)cc->GetNumQ() - 1;
      cfg.hbm_budget_gb = 8.0;
      cfg.mode = mode;
      cfg.inference_share = inf_share;
      cfg.seed = seed;
      cfg.checksum = 1;
      cfg.bootstrap_share = boot_share;
      cfg.bootstrap = boot_cb;
      cfg.bootstrap_user = &user;
      cfg.request_checksums = per[run].data();
      cfg.bootstrap_concurrent = concurrent;
      cfg.ticks_in_flight = ticks_in_flight;
      if (run > 0) {
        cfg.request_outputs = optr[run].data();
        cfg.request_output_capacity = cap;
        cfg.request_output_words = owords[run].data();
      }
      if (fh_serve_run(cc->Native(), &cfg, rk.data(), rg.data(), cc->RelinKeyHandle(), &res[run]) != FH_OK) {
        std::fprintf(stderr, "fh_serve_run: %s\n", fh_last_error());
        return 1;
      }
    }
    const bool same = per[1] == per[3] && per[2] == per[3];
    bool bitwise = true;
    for (int i = 0; i < requests; ++i) {
      const size_t w = owords[3][i];
      bitwise = bitwise && w > 0 && owords[1][i] == w && owords[2][i] == w &&
                std::equal(outs[1][i].begin(), outs[1][i].begin() + w, outs[3][i].begin()) &&
                std::equal(outs[2][i].begin(), outs[2][i].begin() + w, outs[3][i].begin());
    }
    std::printf("{\"logN\": %d, \"slots\": %u, \"L1\": %u, \"boot_depth\": %u, \"rate_per_s\": %.3f, \"requests\": %d, "
                "\"bootstrap_share\": %.3f, \"callbacks\": %d, \"technique\": \"%s\", \"concurrent\": %d, \"ticks_in_flight\": %d, "
                "\"each_result_same_alone_and_in_mix\": %s, "
                "\"bitwise_same_alone_and_in_mix\": %s, \"batched\": %s, "
                "\"per_op\": %s, \"alone\": %s}\n",
                logN, slots, cc->GetNumQ(), cc->GetBootstrapDepth(), rate, requests, boot_share, user.calls, flex ? "FLEXIBLEAUTO" : "FIXEDMANUAL",
                concurrent, ticks_in_flight, same ? "true" : "false", bitwise ? "true" : "false",
                result_json(res[1]).c_str(), result_json(res[2]).c_str(), result_json(res[3]).c_str());
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
