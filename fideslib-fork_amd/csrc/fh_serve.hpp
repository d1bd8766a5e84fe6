// Continuous-batching serving loop over the mixed-level batch engine (fh_serve.cpp).
//
// Requests arrive as a seeded Poisson process. Each request is a fixed program of dependent
// ciphertext operations on its own working registers (device buffers taken from the context pool
// at admission and returned at completion):
//   inference ("dot product + degree-7 polynomial"): x * w (plaintext weights) + Rescale, a
//     rotate-and-sum over 8 doublings (rotations by 1, 2, ..., 128 and additions), then Horner's
//     rule for a degree-7 polynomial (ciphertext products with relinearization and Rescale,
//     plaintext coefficient additions, level alignment of the argument);
//   polynomial ("depth-6 polynomial"): Horner's rule for a degree-6 polynomial of the input;
//   bootstrap: one bootstrap of the input through a caller-supplied function (the bootstrap is
//     host-orchestrated above this layer), run either in the tick that reaches it (after the
//     tick's other operations) or, with bootstrap_concurrent, on a worker thread of the loop while
//     the other requests keep ticking (the function then works on a stream of its own: a second
//     context with the same keys), optionally yielding: one bootstrap operation at a time, issued
//     in the gaps between the loop's ticks.
// Admission is bounded by an HBM budget (each request's register footprint). Every scheduler
// tick gathers the next operation of every admitted request into one batch, whatever its kind
// and level, and executes it as one launch sequence (or, in the per-operation baseline mode, one
// operation call at a time); requests that arrive during a tick join at the next one. Ticks are
// issued without waiting for the GPU, up to ticks_in_flight ahead of it, and each is timed by an
// event on the stream: a request's latency ends at its last tick's completion on the GPU clock
// (or, for a concurrent bootstrap, when the bootstrap function returns).
// The reference has no request loop; its closest analogue is the concurrent-ops lanes of
// api/ConcurrentOps.hpp:39-62 and api/CryptoContext.hpp:174-195.
#pragma once
#include <cstdint>
#include <vector>

#include "fh_batch.hpp"

namespace fh {

constexpr int kDefaultTicksInFlight = 3;
constexpr int kDefaultBatchMinOps = 3;
constexpr double kYieldCapMs = 2.0;  // longest a yielding bootstrap operation waits for the loop

struct ServeConfig {
  double rate_per_s = 100.0;   // mean arrival rate
  int requests = 200;
  int start_level = 24;        // level of every incoming ciphertext
  double hbm_budget_gb = 64.0; // admission budget for request registers
  int mode = 0;                // 0: one fused batch per tick (ticks below batch_min_ops issued directly),
                               // 1: one call per operation, 2: one request at a time (each alone)
  double inference_share = 0.74;
  uint64_t seed = 1;
  bool checksum = false;       // sum every request's result words (to compare the two modes)
  // bootstrap requests: a share of the arrivals, each bootstrapping its input (2 x (start_level+1)
  // limbs) into `out` through `bootstrap`, which returns 0 on success and the output level
  double bootstrap_share = 0.0;
  int (*bootstrap)(void* user, const uint64_t* in, int level, uint64_t* out, int* out_level) = nullptr;
  void* bootstrap_user = nullptr;
  // optional, `requests` entries: the sum of request i's result words (mod 2^64) at index i
  uint64_t* request_checksums = nullptr;
  // optional, `requests` host buffers of request_output_capacity words: each result copied whole
  uint64_t* const* request_outputs = nullptr;
  size_t request_output_capacity = 0;
  uint64_t* request_output_words = nullptr;
  // bootstrap requests on a worker thread, concurrently with the other requests: `bootstrap` is
  // then called from that thread, must use a stream other than the context's and must return only
  // once `out` is written. 1: the worker issues the bootstrap freely; 2 (yielding): each engine
  // call the bootstrap makes waits until its previous one has finished and the loop has no tick
  // on the GPU (or kYieldCapMs have passed), so a tick shares the GPU with at most one bootstrap
  // operation
  int bootstrap_concurrent = 0;
  // ticks the host may issue ahead of the GPU (0: kDefaultTicksInFlight; 1: wait for each tick)
  int ticks_in_flight = 0;
  // mode 0: smallest tick issued as one batch (0: kDefaultBatchMinOps); smaller ticks run per operation
  int batch_min_ops = 0;
};

struct ServeResult {
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
  double requests_per_s = 0, wall_s = 0, mean_batch_ops = 0;
  int ticks = 0, completed = 0, inference = 0, polynomial = 0, bootstraps = 0, max_active = 0;
  uint64_t checksum = 0;       // sum of all result words mod 2^64 (when requested)
  double p50_other_ms = 0, p99_other_ms = 0;  // over the requests other than bootstraps
};

// rot_keys / rot_galois: the 8 keys of rotations by 2^0 .. 2^7.
ServeResult serve_run(Context& c, const ServeConfig& cfg, const uint64_t* const* rot_keys, const uint32_t* rot_galois,
                      const uint64_t* relin_key);

}  // namespace fh
