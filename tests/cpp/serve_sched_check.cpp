// The serving loop's scheduler (fideslib-fork_amd/csrc/fh_serve.cpp) on the CPU: the HIP runtime
// calls it makes and the library operations it issues are replaced by host stand-ins, so that the
// host logic (admission, ticks issued ahead of the GPU and retired by their events, bootstrap
// requests served in the tick or on the worker thread) runs without a GPU. A stream completes its
// events in order, at a random pace while they are polled; each rescaled product adds one to the
// first word of its output register.
//
// Checks, for every mode (batched, per operation, one request at a time), bootstrap placement
// (in the tick, on the worker, on the yielding worker) and ticks_in_flight (1, 3, 16): every
// request completes, with a result written and the same result as in the first session; latencies
// are ordered; bootstrap callbacks run on the worker thread exactly when asked; on the yielding
// worker each engine call of a bootstrap passes the call gate, which returns with the bootstrap's
// stream drained and the loop's stream idle or its wait cap spent (and no gate is set elsewhere);
// the failure of a bootstrap surfaces as an error. Prints one JSON line per session and "ok" at the end; exits non-zero on a failed check.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <thread>

#include "../../fideslib-fork_amd/csrc/fh_keyswitch.hpp"
#include "../../fideslib-fork_amd/csrc/fh_serve.hpp"

using namespace fh;

namespace {
std::mutex g_mu;  // the worker thread records events too
std::mt19937 g_rng(1);
// A stream completes its work in order: an event is complete once the stream's completion mark
// has reached its sequence number. Polling a pending event sometimes advances the mark.
struct Ev {
  hipStream_t s = nullptr;
  uint64_t seq = 0;
  double t = 0;
};
std::map<hipEvent_t, Ev> g_ev;
std::map<hipStream_t, uint64_t> g_issued, g_done;
std::map<hipStream_t, std::vector<hipEvent_t>> g_pending;  // recorded, not yet complete, in order
const auto g_t0 = std::chrono::steady_clock::now();
double host_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - g_t0).count(); }
int g_fail = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "check failed: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");                         \
      ++g_fail;                                           \
    }                                                     \
  } while (0)
bool complete(const Ev& v) { return v.seq <= g_done[v.s]; }
// the stream completes its work up to `mark`: the events it passes take the current time
void advance(hipStream_t s, uint64_t mark) {
  g_done[s] = std::max(g_done[s], mark);
  auto& q = g_pending[s];
  size_t k = 0;
  while (k < q.size() && g_ev[q[k]].seq <= g_done[s]) g_ev[q[k++]].t = host_s();
  q.erase(q.begin(), q.begin() + k);
}
}  // namespace

extern "C" {
hipError_t hipEventCreate(hipEvent_t* e) {
  std::lock_guard<std::mutex> lk(g_mu);
  *e = reinterpret_cast<hipEvent_t>(new int(0));
  g_ev[*e] = Ev();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_ev.erase(e);
  delete reinterpret_cast<int*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  const bool idle = g_done[s] == g_issued[s];
  g_ev[e] = Ev{s, ++g_issued[s], 0};
  g_pending[s].push_back(e);
  if (idle) advance(s, g_issued[s]);  // an idle stream passes the event at once
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(g_mu);
  Ev& v = g_ev[e];
  if (!complete(v) && g_rng() % 3 == 0) advance(v.s, std::min(g_issued[v.s], g_done[v.s] + 1 + g_rng() % 3));
  return complete(v) ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(g_mu);
  Ev& v = g_ev[e];
  advance(v.s, v.seq);
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_mu);
  const Ev &x = g_ev[a], &y = g_ev[b];
  if (!complete(x) || !complete(y)) return hipErrorNotReady;
  *ms = (float)(1000.0 * (y.t - x.t));
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "stand-in runtime error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  advance(s, g_issued[s]);
  return hipSuccess;
}
thread_local int g_last_query = 0;  // this thread's last stream query: 1 idle, 2 busy
hipError_t hipStreamQuery(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[s] < g_issued[s] && g_rng() % 3 == 0) advance(s, std::min(g_issued[s], g_done[s] + 1 + g_rng() % 3));
  g_last_query = g_done[s] == g_issued[s] ? 1 : 2;
  return g_done[s] == g_issued[s] ? hipSuccess : hipErrorNotReady;
}
}

namespace {
bool stream_idle(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_done[s] == g_issued[s];
}
// the bootstrap's stand-in context: its own stream (a distinct handle), with work pending
hipStream_t g_boot_stream = reinterpret_cast<hipStream_t>(0x5eed);
void boot_issue() {
  std::lock_guard<std::mutex> lk(g_mu);
  ++g_issued[g_boot_stream];
}
}  // namespace

// host stand-ins of the library: the rescaling steps (plaintext and ciphertext products) add one to
// the first word of their output register, so a result's first word counts the rescaled products
// that wrote it (the same in every mode)
namespace fh {
Context::Context(int logN, int L1_, int K_, int dnum_, const std::vector<uint64_t>&) {
  N = 1u << logN;
  L1 = L1_;
  K = K_;
  dnum = dnum_;
}
Context::~Context() {}
Pool::~Pool() {}
CallGate& thread_call_gate() {
  thread_local CallGate g;
  return g;
}
Profiler::~Profiler() {}
UploadRing::~UploadRing() {}
void* Pool::get(size_t b) { return std::calloc(1, b); }
void Pool::put(void* p) { std::free(p); }
static void step(uint64_t* dst) { dst[0] += 1; }
void fill_uniform(const Context&, const LimbSet&, uint64_t, hipStream_t) {}
void ew_binary(const Context&, int, const LimbSet&, const LimbSet&, const LimbSet&, hipStream_t) {}
void rescale(Context&, const uint64_t*, int, int, uint64_t* out, hipStream_t) { step(out); }
void eval_rotate(Context&, const uint64_t*, int, const uint64_t*, uint32_t, uint64_t*, hipStream_t) {}
void eval_mult(Context&, const uint64_t*, const uint64_t*, int, const uint64_t*, uint64_t*, hipStream_t) {}
struct BatchPlan {
  std::vector<uint64_t*> outs;
};
Batch::Batch(Context& c) : c_(c), p_(new BatchPlan) {}
Batch::~Batch() = default;
void Batch::reset() { p_->outs.clear(); }
void Batch::add_rotate(const uint64_t*, int, const uint64_t*, uint32_t, uint64_t*) { p_->outs.push_back(nullptr); }
void Batch::add_mult(const uint64_t*, const uint64_t*, int, const uint64_t*, bool, uint64_t* out, int, int) { p_->outs.push_back(out); }
void Batch::add_add(const uint64_t*, const uint64_t*, int, int, uint64_t*) { p_->outs.push_back(nullptr); }
void Batch::add_mult_plain(const uint64_t*, const uint64_t*, int, bool, uint64_t* out) { p_->outs.push_back(out); }
void Batch::add_add_plain(const uint64_t*, const uint64_t*, int, uint64_t*) { p_->outs.push_back(nullptr); }
void Batch::add_level_reduce(const uint64_t*, int, int, int, uint64_t*) { p_->outs.push_back(nullptr); }
void Batch::finalize() {}
void Batch::run(hipStream_t, bool) {
  for (uint64_t* o : p_->outs)
    if (o) step(o);
}
size_t Batch::items() const { return p_->outs.size(); }
}  // namespace fh

namespace {
struct BootUser {
  std::thread::id main_id;
  std::atomic<int> calls{0}, on_worker{0}, gated{0}, gate_bad{0};
  int fail_at = -1;  // the call that reports failure (-1: none)
  Context* boot = nullptr;  // the bootstrap's own context (stand-in)
  hipStream_t loop = nullptr;
};
constexpr int kBootEngineCalls = 4;
int boot_cb(void* user, const uint64_t*, int level, uint64_t* out, int* out_level) {
  auto* u = static_cast<BootUser*>(user);
  const int k = u->calls++;
  if (std::this_thread::get_id() != u->main_id) ++u->on_worker;
  // the bootstrap's engine calls: each runs the thread's call gate first (as the C ABI does), then
  // issues work on the bootstrap's stream
  for (int i = 0; i < kBootEngineCalls; ++i) {
    const CallGate& g = thread_call_gate();
    if (g.fn) {
      const auto t = std::chrono::steady_clock::now();
      g_last_query = 0;
      g.fn(g.user, *u->boot);
      const double waited = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
      ++u->gated;
      // the gate returns once the loop's stream was seen idle (its last query) or the cap is spent
      if (!stream_idle(g_boot_stream) || (g_last_query != 1 && waited < kYieldCapMs)) ++u->gate_bad;
    }
    boot_issue();
    std::this_thread::sleep_for(std::chrono::microseconds(75));  // a long step
  }
  if (k == u->fail_at) return 1;
  out[0] = 1000 + (uint64_t)level;
  *out_level = level + 4;
  return 0;
}
}  // namespace

int main() {
  Context c(8, 16, 2, 3, {});
  const int requests = 40;
  uint64_t keys[9] = {0};
  const uint64_t* rk[8];
  uint32_t rg[8];
  for (int i = 0; i < 8; ++i) {
    rk[i] = &keys[i];
    rg[i] = 5;
  }
  Context boot(8, 16, 2, 3, {});
  boot.stream = g_boot_stream;
  const size_t cap = 2ull * c.L1 * c.N;
  std::vector<uint64_t> words_ref;
  for (int mode = 0; mode < 3; ++mode)
    for (int conc = 0; conc < 3; ++conc)
      for (int tif : {1, 3, 16}) {
        BootUser u;
        u.main_id = std::this_thread::get_id();
        u.boot = &boot;
        u.loop = c.stream;
        std::vector<std::vector<uint64_t>> outs(requests, std::vector<uint64_t>(cap));
        std::vector<uint64_t*> optr(requests);
        for (int i = 0; i < requests; ++i) optr[i] = outs[i].data();
        std::vector<uint64_t> words(requests, 0);
        ServeConfig cfg;
        cfg.rate_per_s = 4000;
        cfg.requests = requests;
        cfg.start_level = 11;
        cfg.mode = mode;
        cfg.inference_share = 0.5;
        cfg.bootstrap_share = 0.25;
        cfg.bootstrap = boot_cb;
        cfg.bootstrap_user = &u;
        cfg.bootstrap_concurrent = conc;
        cfg.ticks_in_flight = tif;
        cfg.seed = 7;
        cfg.request_outputs = optr.data();
        cfg.request_output_capacity = cap;
        cfg.request_output_words = words.data();
        ServeResult r;
        try {
          r = serve_run(c, cfg, rk, rg, &keys[8]);
        } catch (const std::exception& e) {
          CHECK(false, "mode %d concurrent %d ticks_in_flight %d: %s", mode, conc, tif, e.what());
          continue;
        }
        std::printf("{\"mode\": %d, \"concurrent\": %d, \"ticks_in_flight\": %d, \"completed\": %d, \"ticks\": %d, \"bootstraps\": %d, "
                    "\"inference\": %d, \"polynomial\": %d, \"max_active\": %d, \"callbacks\": %d, \"on_worker\": %d, \"p50_ms\": %.3f, "
                    "\"p99_ms\": %.3f, \"p99_other_ms\": %.3f, \"gated\": %d}\n",
                    mode, conc, tif, r.completed, r.ticks, r.bootstraps, r.inference, r.polynomial, r.max_active, u.calls.load(),
                    u.on_worker.load(), r.p50_ms, r.p99_ms, r.p99_other_ms, u.gated.load());
        CHECK(r.completed == requests, "completed %d", r.completed);
        CHECK(r.inference + r.polynomial + r.bootstraps == requests, "kinds do not add up");
        CHECK(r.bootstraps > 0 && u.calls == r.bootstraps, "callbacks %d for %d bootstraps", u.calls.load(), r.bootstraps);
        CHECK(u.on_worker == (conc ? r.bootstraps : 0), "callbacks on the worker: %d", u.on_worker.load());
        CHECK(u.gated == (conc == 2 ? kBootEngineCalls * r.bootstraps : 0), "gated engine calls: %d", u.gated.load());
        CHECK(u.gate_bad == 0, "the gate returned with work in flight %d times", u.gate_bad.load());
        CHECK(!thread_call_gate().fn, "a call gate is set on the loop's thread");
        CHECK(0 <= r.p50_ms && r.p50_ms <= r.p99_ms && r.p99_ms <= r.max_ms, "latency order");
        if (mode == 2) CHECK(r.max_active == 1, "one request at a time");
        // every request's result written once, with the step count of its program
        for (int i = 0; i < requests; ++i) CHECK(words[i] > 0, "request %d has no result", i);
        std::vector<uint64_t> w0(requests);
        for (int i = 0; i < requests; ++i) w0[i] = outs[i][0];
        if (words_ref.empty()) words_ref = w0;
        CHECK(w0 == words_ref, "results differ from the first session (mode %d concurrent %d ticks_in_flight %d)", mode, conc, tif);
      }
  // a failing bootstrap surfaces as an error, in the tick and on the worker
  for (int conc = 0; conc < 3; ++conc) {
    BootUser u;
    u.main_id = std::this_thread::get_id();
    u.boot = &boot;
    u.loop = c.stream;
    u.fail_at = 1;
    ServeConfig cfg;
    cfg.rate_per_s = 4000;
    cfg.requests = requests;
    cfg.start_level = 11;
    cfg.bootstrap_share = 0.25;
    cfg.bootstrap = boot_cb;
    cfg.bootstrap_user = &u;
    cfg.bootstrap_concurrent = conc;
    cfg.seed = 7;
    bool threw = false;
    try {
      serve_run(c, cfg, rk, rg, &keys[8]);
    } catch (const std::exception& e) {
      threw = std::strstr(e.what(), "bootstrap") != nullptr;
    }
    CHECK(threw, "a failing bootstrap (concurrent %d) did not surface", conc);
  }
  if (g_fail) {
    std::printf("FAILED: %d checks\n", g_fail);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
