// Continuous-batching serving loop (see fh_serve.hpp).
#include "fh_serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <memory>
#include <random>
#include <string>
#include <thread>

#include "fh_keyswitch.hpp"

namespace fh {

namespace {

enum OpKind : uint8_t { kMulPtRs, kRotate, kAdd, kAddPt, kDrop, kMultRs, kBoot };

// One step of a request program. Registers: 0 = input x, 1 = t, 2 = r, 3 = y, 4 = aligned argument.
struct Step {
  OpKind kind;
  uint8_t dst, a, b;   // register indices (b: second operand register, or the plaintext index)
  uint8_t rot;         // rotation key index (kRotate)
  bool plain_b;        // b names a plaintext
};

// inference: t = x * w (+Rescale); 8 x (r = Rot(t, 2^k); t = t + r); Horner degree 7 in t
std::vector<Step> inference_program() {
  std::vector<Step> p;
  p.push_back({kMulPtRs, 1, 0, 0, 0, true});  // plaintext 0 = weights
  for (int k = 0; k < 8; ++k) {
    p.push_back({kRotate, 2, 1, 0, (uint8_t)k, false});
    p.push_back({kAdd, 1, 1, 2, 0, false});
  }
  p.push_back({kMulPtRs, 3, 1, 8, 0, true});  // y = t * c7
  p.push_back({kAddPt, 3, 3, 7, 0, true});    // y += c6
  for (int i = 5; i >= 0; --i) {
    p.push_back({kDrop, 4, 1, 3, 0, false});  // align t to y's level
    p.push_back({kMultRs, 3, 3, 4, 0, false});
    p.push_back({kAddPt, 3, 3, (uint8_t)(1 + i), 0, true});
  }
  return p;
}

// polynomial: Horner degree 6 in x
std::vector<Step> polynomial_program() {
  std::vector<Step> p;
  p.push_back({kMulPtRs, 3, 0, 7, 0, true});  // y = x * c6
  p.push_back({kAddPt, 3, 3, 6, 0, true});    // y += c5
  for (int i = 4; i >= 0; --i) {
    p.push_back({kDrop, 4, 0, 3, 0, false});
    p.push_back({kMultRs, 3, 3, 4, 0, false});
    p.push_back({kAddPt, 3, 3, (uint8_t)(1 + i), 0, true});
  }
  return p;
}

// bootstrap: y = Bootstrap(x)
std::vector<Step> bootstrap_program() { return {Step{kBoot, 3, 0, 0, 0, false}}; }

struct Request {
  int id = 0;
  int kind = 1;  // 0 polynomial, 1 inference, 2 bootstrap
  double arrival = 0;
  size_t pc = 0;
  bool in_boot = false;  // a bootstrap step running on the worker thread
  int level[5] = {0, 0, 0, 0, 0};
  DBuf reg[5];
};

// Timing events, created on demand and reused (single-threaded: the loop's thread only).
struct EventPool {
  std::vector<hipEvent_t> all, spare;
  hipEvent_t take() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    FH_HIP(hipEventCreate(&e));
    all.push_back(e);
    return e;
  }
  void give(hipEvent_t e) { spare.push_back(e); }
  ~EventPool() {
    for (hipEvent_t e : all) (void)hipEventDestroy(e);
  }
};

// Bootstrap steps served concurrently: a worker thread waits until the request's input is
// written (an event on the loop's stream), runs the caller's bootstrap function (which works on
// its own stream and returns once the output is written) and reports the completion time, while
// the loop keeps ticking the other requests.
class BootWorker {
 public:
  struct Job {
    Request* r;
    hipEvent_t ready;
  };
  struct Done {
    Request* r;
    hipEvent_t ready;
    int rc, level;
    double t;
  };
  BootWorker(const ServeConfig& cfg, std::chrono::steady_clock::time_point t0, hipStream_t loop_stream)
      : cfg_(cfg), t0_(t0), loop_stream_(loop_stream), th_([this] { loop(); }) {}
  ~BootWorker() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    th_.join();
  }
  void submit(const Job& j) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      q_.push_back(j);
    }
    cv_.notify_all();
  }
  bool pop(Done* d) {
    std::lock_guard<std::mutex> lk(mu_);
    if (done_.empty()) return false;
    *d = done_.front();
    done_.pop_front();
    return true;
  }
  // until a bootstrap completes or `seconds` pass
  void wait(double seconds) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait_for(lk, std::chrono::duration<double>(std::max(seconds, 0.0)), [&] { return !done_.empty(); });
  }

 private:
  // The yielding gate (run before each engine call of the bootstrap on this thread): the
  // bootstrap's previous operation has finished, and the loop's stream holds no tick (or
  // kYieldCapMs have passed, so that a loop that never idles does not starve the bootstrap).
  static void yield_gate(void* user, Context& bc) {
    const hipStream_t loop = static_cast<const BootWorker*>(user)->loop_stream_;
    FH_HIP(hipStreamSynchronize(bc.stream));
    const auto start = std::chrono::steady_clock::now();
    for (;;) {
      const hipError_t e = hipStreamQuery(loop);
      if (e == hipSuccess) return;
      if (e != hipErrorNotReady) FH_HIP(e);
      (void)hipGetLastError();  // a not-ready query leaves its status behind on this thread
      if (std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count() >= kYieldCapMs) return;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  void loop() {
    if (cfg_.bootstrap_concurrent == 2) thread_call_gate() = CallGate{&BootWorker::yield_gate, this};
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
        if (stop_) return;  // the loop is unwinding: queued bootstraps are dropped
        j = q_.front();
        q_.pop_front();
      }
      int lo = -1, rc = -1;
      if (hipEventSynchronize(j.ready) == hipSuccess)
        rc = cfg_.bootstrap(cfg_.bootstrap_user, j.r->reg[0].p, j.r->level[0], j.r->reg[3].p, &lo);
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
      {
        std::lock_guard<std::mutex> lk(mu_);
        done_.push_back({j.r, j.ready, rc, lo, t});
      }
      cv_done_.notify_all();
    }
  }
  const ServeConfig& cfg_;
  const std::chrono::steady_clock::time_point t0_;
  const hipStream_t loop_stream_;
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::deque<Job> q_;
  std::deque<Done> done_;
  bool stop_ = false;
  std::thread th_;  // last: started once the members above exist
};

// A tick issued on the stream: its completion event and the requests it completed.
struct Tick {
  hipEvent_t ev = nullptr;
  std::vector<std::unique_ptr<Request>> done;
};

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, (size_t)std::ceil(p * v.size()) - (p > 0 ? 1 : 0));
  return 1000.0 * v[i];
}

}  // namespace

ServeResult serve_run(Context& c, const ServeConfig& cfg, const uint64_t* const* rot_keys, const uint32_t* rot_galois,
                      const uint64_t* relin_key) {
  if (cfg.requests < 1 || cfg.rate_per_s <= 0) throw Error(-1, "serving needs requests >= 1 and a positive rate");
  const int L = cfg.start_level;
  if (L < 9 || L >= c.L1) throw Error(-1, "start level must leave room for the 8 levels a request consumes");
  for (int k = 0; k < 8; ++k)
    if (!rot_keys[k]) throw Error(-1, "null rotation key");
  if (!relin_key) throw Error(-1, "null relinearization key");
  if (cfg.ticks_in_flight < 0 || cfg.ticks_in_flight > 64) throw Error(-1, "ticks_in_flight must be in [0, 64]");
  const size_t N = c.N;
  const size_t reg_words = 2ull * (L + 1) * N;
  const hipStream_t st = c.stream;
  const int depth = cfg.ticks_in_flight > 0 ? cfg.ticks_in_flight : kDefaultTicksInFlight;
  if (cfg.batch_min_ops < 0) throw Error(-1, "batch_min_ops must be >= 0");
  if (cfg.bootstrap_concurrent < 0 || cfg.bootstrap_concurrent > 2)
    throw Error(-1, "bootstrap_concurrent must be 0 (in the tick), 1 (worker thread) or 2 (worker thread, yielding)");
  const int batch_min = cfg.batch_min_ops > 0 ? cfg.batch_min_ops : kDefaultBatchMinOps;

  // shared plaintexts: 0 = weights, 1..8 = polynomial coefficients c0..c7 (synthetic residues)
  DBuf pts(c.pool, 9 * (size_t)(L + 1) * N);
  {
    std::vector<int> ids(L + 1);
    for (int i = 0; i <= L; ++i) ids[i] = i;
    fill_uniform(c, make_set(pts.p, L + 1, ids.data(), 9, (uint64_t)(L + 1) * N), cfg.seed ^ 0x5eedULL, st);
  }
  auto pt = [&](int i) { return pts.p + (size_t)i * (L + 1) * N; };
  // per-operation mode scratch (one multiplication at a time)
  DBuf tmp(c.pool, reg_words);

  const std::vector<Step> progs[3] = {polynomial_program(), inference_program(), bootstrap_program()};
  if (cfg.bootstrap_share < 0 || cfg.bootstrap_share + cfg.inference_share > 1.0) throw Error(-1, "request shares must sum to at most 1");
  if (cfg.bootstrap_share > 0 && !cfg.bootstrap) throw Error(-1, "bootstrap requests need a bootstrap function");
  // a bootstrapped register can hold up to the whole chain
  const size_t boot_words = cfg.bootstrap_share > 0 ? 2ull * c.L1 * N : 0;

  // arrivals
  std::mt19937_64 rng(cfg.seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> arrival(cfg.requests);
  std::vector<int> kind(cfg.requests);
  double t = 0;
  for (int i = 0; i < cfg.requests; ++i) {
    t += -std::log(1.0 - U(rng)) / cfg.rate_per_s;
    arrival[i] = t;
    const double u = U(rng);
    kind[i] = u < cfg.inference_share ? 1 : (u >= 1.0 - cfg.bootstrap_share ? 2 : 0);
  }
  const size_t per_request = 5 * reg_words * 8;
  const int max_active = cfg.mode == 2 ? 1 : std::max(1, (int)std::min<double>(1e6, cfg.hbm_budget_gb * 1e9 / (double)per_request));

  // device-side store of the results (one slot of 2 x L1 limbs per request) when they are asked for
  const bool capturing = cfg.checksum || cfg.request_checksums || cfg.request_outputs;
  const size_t slot_words = 2ull * c.L1 * N;
  DBuf store;
  if (capturing) store = DBuf(c.pool, (size_t)cfg.requests * slot_words);
  std::vector<size_t> result_words(cfg.requests, 0);

  ServeResult res;
  std::vector<double> lat, lat_other;
  lat.reserve(cfg.requests);
  EventPool evs;
  std::deque<Tick> inflight;
  std::deque<std::unique_ptr<Request>> active;
  Batch batch(c);
  int next = 0, boots_running = 0;
  double ops_total = 0;
  using clock = std::chrono::steady_clock;
  FH_HIP(hipStreamSynchronize(st));
  const auto t0 = clock::now();
  auto now = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  // GPU-side time origin: the stream is idle, so the event completes at once
  const hipEvent_t ev0 = evs.take();
  FH_HIP(hipEventRecord(ev0, st));
  std::unique_ptr<BootWorker> worker;  // after every buffer it uses: stopped and joined first
  if (cfg.bootstrap_concurrent && cfg.bootstrap_share > 0) worker.reset(new BootWorker(cfg, t0, st));

  // A finished request's result is copied on the stream into its slot of a device-side store
  // (ordered after the work that wrote it) and read back after the session, so capturing results
  // never stalls the loop.
  auto capture = [&](Request* r) {
    if (!capturing) return;
    const size_t words = 2ull * (r->level[3] + 1) * N;
    if (words > slot_words) throw Error(-6, "a request's result exceeds its capture slot");
    if (cfg.request_outputs && cfg.request_outputs[r->id] && words > cfg.request_output_capacity)
      throw Error(-1, "request_output_capacity below a request's result");
    FH_HIP(hipMemcpyAsync(store.p + (size_t)r->id * slot_words, r->reg[3].p, words * 8, hipMemcpyDeviceToDevice, st));
    result_words[r->id] = words;
  };
  auto finish = [&](Request* r, double td) {
    lat.push_back(td - r->arrival);
    if (r->kind != 2) lat_other.push_back(td - r->arrival);
    ++res.completed;
    (r->kind == 1 ? res.inference : r->kind == 2 ? res.bootstraps : res.polynomial)++;
  };
  // the oldest tick in flight, once complete (wait: block until it is): its requests finish at
  // the tick's completion time on the GPU clock
  auto retire = [&](bool wait) -> bool {
    if (inflight.empty()) return false;
    Tick& k = inflight.front();
    if (wait) {
      FH_HIP(hipEventSynchronize(k.ev));
    } else {
      const hipError_t e = hipEventQuery(k.ev);
      if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // a pending event is not an error
        return false;
      }
      FH_HIP(e);
    }
    float ms = 0;
    FH_HIP(hipEventElapsedTime(&ms, ev0, k.ev));
    for (auto& r : k.done) finish(r.get(), ms / 1000.0);
    evs.give(k.ev);
    inflight.pop_front();
    return true;
  };
  // bootstraps the worker completed: the request takes its next step (the program's last)
  auto collect_boots = [&] {
    BootWorker::Done d;
    while (worker && worker->pop(&d)) {
      evs.give(d.ready);
      --boots_running;
      if (d.rc != 0 || d.level < 0) throw Error(-6, "the bootstrap function failed");
      Request* r = d.r;
      r->in_boot = false;
      r->level[3] = d.level;
      auto it = std::find_if(active.begin(), active.end(), [&](const std::unique_ptr<Request>& p) { return p.get() == r; });
      if (it == active.end()) throw Error(-6, "bootstrap completion for an unknown request");
      if (++r->pc == progs[r->kind].size()) {
        capture(r);  // the bootstrap function returned once the output was written
        std::unique_ptr<Request> up = std::move(*it);
        active.erase(it);
        finish(up.get(), d.t);
      }
    }
  };

  while (res.completed < cfg.requests) {
    while (retire(false)) {
    }
    collect_boots();
    if (res.completed >= cfg.requests) break;
    const double tn = now();
    while (next < cfg.requests && arrival[next] <= tn && (int)active.size() < max_active) {
      std::unique_ptr<Request> r(new Request);
      r->id = next;
      r->kind = kind[next];
      r->arrival = arrival[next];
      for (int k = 0; k < 5; ++k) r->reg[k] = DBuf(c.pool, k == 3 && r->kind == 2 ? std::max(reg_words, boot_words) : reg_words);
      std::vector<int> ids(L + 1);
      for (int i = 0; i <= L; ++i) ids[i] = i;
      // the request's input ciphertext (synthetic residues at the start level)
      fill_uniform(c, make_set(r->reg[0].p, L + 1, ids.data(), 2, (uint64_t)(L + 1) * N), cfg.seed * 1000003ull + next, st);
      r->level[0] = L;
      active.push_back(std::move(r));
      ++next;
    }
    res.max_active = std::max(res.max_active, (int)active.size());
    size_t runnable = 0;
    for (auto& r : active) runnable += !r->in_boot;
    if (runnable == 0) {
      if (retire(true)) continue;  // the loop condition sees what it completed
      // until the next arrival, unless none is left or admission is full
      const bool admit = next < cfg.requests && (int)active.size() < max_active;
      const double wait = admit ? arrival[next] - now() : 1.0;
      if (boots_running > 0) {
        worker->wait(std::min(wait, 1.0));
        continue;
      }
      if (next >= cfg.requests)
        throw Error(-6, "serving loop stalled: " + std::to_string(res.completed) + " of " + std::to_string(cfg.requests) +
                            " requests completed, " + std::to_string(active.size()) + " admitted");
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    // at most `depth` ticks in flight: the host builds the next tick while the GPU runs these
    if ((int)inflight.size() >= depth) retire(true);

    // one tick: the next step of every runnable request, as one batch from batch_min_ops
    // operations on (a single operation is issued directly: its batch would only add a table)
    const bool use_batch = cfg.mode == 0 && (int)runnable >= batch_min;
    if (use_batch) batch.reset();
    std::vector<Request*> boots;  // bootstraps run in the tick, after its other operations
    size_t ops = 0;
    for (auto& rp : active) {
      Request* r = rp.get();
      if (r->in_boot) continue;
      const Step& s = progs[r->kind][r->pc];
      if (s.kind == kBoot && worker) {
        // to the worker, once everything issued so far (the input included) is done
        const hipEvent_t ready = evs.take();
        FH_HIP(hipEventRecord(ready, st));
        r->in_boot = true;
        ++boots_running;
        worker->submit({r, ready});
        continue;
      }
      ++ops;
      const int la = r->level[s.a];
      uint64_t* a = r->reg[s.a].p;
      uint64_t* d = r->reg[s.dst].p;
      const uint64_t* b = s.plain_b ? pt(s.b) : r->reg[s.b].p;
      const size_t lq = la + 1;
      int ld = la;
      switch (s.kind) {
        case kMulPtRs:
          ld = la - 1;
          if (use_batch) batch.add_mult_plain(a, b, la, true, d);
          else {
            std::vector<int> ids(lq);
            for (size_t i = 0; i < lq; ++i) ids[i] = (int)i;
            ew_binary(c, 2, make_set(a, lq, ids.data(), 2, lq * N), make_set(const_cast<uint64_t*>(b), lq, ids.data(), 2, 0),
                      make_set(tmp.p, lq, ids.data(), 2, lq * N), st);
            rescale(c, tmp.p, la, 2, d, st);
          }
          break;
        case kRotate:
          if (use_batch) batch.add_rotate(a, la, rot_keys[s.rot], rot_galois[s.rot], d);
          else eval_rotate(c, a, la, rot_keys[s.rot], rot_galois[s.rot], d, st);
          break;
        case kAdd:
          if (r->level[s.b] != la) throw Error(-6, "serving program adds operands at different levels");
          if (use_batch) batch.add_add(a, b, la, 2, d);
          else {
            std::vector<int> ids(2 * lq);
            for (size_t i = 0; i < 2 * lq; ++i) ids[i] = (int)(i % lq);
            ew_binary(c, 0, make_set(a, 2 * lq, ids.data()), make_set(const_cast<uint64_t*>(b), 2 * lq, ids.data()),
                      make_set(d, 2 * lq, ids.data()), st);
          }
          break;
        case kAddPt:
          if (use_batch) batch.add_add_plain(a, b, la, d);
          else {
            std::vector<int> ids(lq);
            for (size_t i = 0; i < lq; ++i) ids[i] = (int)i;
            ew_binary(c, 0, make_set(a, lq, ids.data()), make_set(const_cast<uint64_t*>(b), lq, ids.data()), make_set(d, lq, ids.data()), st);
            if (d != a) FH_HIP(hipMemcpyAsync(d + lq * N, a + lq * N, lq * N * 8, hipMemcpyDeviceToDevice, st));
          }
          break;
        case kDrop: {
          ld = r->level[s.b];  // the level of the register to align with
          if (use_batch) batch.add_level_reduce(a, la, ld, 2, d);
          else {
            const size_t lo = ld + 1;
            for (int comp = 0; comp < 2; ++comp)
              FH_HIP(hipMemcpyAsync(d + comp * lo * N, a + comp * lq * N, lo * N * 8, hipMemcpyDeviceToDevice, st));
          }
          break;
        }
        case kBoot:
          boots.push_back(r);  // after the tick's other operations
          ld = -1;
          break;
        case kMultRs:
          if (r->level[s.b] != la) throw Error(-6, "serving program multiplies operands at different levels");
          ld = la - 1;
          if (use_batch) batch.add_mult(a, b, la, relin_key, true, d);
          else {
            eval_mult(c, a, b, la, relin_key, tmp.p, st);
            rescale(c, tmp.p, la, 2, d, st);
          }
          break;
      }
      if (s.kind != kBoot) r->level[s.dst] = ld;
    }
    if (ops == 0) continue;  // every runnable request went to the worker
    ops_total += (double)ops;
    if (use_batch && batch.items() > 0) {
      batch.finalize();
      batch.run(st, false);
    }
    for (Request* r : boots) {
      int lo = -1;
      if (cfg.bootstrap(cfg.bootstrap_user, r->reg[0].p, r->level[0], r->reg[3].p, &lo) != 0 || lo < 0)
        throw Error(-6, "the bootstrap function failed");
      r->level[3] = lo;
    }
    ++res.ticks;
    Tick tick;
    // the stepped requests advance; those at the end of their program have their results
    // captured and leave with the tick
    for (auto it = active.begin(); it != active.end();) {
      Request* r = it->get();
      if (!r->in_boot && ++r->pc == progs[r->kind].size()) {
        capture(r);
        tick.done.push_back(std::move(*it));
        it = active.erase(it);
      } else {
        ++it;
      }
    }
    tick.ev = evs.take();
    FH_HIP(hipEventRecord(tick.ev, st));
    inflight.push_back(std::move(tick));
  }
  res.wall_s = now();
  if (capturing) {
    FH_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> h(slot_words);
    for (int i = 0; i < cfg.requests; ++i) {
      const size_t words = result_words[i];
      FH_HIP(hipMemcpy(h.data(), store.p + (size_t)i * slot_words, words * 8, hipMemcpyDeviceToHost));
      uint64_t sum = 0;
      for (size_t k = 0; k < words; ++k) sum += h[k];
      res.checksum += sum;
      if (cfg.request_checksums) cfg.request_checksums[i] = sum;
      if (cfg.request_outputs && cfg.request_outputs[i]) {
        std::copy(h.begin(), h.begin() + words, cfg.request_outputs[i]);
        if (cfg.request_output_words) cfg.request_output_words[i] = words;
      }
    }
  }
  res.p50_ms = percentile(lat, 0.50);
  res.p99_ms = percentile(lat, 0.99);
  res.p50_other_ms = percentile(lat_other, 0.50);
  res.p99_other_ms = percentile(lat_other, 0.99);
  res.max_ms = 1000.0 * *std::max_element(lat.begin(), lat.end());
  double sum = 0;
  for (double v : lat) sum += v;
  res.mean_ms = 1000.0 * sum / lat.size();
  res.requests_per_s = cfg.requests / (res.wall_s - arrival[0]);
  res.mean_batch_ops = ops_total / std::max(1, res.ticks);
  return res;
}

}  // namespace fh
