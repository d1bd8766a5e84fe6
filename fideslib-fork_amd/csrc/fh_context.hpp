// Device context for the RNS-CKKS evaluation path: primes, twiddle tables and
// every basis-conversion constant, resident in HBM for the context's lifetime.
//
// Prime indices: 0..L1-1 are the Q chain (q_0 first), L1..L1+K-1 the special
// primes P. A polynomial over Q_l ∪ P keeps its limbs in the order
// q_0..q_l, p_0..p_{K-1}; every limb is N contiguous uint64 words in the
// bit-reversed EVALUATION order of OpenFHE (no reordering anywhere).
// Reference: src/ConstantsGPU.cuh:49-198 (constants), src/CKKS/Context.cuh:56-424.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fh_arith.hpp"

namespace fh {

constexpr int kMaxPrimes = 72;   // Q + P primes a context may hold
constexpr int kMaxLimbs = 72;    // limbs of one polynomial in one launch descriptor

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FH_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::fh::Error(-5, std::string("HIP error: ") + hipGetErrorString(e_) + " at " +      \
                                 __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

// ---- host number theory (product-side; u128 arithmetic) ----
uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t q);
uint64_t h_invmod(uint64_t a, uint64_t q);
uint64_t h_shoup(uint64_t w, uint64_t q);  // floor(w 2^64 / q)
uint64_t h_inv_odd(uint64_t g, uint64_t m); // inverse of odd g modulo a power of two m
bool h_is_prime(uint64_t n);
uint64_t h_min_root_of_unity(uint64_t m, uint64_t q);
// OpenFHE FIXEDMANUAL modulus chain (L1 primes) + K special primes
std::vector<uint64_t> h_gen_primes(int logN, int L1, int first_bits, int scale_bits, int K, int aux_bits);
int h_default_num_special(const std::vector<uint64_t>& q, int L1, int dnum, int aux_bits);

// Simple size-bucketed caching allocator over hipMalloc (stream-ordered reuse on the context stream).
class Pool {
 public:
  ~Pool();
  void* get(size_t bytes);
  void put(void* p);
  // p is returned once the work queued on st so far has completed: an event recorded on st marks
  // it, and get() reuses the block only after that event has completed, so no caller on any
  // stream (or the host) receives a block that queued work still reads or writes. The host does
  // not wait, unless more than kPendingCap bytes are pending (then get() waits for the oldest).
  void put_after(void* p, hipStream_t st);
  void release_all();
  // while held, returned blocks are parked instead of reused: work issued on two streams between
  // hold() and release_held() never gets a block another stream may still be using
  void hold();
  void release_held();
  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_cached() const { return cached_; }

 private:
  struct Pending {
    void* p;
    size_t bytes;
    hipEvent_t ev;
  };
  static constexpr size_t kPendingCap = (size_t)8 << 30;
  void reap(bool wait_one);  // moves completed pending blocks (oldest first) to the free list
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> sizes_;
  std::deque<Pending> pending_;
  std::vector<hipEvent_t> spare_ev_;
  size_t pending_bytes_ = 0;
  int hold_ = 0;
  std::vector<void*> held_;
  size_t in_use_ = 0, cached_ = 0;
};

// Pinned host staging for small host-to-device uploads (batch job tables): a ring of slots, each
// rewritten only once the copy that read it has completed, so an upload never waits for work
// queued before it. Slots are allocated on first use and kept for the context's lifetime.
class UploadRing {
 public:
  ~UploadRing();
  // Copy `bytes` from host `src` to device `dst` on `st` through the next slot; returns the event
  // recorded after the copy (valid until that slot's next upload, which completes no earlier).
  hipEvent_t upload(void* dst, const void* src, size_t bytes, hipStream_t st);

 private:
  static constexpr int kSlots = 8;
  struct Slot {
    void* h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
  };
  Slot slot_[kSlots];
  int next_ = 0;
  std::mutex mu_;
};

// RAII device buffer from the pool.
struct DBuf {
  Pool* pool = nullptr;
  uint64_t* p = nullptr;
  size_t words = 0;
  DBuf() = default;
  DBuf(Pool& pl, size_t w) : pool(&pl), p((uint64_t*)pl.get(w * 8)), words(w) {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool(o.pool), p(o.p), words(o.words) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    reset();
    pool = o.pool; p = o.p; words = o.words; o.p = nullptr;
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() { if (p && pool) pool->put(p); p = nullptr; }
};

// Per-kernel-family launch timing with HIP events on the launching stream
// (enabled on demand; the reference wraps ops in NVTX ranges, src/CudaUtils.cuh:664-694).
struct Profiler {
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  bool on = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> spare;
  std::mutex mu;
  hipEvent_t take();
  void give(hipEvent_t e) { spare.push_back(e); }
  ~Profiler();
};

// Launch-shape tunables (set by the autotuner or by the caller through fh_tune_set).
struct Tune {
  // limb data per NTT launch pair, in MB. Measured at N=2^16 (3200 limbs): 96 MB chunks take 3.97 ms
  // per forward+inverse, one launch pair over all limbs 3.23 ms; launch tails outweigh MALL reuse.
  int ntt_chunk_mb = 1 << 16;
  // 1: spread the NTT jobs of 60-bit primes (integer Shoup arithmetic, VALU-heavy) evenly among
  // the jobs of 50-bit primes (fp64 arithmetic, memory-bound) in ModUp target and batch launches,
  // instead of running them as one prime-major block at the end.
  int ntt_mix = 0;
  // NTT radix split: log2 of the column-pass length (0: floor(logN / 2)); see ntt_split_ok
  int ntt_split = 0;
  // workgroup size (256 or 512 threads) of the forward column, forward row, inverse row and
  // inverse column passes; 512 applies with the default split only
  int ntt_threads[4] = {256, 256, 256, 256};
  // 2: a batched rotation of 8 or more ciphertexts runs as two halves on two streams (ModUp,
  // key inner product and ModDown of one half can overlap those of the other)
  int rot_streams = 1;
  // 1: a batched rotation runs the row stages of its ModUp targets' forward NTT fused with the key
  // inner product (the transformed digits are not written back); 0: separate row pass and key
  // inner product. Measured at N=2^16 (64 ciphertexts, level 24): the fused kernel takes 3.62 ms
  // per step against 2.91 ms for the two it replaces. Its per-item accumulators stay live through
  // the row stages (240 VGPRs, 2 waves per SIMD), and the work is VALU-bound, so the bytes it
  // saves do not pay; off by default.
  int rot_fused = 0;
  // key-switch items per workgroup group of a batch's key inner product (items sharing a key are
  // cut into chunks of this many; each chunk reads the key once). On the 256-operation mixed-level
  // stream at N=2^16, 4 to 64 measured within 1% of each other.
  int ksk_group_chunk = 16;
  // 2: the chunks of a chunked NTT (ntt_chunk_mb below the batch) alternate between two streams
  int ntt_streams = 1;
  // 1: NTTs with equal column and row lengths run as one launch of workgroup teams (one HBM round
  // trip per limb: the intermediate stays in the L2 of the team's XCD); 0: two launches
  int ntt_team = 0;
  // teams per XCD of a team launch (each team holds one limb of intermediate in its XCD's L2)
  int ntt_team_per_xcd = 4;
};

struct Context;
// Records an event pair around the launches issued in its scope when profiling is on.
struct ProfScope {
  const Context& c;
  const char* name;
  hipStream_t st;
  hipEvent_t b = nullptr;
  ProfScope(const Context& ctx, const char* n, hipStream_t s);
  ~ProfScope();
};

struct Context {
  int logN = 0;
  uint32_t N = 0;
  int device = -1;   // HIP device current at creation (set by noalloc_init)
  int L1 = 0, K = 0, dnum = 0, alpha = 0, np = 0;
  std::vector<uint64_t> q, psi;
  std::vector<uint8_t> fp64;   // per prime: 1 when its NTT runs in fp64 arithmetic

  // device tables
  PrimeConst* d_prime = nullptr;   // [np]
  uint64_t* d_tw = nullptr;        // [np][N] psi^brev(k)
  uint64_t* d_tw_sh = nullptr;     // Shoup quotients
  uint64_t* d_itw = nullptr;       // [np][N] psi^-brev(k)
  uint64_t* d_itw_sh = nullptr;
  double* d_twd = nullptr;         // [np][N] twiddles as doubles (fp64 NTT path, q < 2^51)
  double* d_itwd = nullptr;
  // Row twists of the two-pass NTT for radix split l1 (N = S1 x S2, S1 = 2^l1): after the column
  // stages, row r is finished by a standard S2-point negacyclic transform of x_c tau_r^c, with
  // tau_r = psi^(1 + 2 brev_l1(r) - S1); tw[r S2 + c] = tau_r^c, itw its inverse. Built at
  // creation for the default split, on first use for another (ensure_twists).
  struct TwistSet {
    uint64_t* tw = nullptr;        // [np][N]
    uint64_t* tw_sh = nullptr;
    uint64_t* itw = nullptr;
    uint64_t* itw_sh = nullptr;
    double* twd = nullptr;         // centred doubles (fp64 path)
    double* itwd = nullptr;
  };
  TwistSet twists[9];
  void ensure_twists(int l1);

  // ModUp: for level l and digit d, source prescale [Qhat_i^-1]_{q_i} (+Shoup), and
  // [Qhat_i]_{p_j} for every extended position j of level l.
  uint64_t* d_modup_qhinv = nullptr;     // [L1][dnum][alpha]
  uint64_t* d_modup_qhinv_sh = nullptr;  // [L1][dnum][alpha]
  uint64_t* d_modup_qhat = nullptr;      // [L1][dnum][alpha][L1+K]
  uint64_t* d_modup_qhat_m = nullptr;    // same, Montgomery form [Qhat_i 2^64]_{p_j}
  uint64_t* d_md_phat_m = nullptr;       // [K][L1] Montgomery form [P^_k 2^64]_{q_i}
  bool redc_conv_ok = false;             // conversion sums fit a single Montgomery reduction
  bool split_mac_ok = false;             // every prime < 2^60 and alpha, K <= 16: 30-bit split products
  // ReducedNoise basis conversions (the reference engine's reducedNoise option): prescaled
  // residues above q/2 read as v - q, i.e. the sums gain -count * Q_src. Off by default.
  bool reduced_noise = false;
  bool rn_conv_ok = false;               // one more term still fits a single Montgomery reduction
  uint64_t* d_modup_negq_m = nullptr;    // [L1][dnum][L1+K] Montgomery form of [-Q_d]_{p_j}
  uint64_t* d_md_negp_m = nullptr;       // [L1] Montgomery form of [-P]_{q_i}
  bool mac30_ok() const { return split_mac_ok && (!reduced_noise || (alpha < 16 && K < 16)); }
  // ModDown
  uint64_t* d_md_phinv = nullptr;        // [K]     [P^_k^-1]_{p_k}
  uint64_t* d_md_phinv_sh = nullptr;     // [K]
  uint64_t* d_md_phat = nullptr;         // [K][L1] [P^_k]_{q_i}
  uint64_t* d_md_pinv = nullptr;         // [L1]    [P^-1]_{q_i}
  uint64_t* d_md_pinv_sh = nullptr;
  uint64_t* d_pmod = nullptr;            // [L1]    [P]_{q_i}
  uint64_t* d_pmod_sh = nullptr;
  // Rescale
  uint64_t* d_rs_qlinv = nullptr;        // [L1][L1] [q_l^-1]_{q_i}
  uint64_t* d_rs_qlinv_sh = nullptr;
  uint64_t* d_rs_QlQl = nullptr;         // [L1][L1] QlQlInvModqlDivqlModq
  uint64_t* d_rs_QlQl_sh = nullptr;

  // device encoder (built on first use): zeta^idx = (cos, sin)(2 pi idx / 2N) for idx < 2N, and
  // 5^j mod 2N for j < N/2
  double* d_ifft_w = nullptr;
  uint32_t* d_ifft_rot = nullptr;

  // host copies of the conversion constants (for table checks and host helpers)
  std::vector<uint64_t> h_pmod, h_pinv;

  // team NTT launches: arrival counters [kMaxTeams][3], the XCD of every member [kMaxTeams][16], an
  // error word (a barrier wait ran past its bound) and the count of teams that spanned XCDs,
  // zeroed at creation
  static constexpr int kMaxTeams = 1024;
  uint32_t* d_team = nullptr;
  uint32_t* team_ctr() const { return d_team; }
  uint32_t* team_xcc() const { return d_team + 3 * kMaxTeams; }
  uint32_t* team_err() const { return d_team + 19 * kMaxTeams; }
  int cu_count = 0;

  hipStream_t stream = nullptr;
  UploadRing uploads;  // staging of job-table uploads on `stream`
  // second stream (default priority) and the fork / join events of two-stream work
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void ensure_aux_stream();
  Pool pool;
  mutable Profiler prof;
  Tune tune;
  std::vector<void*> owned;   // table allocations freed at destruction

  Context(int logN, int L1, int K, int dnum, const std::vector<uint64_t>& primes);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int num_digits(int l) const {
    int b = (l + 1 + alpha - 1) / alpha;
    return b > dnum ? dnum : b;
  }
  int ext_prime(int l, int pos) const { return pos <= l ? pos : L1 + (pos - l - 1); }
  size_t limb_words() const { return N; }

 private:
  template <class T>
  T* upload(const std::vector<T>& v);
};

// A hook run at the start of every engine call this thread makes on a context (before the call
// issues any work): the serving loop's bootstrap worker sets it to issue a bootstrap one
// operation at a time, in the gaps between the loop's ticks. Unset (fn null) by default.
struct CallGate {
  void (*fn)(void* user, Context& c) = nullptr;
  void* user = nullptr;
};
CallGate& thread_call_gate();

}  // namespace fh
