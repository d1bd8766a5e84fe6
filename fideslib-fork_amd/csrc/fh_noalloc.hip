// Allocation-free Rescale: the work buffer of fh_rescale lives in a fixed device array instead of
// the context's pool, for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// `scratch` is one file-scope device array per GPU (each device that loads this library holds its
// own copy until the process ends). In host code its name is a host-side stand-in, never a device
// address: the device address is looked up once per device with hipGetSymbolAddress (noalloc_init,
// called by fh_context_create with the context's device current) and only that address is handed
// to kernels. Every call on one device shares that device's array, so calls that overlap on one
// GPU (several streams or threads at once) race; calls issued in order on one stream do not.
#include <mutex>

#include "fh_keyswitch.hpp"

namespace fh {

namespace {

// 2^22 words = 32 MiB: Rescale of ncomp components at level l needs ncomp (l + 1) N words, i.e.
// up to 64 limb-rows at N = 2^16 (two components up to level 31).
constexpr size_t kScratchWords = size_t(1) << 22;
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_ptr[kMaxDevices] = {};   // device address of `scratch` per device (null until looked up)

}  // namespace

size_t noalloc_scratch_words() { return kScratchWords; }

void noalloc_init(Context& c) {
  int dev = 0;
  FH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(-5, "device ordinal above the scratch table");
  c.device = dev;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) throw Error(-5, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_ptr[dev] = static_cast<uint64_t*>(p);
}

void rescale_noalloc(Context& c, const uint64_t* x, int l, int ncomp, uint64_t* out, hipStream_t st) {
  const size_t need = rescale_work_words(c, l, ncomp);
  if (need > kScratchWords)
    throw Error(-1, "rescale_noalloc needs " + std::to_string(need * 8) + " bytes of scratch, above the fixed " +
                        std::to_string(kScratchWords * 8) + "-byte array (ncomp * (level + 1) * N <= 2^22 words)");
  uint64_t* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c.device >= 0 && c.device < kMaxDevices) p = g_ptr[c.device];
  }
  if (!p) throw Error(-5, "scratch array not set up for this context's device");
  rescale_core(c, x, l, ncomp, out, p, st);
}

}  // namespace fh
