// Forward NTT of the new limbs of a ModUp (the ModUpAddr job space). The kernels are in
// fh_ntt_impl.hpp.
#include "fh_ntt_impl.hpp"

namespace fh {

void ntt_fwd_modup(const Context& c, const ModUpAddr& a, uint32_t n, hipStream_t st) {
  ModUpAddr m = a;
  m.mix = c.tune.ntt_mix ? 1u : 0u;
  dispatch(c, m, m, n, st, true);
}

}  // namespace fh
