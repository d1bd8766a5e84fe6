// NTT / INTT entry points over device job tables (mixed-level batches), with the ModDown and
// Rescale job-table epilogues. The kernels are in fh_ntt_impl.hpp.
#include "fh_ntt_impl.hpp"

namespace fh {

void ntt_fwd_jobs(const Context& c, const LimbJob* j, uint32_t n, hipStream_t st) { dispatch(c, JobAddr{j}, JobAddr{j}, n, st, true); }
void ntt_inv_jobs(const Context& c, const LimbJob* j, uint32_t n, hipStream_t st) { dispatch(c, JobAddr{j}, JobAddr{j}, n, st, false); }
void ntt_fwd_jobs(const Context& c, const LimbJob* in, const LimbJob* out, uint32_t n, hipStream_t st) {
  dispatch(c, JobAddr{in}, JobAddr{out}, n, st, true);
}
void ntt_inv_jobs(const Context& c, const LimbJob* in, const LimbJob* out, uint32_t n, hipStream_t st) {
  dispatch(c, JobAddr{in}, JobAddr{out}, n, st, false);
}
void ntt_fwd_jobs_moddown(const Context& c, const LimbJob* jobs, const EpiJob* epi, uint32_t n, hipStream_t st) {
  EpiModDownJobs e{epi, c.d_md_pinv, c.d_md_pinv_sh, nullptr, nullptr, nullptr, 0, 0, 0};
  dispatch(c, JobAddr{jobs}, JobAddr{jobs}, n, st, true, e, "ntt_fwd_rows_moddown");
}
void ntt_fwd_jobs_rescale(const Context& c, const LimbJob* jobs, const EpiJob* epi, uint32_t n, hipStream_t st) {
  EpiRescaleJobs e{epi, c.d_rs_qlinv, c.d_rs_qlinv_sh, (uint32_t)c.L1, nullptr, nullptr, nullptr, 0, 0, PrimeConst{}};
  dispatch(c, JobAddr{jobs}, JobAddr{jobs}, n, st, true, e, "ntt_fwd_rows_rescale");
}

}  // namespace fh
