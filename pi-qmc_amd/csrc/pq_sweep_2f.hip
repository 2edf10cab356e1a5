// pq_sweep_2f.hip -- sweep_kernel instantiations for D=2, float arithmetic.
#include "pq_sweep.h"

namespace pq {
hipError_t launch_sweep_2f(int NC, bool multi, const DevCtx& x, int nchain, int nsweep,
                           const SweepLayoutH& L, hipStream_t st) {
  return launch_sweep_dr<2, float>(NC, multi, x, nchain, nsweep, L, st);
}
}  // namespace pq
