// pq_sweep_3f.hip -- sweep_kernel instantiations for D=3, float arithmetic.
#include "pq_sweep.h"

namespace pq {
hipError_t launch_sweep_3f(int NC, bool multi, const DevCtx& x, int nchain, int nsweep,
                           const SweepLayoutH& L, hipStream_t st) {
  return launch_sweep_dr<3, float>(NC, multi, x, nchain, nsweep, L, st);
}
}  // namespace pq
