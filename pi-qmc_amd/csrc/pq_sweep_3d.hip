// pq_sweep_3d.hip -- sweep_kernel instantiations for D=3, double arithmetic.
#include "pq_sweep.h"

namespace pq {
hipError_t launch_sweep_3d(int NC, bool multi, const DevCtx& x, int nchain, int nsweep,
                           const SweepLayoutH& L, hipStream_t st) {
  return launch_sweep_dr<3, double>(NC, multi, x, nchain, nsweep, L, st);
}
}  // namespace pq
