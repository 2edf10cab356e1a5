// pq_launch.h -- host-side launch entry points of the kernels (pq_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "pq_device.h"

namespace pq {

struct SweepLayoutH {
  int off_sec1, off_sec2, off_mov, off_fst, off_wdet, off_werr, off_wscale, off_scale, off_notmine;
  int off_red, off_cnt, off_inv;   // off_inv < 0: inverse cache in global memory
  int off_gj;                      // Gauss-Jordan pivot rows, 2 doubles per thread
  int off_tab;                     // pair tables (u) staged in LDS, < 0: read from global memory
  int off_tabofs;                  // [PQ_MAX_PAIRS] LDS offset of each pair's table
  int off_choose;                  // chooser results (Eng::ChooseBuf)
  int off_perm;                    // path permutation, its inverse, section permutation, scratch: [5][N] ints
  int off_walk;                    // WalkingChooser rows, < 0: no walking Sample
  int bytes;
  int team;                        // threads per chain: 64 to 512
  int* queue;                      // queue counter (zeroed before the launch), or null: one chain per workgroup
  int nwork;                       // chains of the launch (set by the launcher)
};

struct CallBufsH {
  const int* qchain;
  const int* qlevel;
  const double* qsec;
  const double* qmov;
  const int* qidx;
  int nmov;
  int64_t sec_stride;
  int64_t mov_stride;
  double* out;
};

struct FnBufsH {
  double* sec1;
  double* sec2;
  double* mov;
  double* wdet;
  int* werr;
  double* wscale;
  int* notmine;
  int op;
  int level;
  const int* idx;
  int nmov;
  double* out;
};

// A Measure work item: slices [t0, t1) of one chain.
struct MeasItem {
  int chain, t0, t1;
};

// Stride in doubles of the energy items' staged slices: odd, so that lanes
// reading one particle on consecutive slices fall on different LDS banks.
inline __host__ __device__ int meas_slice_stride(int nd) { return nd | 1; }


#define PQ_MAX_CF_SPECIES 8   // species1 and up to seven partner species of a pair-correlation histogram

// PairCFEstimator<1> with PairRadial: sp[0] against partners sp[1..nsp-1]
struct PairCfArgs {
  int nsp;
  int sp[PQ_MAX_CF_SPECIES];
  int nbin;
  double rmin, rmax, deltaInv;
};

hipError_t launch_energy_items(int D, const DevCtx& x, const MeasItem* items, int nitems, const int* first_item,
                               int nchain, int max_slice_doubles, double* part, double* energy, hipStream_t st);
hipError_t launch_pair_cf_items(int D, const DevCtx& x, const MeasItem* items, int nitems, int max_slice_doubles,
                                const PairCfArgs& g, uint64_t* counts, hipStream_t st);
hipError_t launch_fill_u64(uint64_t* p, int64_t n, hipStream_t st);

hipError_t launch_sweep(int D, int NC, bool mixed, bool multi, const DevCtx& x, int nchain, int nsweep,
                        const SweepLayoutH& L, hipStream_t st);
hipError_t launch_action_diff(int D, int NC, bool mixed, const DevCtx& x, int nq, const CallBufsH& b,
                              hipStream_t st);
hipError_t launch_fn(int D, int NC, bool mixed, const DevCtx& x, int chain, const FnBufsH& b,
                     hipStream_t st);
hipError_t launch_node_eval(int D, int NC, bool mixed, const DevCtx& x, int chain, int fa, int nmat,
                            double* r1, double* r2, double* inv, double* egr, double* det, int* err,
                            double* dist, hipStream_t st);
hipError_t launch_coulomb_energy(int D, const DevCtx& x, int nchain, double eps, double* energy, hipStream_t st);
hipError_t launch_virial_energy(int D, const DevCtx& x, int nchain, int nwindow, double* energy, hipStream_t st);
hipError_t launch_ewald_coulomb_energy(int D, const DevCtx& x, int nchain, const EwP* ew, const double* kpool,
                                       const int64_t* voff, const double* vgrid, double eps, double* energy,
                                       hipStream_t st);
hipError_t launch_bead_action(int D, const DevCtx& x, int chain, const double* beads, const int* perm, int nq,
                              const int* qi, const int* qt, double* out, hipStream_t st);

}  // namespace pq
