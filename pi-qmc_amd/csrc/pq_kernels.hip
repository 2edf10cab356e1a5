// pq_kernels.hip -- gfx950 kernels of the bisection-move hot path.
//
// sweep_kernel: the device-resident sampler.  One single-wave workgroup per
// chain runs nsweep sweeps; a sweep is nsection x (choose the section, copy it
// from HBM to LDS, initialize the node caches, run every Sample's tryMove over
// levels L..0, write the moved slices back).  Chains never communicate, so the
// launch needs no grid synchronisation.
//
// The call-level kernels (action_diff_kernel, fn_*_kernel, node_eval_kernel)
// run the same Eng routines on caller-provided sections in global memory; the
// estimator kernels compute the thermodynamic energy and g(r) counts of the
// resident paths.
#include "pq_engine.h"
#include "pq_launch.h"

namespace pq {

// LDS carve-up of the sampler workgroup
struct SweepLayout {
  int off_sec1, off_sec2, off_mov, off_fst, off_wdet, off_werr, off_wscale, off_scale, off_notmine;
  int off_red, off_cnt, off_inv;   // off_inv < 0: inverse cache in global memory
  int off_gj;
  int bytes;
};

__device__ __forceinline__ void copy_chain(ChainP* dst, const ChainP* src) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (int k = threadIdx.x; k < (int)(sizeof(ChainP) / 4); k += blockDim.x) d[k] = s[k];
  __syncthreads();
}

template <int D, int G, class R, int NT>
__global__ __launch_bounds__(NT) void sweep_kernel(DevCtx x, int nsweep, SweepLayout lay) {
  extern __shared__ double smem[];
  const int chain = x.order[blockIdx.x];
  ChainP& c = *reinterpret_cast<ChainP*>(smem);
  copy_chain(&c, &x.chains[chain]);
  char* base = reinterpret_cast<char*>(smem);
  Eng<D, G, R, NT> e(x, c, chain);
  e.sec1 = reinterpret_cast<double*>(base + lay.off_sec1);
  e.sec2 = reinterpret_cast<double*>(base + lay.off_sec2);
  e.mov = reinterpret_cast<double*>(base + lay.off_mov);
  e.fst = reinterpret_cast<double*>(base + lay.off_fst);
  e.wdet = reinterpret_cast<double*>(base + lay.off_wdet);
  e.werr = reinterpret_cast<int*>(base + lay.off_werr);
  e.wscale = reinterpret_cast<double*>(base + lay.off_wscale);
  e.scale = reinterpret_cast<double*>(base + lay.off_scale);
  e.notmine = reinterpret_cast<int*>(base + lay.off_notmine);
  e.red = reinterpret_cast<double*>(base + lay.off_red);
  e.gjrow = reinterpret_cast<double*>(base + lay.off_gj);
  e.inv = (lay.off_inv >= 0) ? reinterpret_cast<double*>(base + lay.off_inv) : x.inv + c.inv_off;
  e.egr = x.egr + c.egr_off;
  // per-level try/accept counters, bead updates, accepted, node errors
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base + lay.off_cnt);
  const int tid = e.tid;
  if (tid < 19) cnt[tid] = 0;
  if (tid < c.nfn) e.scale[tid] = x.scale[(int64_t)chain * PQ_MAX_NODES + tid];
  __syncthreads();

  const int S = c.S, N = c.N, M = c.M, L = c.L;
  const int64_t slice = (int64_t)N * D;
  double* beads = x.beads + c.bead_off;
  uint64_t k = x.counter[chain];

  for (int sw = 0; sw < nsweep; ++sw) {
    for (int isec = 0; isec < c.nsection; ++isec) {
      // SectionChooser::run (SectionChooser.cc:30-48) / DoubleSectionChooser::run (:33-65)
      int t0, t0b = 0;
      if (c.nfn) {
        double xx = e.U(k++) * (1 - 1e-8);
        const int ilo = -1, ihi = M / 2 - 1;
        t0 = ilo + (int)((ihi + 1 - ilo) * xx);
        if (t0 > ihi) t0 = ihi;
        if (e.U(k++) >= 0.5) t0 = (t0 + M / 2 + 1) % M - 1;
        t0b = (t0 + M / 2 + 1) % M - 1;
      } else {
        double xx = e.U(k++) * (1 - 1e-8);
        const int ilo = -1, ihi = M - 1;
        t0 = ilo + (int)((ihi + 1 - ilo) * xx);
        if (t0 > ihi) t0 = ihi;
      }
      // SerialPaths::getBeads (SerialPaths.cc:119-132), identity permutation
      for (int64_t q = tid; q < S * slice; q += NT) {
        int kk = (int)(q / slice);
        int64_t rem = q - kk * slice;
        int t = kk + t0;
        t = (t < 0) ? t + M : ((t >= M) ? t - M : t);
        e.sec1[q] = beads[t * slice + rem];
        if (c.nfn) {
          int tb = kk + t0b;
          tb = (tb < 0) ? tb + M : ((tb >= M) ? tb - M : tb);
          e.sec2[q] = beads[tb * slice + rem];
        }
      }
      __syncthreads();
      if (c.nfn) {
        const int ne = e.fn_initialize();
        if (tid == 0) cnt[18] += ne;
      }

      for (int smp = 0; smp < c.nsample; ++smp) {
        const int sp = c.sample_sp[smp];
        // SpeciesParticleChooser::chooseParticles (SpeciesParticleChooser.cc:19-31), npart=1
        int i = (int)(c.count[sp] * e.U(k++) + c.ifirst[sp]);
        if (i == c.count[sp] + c.ifirst[sp]) i = c.count[sp] + c.ifirst[sp] - 1;
        if (tid < D) {
          e.mov[tid] = e.sb(e.sec1, 0, i)[tid];
          e.mov[(S - 1) * D + tid] = e.sb(e.sec1, S - 1, i)[tid];
        }
        __syncthreads();
        // MultiLevelSampler::tryMove (MultiLevelSampler.cc:100-177)
        double old = 0.0;
        bool accepted = true;
        for (int l = L; l >= 0; --l) {
          if (tid == 0) {
            cnt[l]++;
            if (l < L) cnt[16] += 1ull << (L - 1 - l);
          }
          int nd;
          const double lnT = e.mover(l, i, k, &nd);
          k += nd;
          __syncthreads();
          double dA = e.action_diff(i, l);
          if (c.nfn) dA += e.fn_diff(i, l);
          const double piRatio = -dA + old;
          const double acceptProb = exp(lnT + piRatio);
          if (e.U(k++) > acceptProb) { accepted = false; break; }
          old = dA;
          if (tid == 0) cnt[8 + l]++;
        }
        if (accepted) {
          if (c.nfn) e.fn_accept();
          for (int q = tid; q < S * D; q += NT) {
            const int t = q / D, d = q - t * D;
            e.sb(e.sec1, t, i)[d] = e.mov[q];
          }
          if (tid == 0) cnt[17] += (1ull << L) - 1;
          __syncthreads();
        }
      }
      // SerialPaths::putBeads (SerialPaths.cc:144-165): interior slices only,
      // the end slices are never moved
      for (int64_t q = slice + tid; q < (S - 1) * slice; q += NT) {
        int kk = (int)(q / slice);
        int64_t rem = q - kk * slice;
        int t = kk + t0;
        t = (t >= M) ? t - M : t;
        beads[t * slice + rem] = e.sec1[q];
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (tid == 0) {
    x.counter[chain] = k;
    pq_chain_stats& st = x.stats[chain];
    st.draws = k;
    st.bead_updates += cnt[16];
    st.accepted_bead_updates += cnt[17];
    st.node_errors += cnt[18];
    for (int l = 0; l < 8; ++l) { st.tries[l] += cnt[l]; st.accepts[l] += cnt[8 + l]; }
  }
  if (tid < c.nfn) x.scale[(int64_t)chain * PQ_MAX_NODES + tid] = e.scale[tid];
}

// ---------------------------------------------------- call-level kernels

struct CallBufs {
  const int* qchain;
  const int* qlevel;
  const double* qsec;      // [nq][S][N][D] (padded to the largest S*N*D)
  const double* qmov;      // [nq][S][D]
  const int* qidx;
  int64_t sec_stride;
  int64_t mov_stride;
  double* out;
};

template <int D, int G, class R>
__global__ __launch_bounds__(PQ_WAVE) void action_diff_kernel(DevCtx x, CallBufs b) {
  __shared__ ChainP c;
  const int q = blockIdx.x;
  const int chain = b.qchain[q];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, G, R, PQ_WAVE> e(x, c, chain);
  e.sec1 = const_cast<double*>(b.qsec + q * b.sec_stride);
  e.mov = const_cast<double*>(b.qmov + q * b.mov_stride);
  double v = e.action_diff(b.qidx[q], b.qlevel[q]);
  if (e.lane == 0) b.out[q] = v;
}

// The fixed-node protocol of one chain on global buffers:
// op 0 initialize, op 1 action difference at level, op 2 accept.
struct FnBufs {
  double* sec1;
  double* sec2;
  double* mov;
  double* wdet;
  int* werr;
  double* wscale;
  int* notmine;
  int op;
  int level;
  int idx;
  double* out;
};

template <int D, int G, class R>
__global__ __launch_bounds__(PQ_WAVE) void fn_kernel(DevCtx x, int chain, FnBufs b) {
  __shared__ ChainP c;
  __shared__ double sscale[PQ_MAX_NODES];
  __shared__ double gj[2 * PQ_WAVE];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, G, R, PQ_WAVE> e(x, c, chain);
  e.gjrow = gj;
  e.sec1 = b.sec1; e.sec2 = b.sec2; e.mov = b.mov;
  e.fst = x.fnstate + c.fn_off;
  e.inv = x.inv + c.inv_off;
  e.egr = x.egr + c.egr_off;
  e.wdet = b.wdet; e.werr = b.werr; e.wscale = b.wscale; e.notmine = b.notmine;
  e.scale = sscale;
  if (e.lane < c.nfn) sscale[e.lane] = x.scale[(int64_t)chain * PQ_MAX_NODES + e.lane];
  __syncthreads();
  if (b.op == 0) {
    int nerr = e.fn_initialize();
    if (e.lane == 0) b.out[0] = (double)nerr;
  } else if (b.op == 1) {
    double v = e.fn_diff(b.idx, b.level);
    if (e.lane == 0) b.out[0] = v;
  } else {
    e.fn_accept();
  }
  __syncthreads();
  if (e.lane < c.nfn) x.scale[(int64_t)chain * PQ_MAX_NODES + e.lane] = sscale[e.lane];
}

// NodeModel::evaluate + evaluateDistance on nmat (r1, r2) slice pairs, staged
// as a pseudo-section of nmat slices.
template <int D, int G, class R>
__global__ __launch_bounds__(PQ_WAVE) void node_eval_kernel(DevCtx x, int chain, int fa, int nmat,
                                                             double* r1, double* r2, double* inv,
                                                             double* egr, double* det, int* err,
                                                             double* dist) {
  __shared__ ChainP c;
  __shared__ double wd[PQ_WAVE];
  __shared__ int we[PQ_WAVE];
  __shared__ double gj[2 * PQ_WAVE];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, G, R, PQ_WAVE> e(x, c, chain);
  e.gjrow = gj;
  e.sec1 = r1; e.sec2 = r2;
  e.inv = inv;
  e.egr = egr;
  e.wdet = wd; e.werr = we;
  const double scl = x.scale[(int64_t)chain * PQ_MAX_NODES + fa];
  constexpr int NG = PQ_WAVE / G;
  for (int t = 0; t < nmat; t += NG) {
    const int nt = min(NG, nmat - t);
    e.node_matrices(fa, t, 1, nt, false, -1, scl);
    if (e.lane < nt) { det[t + e.lane] = wd[e.lane]; err[t + e.lane] = we[e.lane]; }
    __syncthreads();
  }
  e.node_distances(fa, 0, nmat, false, -1, nullptr, scl, dist);
}

// ------------------------------------------------------------ estimators

// getBeadAction utau of spring, SHO and pair actions for bead (i,t) of a
// chain, summed as CompositeAction::getBeadAction (CompositeAction.cc:57-67).
template <int D>
__device__ double bead_utau(const DevCtx& x, const ChainP& c, const double* beads, int i, int t) {
  auto P = [&](int p, int tt) -> const double* {
    tt %= c.M;
    if (tt < 0) tt += c.M;
    return beads + ((int64_t)tt * c.N + p) * D;
  };
  const int sp = species_of(c, i);
  double utau = 0;
  const double* b1 = P(i, t);
  const double* b0 = P(i, t - 1);
  if (c.spring) {
    double delta[D];
#pragma unroll
    for (int d = 0; d < D; ++d) delta[d] = b1[d] - b0[d];
    pbc<D>(c, delta);
    double dd = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) dd += delta[d] * delta[d];
    utau += D / (2 * c.tau) - dd / (4.0 * c.lambda[sp] * c.tau * c.tau);
  }
  for (int a = 0; a < c.nsho; ++a) {
    const int ssp = c.sho_sp[a];
    double ui = 0;
    if (!(i < c.ifirst[ssp] || i >= c.ifirst[ssp] + c.count[ssp])) {
      const int ndim = c.sho_ndim[a] ? c.sho_ndim[a] : D;
      const double omega = c.sho_omega[a], m = c.mass[ssp];
      const double wt = omega * c.tau, sinhwt = sinh(wt), coshwt = cosh(wt);
      const double cschwt = 1.0 / sinhwt, cothwt = coshwt * cschwt, tin = 1.0 / c.tau;
      double delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = b1[d] - b0[d];
      pbc<D>(c, delta);
      double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
      for (int d = D - ndim; d < D; ++d) {
        double r1 = b1[d] - c.sho_center[a][d], r0 = b0[d] - c.sho_center[a][d];
        r0r0 += r0 * r0; r0r1 += r0 * r1; r1r1 += r1 * r1; delta2 += delta[d] * delta[d];
      }
      ui = 0.5 * ndim * omega * cothwt
           + 0.5 * m * omega * omega * ((r1r1 + r0r0) * (1 - cothwt * cothwt) + 2.0 * r0r1 * coshwt * cschwt * cschwt)
           + 0.5 * m * delta2 * tin * tin - 0.5 * ndim * tin;
    }
    utau += ui;
  }
  double uc = 0;
  for (int pp = 0; pp < c.npair; ++pp) {
    const PairP& PP = c.pair[pp];
    int js;
    if (sp == PP.s1) js = PP.s2;
    else if (sp == PP.s2) js = PP.s1;
    else continue;
    const TableP T = x.tables[PP.table];
    const double* u = x.tab_u + T.off;
    const double* ut = x.tab_ut + T.off;
    double up = 0;
    for (int j = c.ifirst[js]; j < c.ifirst[js] + c.count[js]; ++j) {
      if (j == i) continue;
      const double* a1 = P(j, t);
      const double* a0 = P(j, t - 1);
      double dl[D], pd[D];
#pragma unroll
      for (int d = 0; d < D; ++d) { dl[d] = b1[d] - a1[d]; pd[d] = b0[d] - a0[d]; }
      pbc<D>(c, dl);
      pbc<D>(c, pd);
      double r = 0, pr = 0, s2 = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) { r += dl[d] * dl[d]; pr += pd[d] * pd[d]; }
      r = sqrt(r);
      pr = sqrt(pr);
      const double q = 0.5 * (r + pr);
#pragma unroll
      for (int d = 0; d < D; ++d) { double sv = dl[d] - pd[d]; s2 += sv * sv; }
      s2 /= q * q;
      // uk0CalcDerivatives, utau part (PairAction.cc:640-657)
      double qq = q * T.rgridinv;
      double xx = log(qq) * T.logrratioinv;
      int ig = (int)xx;
      if (qq < 1) { ig = 0; xx = 0; }
      else if (ig > T.ngpts - 2) { ig = T.ngpts - 2; xx = 1; }
      else xx -= ig;
      double vt = 0;
      for (int kk = T.K1 - 1; kk >= 0; --kk) {
        vt *= s2;
        vt += (1 - xx) * ut[ig * T.K1 + kk] + xx * ut[(ig + 1) * T.K1 + kk];
      }
      (void)u;
      up += 0.5 * vt;
    }
    uc += up;
  }
  utau += uc;
  return utau;
}

// ThermoEnergyEstimator (ThermoEnergyEstimator.cc:32-54): one workgroup per chain
template <int D>
__global__ __launch_bounds__(256) void energy_kernel(DevCtx x, double* energy) {
  __shared__ ChainP c;
  __shared__ double part[4];
  const int chain = blockIdx.x;
  {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&x.chains[chain]);
    uint32_t* d = reinterpret_cast<uint32_t*>(&c);
    for (int k = threadIdx.x; k < (int)(sizeof(ChainP) / 4); k += blockDim.x) d[k] = s[k];
    __syncthreads();
  }
  const double* beads = x.beads + c.bead_off;
  double acc = 0;
  const int nl = c.M * c.N;
  for (int q = threadIdx.x; q < nl; q += blockDim.x) acc += bead_utau<D>(x, c, beads, q % c.N, q / c.N);
  acc = wave_sum(acc);
  if (lane_id() == 0) part[threadIdx.x / PQ_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x / PQ_WAVE); ++w) v += part[w];
    energy[chain] = v / c.M;
  }
}

template <int D>
__global__ __launch_bounds__(256) void bead_utau_kernel(DevCtx x, int nq, const int* qc, const int* qi,
                                                         const int* qt, double* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const ChainP& c = x.chains[qc[q]];
  out[q] = bead_utau<D>(x, c, x.beads + c.bead_off, qi[q], qt[q]);
}

// PairCFEstimator<1>::handleLink with PairRadial (PairCFEstimator.h:72-93).
// Integer counts per chain; contraction is off so the bin of every distance
// is the one the FP64 expression gives on the host.
template <int D>
__global__ __launch_bounds__(256) void histogram_kernel(DevCtx x, int s1, int s2, double rmin,
                                                         double rmax, int nbin, uint64_t* counts) {
#pragma clang fp contract(off)
  extern __shared__ unsigned int hist[];
  __shared__ ChainP c;
  const int chain = blockIdx.x;
  {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&x.chains[chain]);
    uint32_t* d = reinterpret_cast<uint32_t*>(&c);
    for (int k = threadIdx.x; k < (int)(sizeof(ChainP) / 4); k += blockDim.x) d[k] = s[k];
  }
  for (int b = threadIdx.x; b < nbin; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  if (s1 < c.nsp && s2 < c.nsp) {
    const double* beads = x.beads + c.bead_off;
    const double deltaInv = nbin / (rmax - rmin);
    const int n1 = c.count[s1], n2 = c.count[s2];
    const int64_t items = (int64_t)c.M * n1 * n2;
    for (int64_t q = threadIdx.x; q < items; q += blockDim.x) {
      const int t = (int)(q / ((int64_t)n1 * n2));
      const int rem = (int)(q - (int64_t)t * n1 * n2);
      const int i = c.ifirst[s1] + rem / n2, j = c.ifirst[s2] + rem % n2;
      if (i == j) continue;
      const double* a = beads + ((int64_t)t * c.N + i) * D;
      const double* b = beads + ((int64_t)t * c.N + j) * D;
      double delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = a[d] - b[d];
      pbc<D>(c, delta);
      double r2 = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) r2 += delta[d] * delta[d] * 1.0;
      const double dd = sqrt(r2);
      const int ibin = (int)((dd - rmin) * deltaInv);
      if (dd < rmin || ibin > nbin - 1) continue;
      atomicAdd(&hist[ibin], 1u);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbin; b += blockDim.x) counts[(int64_t)chain * nbin + b] += hist[b];
}

// ------------------------------------------------------------ launchers

template <int D, int G, class R, int NT>
static hipError_t launch_sweep_nt(const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                  hipStream_t st) {
  SweepLayout lay;
  lay.off_sec1 = L.off_sec1; lay.off_sec2 = L.off_sec2; lay.off_mov = L.off_mov;
  lay.off_fst = L.off_fst; lay.off_wdet = L.off_wdet; lay.off_werr = L.off_werr;
  lay.off_wscale = L.off_wscale; lay.off_scale = L.off_scale; lay.off_notmine = L.off_notmine;
  lay.off_red = L.off_red; lay.off_cnt = L.off_cnt; lay.off_inv = L.off_inv;
  lay.off_gj = L.off_gj;
  lay.bytes = L.bytes;
  auto k = sweep_kernel<D, G, R, NT>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, L.bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(nchain), dim3(NT), L.bytes, st, x, nsweep, lay);
  return hipGetLastError();
}

template <int D, int G, class R>
static hipError_t launch_sweep_t(const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                 hipStream_t st) {
  if (L.team == 256) return launch_sweep_nt<D, G, R, 256>(x, nchain, nsweep, L, st);
  return launch_sweep_nt<D, G, R, PQ_WAVE>(x, nchain, nsweep, L, st);
}

#define PQ_DISPATCH(FN, ...)                                                          \
  do {                                                                                \
    if (mixed) {                                                                      \
      if (D == 2) { if (G <= 4) return FN<2, 4, float>(__VA_ARGS__);                  \
                    if (G <= 16) return FN<2, 16, float>(__VA_ARGS__);                \
                    return FN<2, 32, float>(__VA_ARGS__); }                           \
      if (G <= 4) return FN<3, 4, float>(__VA_ARGS__);                                \
      if (G <= 16) return FN<3, 16, float>(__VA_ARGS__);                              \
      return FN<3, 32, float>(__VA_ARGS__);                                           \
    }                                                                                 \
    if (D == 2) { if (G <= 4) return FN<2, 4, double>(__VA_ARGS__);                   \
                  if (G <= 16) return FN<2, 16, double>(__VA_ARGS__);                 \
                  return FN<2, 32, double>(__VA_ARGS__); }                            \
    if (G <= 4) return FN<3, 4, double>(__VA_ARGS__);                                 \
    if (G <= 16) return FN<3, 16, double>(__VA_ARGS__);                               \
    return FN<3, 32, double>(__VA_ARGS__);                                            \
  } while (0)

hipError_t launch_sweep(int D, int G, bool mixed, const DevCtx& x, int nchain, int nsweep,
                        const SweepLayoutH& L, hipStream_t st) {
  PQ_DISPATCH(launch_sweep_t, x, nchain, nsweep, L, st);
}

template <int D, int G, class R>
static hipError_t launch_action_diff_t(const DevCtx& x, int nq, const CallBufsH& b, hipStream_t st) {
  CallBufs cb{b.qchain, b.qlevel, b.qsec, b.qmov, b.qidx, b.sec_stride, b.mov_stride, b.out};
  hipLaunchKernelGGL((action_diff_kernel<D, G, R>), dim3(nq), dim3(PQ_WAVE), 0, st, x, cb);
  return hipGetLastError();
}

hipError_t launch_action_diff(int D, int G, bool mixed, const DevCtx& x, int nq, const CallBufsH& b,
                              hipStream_t st) {
  PQ_DISPATCH(launch_action_diff_t, x, nq, b, st);
}

template <int D, int G, class R>
static hipError_t launch_fn_t(const DevCtx& x, int chain, const FnBufsH& b, hipStream_t st) {
  FnBufs fb{b.sec1, b.sec2, b.mov, b.wdet, b.werr, b.wscale, b.notmine, b.op, b.level, b.idx, b.out};
  hipLaunchKernelGGL((fn_kernel<D, G, R>), dim3(1), dim3(PQ_WAVE), 0, st, x, chain, fb);
  return hipGetLastError();
}

hipError_t launch_fn(int D, int G, bool mixed, const DevCtx& x, int chain, const FnBufsH& b,
                     hipStream_t st) {
  PQ_DISPATCH(launch_fn_t, x, chain, b, st);
}

template <int D, int G, class R>
static hipError_t launch_node_eval_t(const DevCtx& x, int chain, int fa, int nmat, double* r1,
                                     double* r2, double* inv, double* egr, double* det, int* err,
                                     double* dist, hipStream_t st) {
  hipLaunchKernelGGL((node_eval_kernel<D, G, R>), dim3(1), dim3(PQ_WAVE), 0, st, x, chain, fa, nmat,
                     r1, r2, inv, egr, det, err, dist);
  return hipGetLastError();
}

hipError_t launch_node_eval(int D, int G, bool mixed, const DevCtx& x, int chain, int fa, int nmat,
                            double* r1, double* r2, double* inv, double* egr, double* det, int* err,
                            double* dist, hipStream_t st) {
  PQ_DISPATCH(launch_node_eval_t, x, chain, fa, nmat, r1, r2, inv, egr, det, err, dist, st);
}

hipError_t launch_energy(int D, const DevCtx& x, int nchain, double* energy, hipStream_t st) {
  if (D == 2) hipLaunchKernelGGL(energy_kernel<2>, dim3(nchain), dim3(256), 0, st, x, energy);
  else hipLaunchKernelGGL(energy_kernel<3>, dim3(nchain), dim3(256), 0, st, x, energy);
  return hipGetLastError();
}

hipError_t launch_bead_utau(int D, const DevCtx& x, int nq, const int* qc, const int* qi, const int* qt,
                            double* out, hipStream_t st) {
  const int nb = (nq + 255) / 256;
  if (D == 2) hipLaunchKernelGGL(bead_utau_kernel<2>, dim3(nb), dim3(256), 0, st, x, nq, qc, qi, qt, out);
  else hipLaunchKernelGGL(bead_utau_kernel<3>, dim3(nb), dim3(256), 0, st, x, nq, qc, qi, qt, out);
  return hipGetLastError();
}

hipError_t launch_histogram(int D, const DevCtx& x, int nchain, int s1, int s2, double rmin, double rmax,
                            int nbin, uint64_t* counts, hipStream_t st) {
  const size_t lds = sizeof(unsigned int) * nbin;
  if (D == 2) hipLaunchKernelGGL(histogram_kernel<2>, dim3(nchain), dim3(256), lds, st, x, s1, s2, rmin, rmax, nbin, counts);
  else hipLaunchKernelGGL(histogram_kernel<3>, dim3(nchain), dim3(256), lds, st, x, s1, s2, rmin, rmax, nbin, counts);
  return hipGetLastError();
}

}  // namespace pq
