// pq_kernels.hip -- gfx950 kernels of the bisection-move hot path.
//
// sweep_kernel (pq_sweep.h): the device-resident sampler.  One workgroup per
// chain runs nsweep sweeps; a sweep is nsection x (choose the section, copy it
// from HBM to LDS, initialize the node caches, run every Sample's tryMove over
// levels L..0, write the moved slices back).  Chains never communicate, so the
// launch needs no grid synchronisation.
//
// The call-level kernels (action_diff_kernel, fn_*_kernel, node_eval_kernel)
// run the same Eng routines on caller-provided sections in global memory; the
// estimator kernels here compute the Coulomb and virial energies of the
// resident paths (the thermodynamic energy and g(r) are in pq_measure.hip).
#include "pq_engine.h"
#include "pq_launch.h"

namespace pq {

// ---------------------------------------------------- call-level kernels

struct CallBufs {
  const int* qchain;
  const int* qlevel;
  const double* qsec;      // [nq][S][N][D] (padded to the largest S*N*D)
  const double* qmov;      // [nq][S][nmov][D]
  const int* qidx;         // [nq][nmov]
  int nmov;
  int64_t sec_stride;
  int64_t mov_stride;
  double* out;
};

template <int D, int NC, class R>
__global__ __launch_bounds__(PQ_WAVE) void action_diff_kernel(DevCtx x, CallBufs b) {
  __shared__ ChainP c;
  const int q = blockIdx.x;
  const int chain = b.qchain[q];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, NC, R, PQ_WAVE> e(x, c, chain);
  e.sec1 = const_cast<double*>(b.qsec + q * b.sec_stride);
  e.mov = const_cast<double*>(b.qmov + q * b.mov_stride);
  e.nmov_ = b.nmov;
  e.midx = b.qidx + (int64_t)q * b.nmov;
  double v = e.action_diff(b.qlevel[q]);
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
  const int* idx;          // [nmov] moving particles
  int nmov;
  double* out;
};

template <int D, int NC, class R>
__global__ __launch_bounds__(PQ_WAVE) void fn_kernel(DevCtx x, int chain, FnBufs b) {
  __shared__ ChainP c;
  __shared__ double sscale[PQ_MAX_NODES];
  __shared__ double gj[NC == 32 ? PQ_GJ32_SCRATCH_DOUBLES : 2 * PQ_WAVE];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, NC, R, PQ_WAVE> e(x, c, chain);
  e.gjrow = gj;
  e.sec1 = b.sec1; e.sec2 = b.sec2; e.mov = b.mov;
  e.nmov_ = b.nmov; e.midx = b.idx;
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
    double v = e.fn_diff(b.level);
    if (e.lane == 0) b.out[0] = v;
  } else {
    e.fn_accept();
  }
  __syncthreads();
  if (e.lane < c.nfn) x.scale[(int64_t)chain * PQ_MAX_NODES + e.lane] = sscale[e.lane];
}

// NodeModel::evaluate + evaluateDistance on nmat (r1, r2) slice pairs, staged
// as a pseudo-section of nmat slices.
template <int D, int NC, class R>
__global__ __launch_bounds__(PQ_WAVE) void node_eval_kernel(DevCtx x, int chain, int fa, int nmat,
                                                             double* r1, double* r2, double* inv,
                                                             double* egr, double* det, int* err,
                                                             double* dist) {
  __shared__ ChainP c;
  __shared__ double wd[PQ_WAVE];
  __shared__ int we[PQ_WAVE];
  __shared__ double gj[NC == 32 ? PQ_GJ32_SCRATCH_DOUBLES : 2 * PQ_WAVE];
  copy_chain(&c, &x.chains[chain]);
  Eng<D, NC, R, PQ_WAVE> e(x, c, chain);
  e.gjrow = gj;
  e.sec1 = r1; e.sec2 = r2;
  e.inv = inv;
  e.egr = egr;
  e.wdet = wd; e.werr = we;
  const double scl = x.scale[(int64_t)chain * PQ_MAX_NODES + fa];
  constexpr int NG = PQ_WAVE / gj_group(NC);
  for (int t = 0; t < nmat; t += NG) {
    const int nt = min(NG, nmat - t);
    e.node_matrices(fa, t, 1, nt, false, scl);
    if (e.lane < nt) { det[t + e.lane] = wd[e.lane]; err[t + e.lane] = we[e.lane]; }
    __syncthreads();
  }
  e.node_distances(fa, 0, nmat, false, nullptr, scl, dist);
}

// ------------------------------------------------------------ estimators

// sum over slices and half-space k-vectors of 2 f(k)/(2V) |rho_k(t)|^2 plus
// M self energies: the long-range Ewald sums of all M slices, by the threads
// of a workgroup (each returns its share)
template <int D>
__device__ double ewald_all_slices(const double* kp, const EwP& ew, const ChainP& c, const double* beads) {
  double acc = 0;
  for (int q = threadIdx.x; q < c.M * ew.nk; q += blockDim.x) {
    const int t = q / ew.nk;
    const double* kv = kp + 4 * (q - t * ew.nk);
    double rr = 0, ri = 0;
    for (int j = 0; j < c.N; ++j) {
      double co, si;
      ewald_phase<D>(kv, beads + ((int64_t)t * c.N + j) * D, co, si);
      const double qj = c.charge[species_of(c, j)];
      rr += qj * co;
      ri += qj * si;
    }
    acc += kv[3] * (rr * rr + ri * ri);
  }
  if (threadIdx.x == 0) acc += c.M * ew.self;
  return acc;
}

// CoulombEnergyEstimator::handleLink (CoulombEnergyEstimator.cc:30-40) summed
// over every link by SerialPaths::sumOverLinks and divided by M
// (SimpleScalarAccumulator::storeValue): sum over slices t and pairs j < i of
// q_i q_j / (epsilon |pbc(r_i(t) - r_j(t))|).  One workgroup per chain, lanes
// over (slice, particle).
template <int D>
__global__ __launch_bounds__(256) void coulomb_energy_kernel(DevCtx x, double eps, double* energy) {
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
  for (int q = threadIdx.x; q < c.M * c.N; q += blockDim.x) {
    const int t = q / c.N, i = q - t * c.N;
    const double* ri = beads + ((int64_t)t * c.N + i) * D;
    const double qi = c.charge[species_of(c, i)];
    double e = 0;
    for (int j = 0; j < i; ++j) {
      const double* rj = beads + ((int64_t)t * c.N + j) * D;
      double delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = ri[d] - rj[d];
      pbc<D>(c, delta);
      double dd = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) dd += delta[d] * delta[d];
      e += qi * c.charge[species_of(c, j)] / (eps * sqrt(dd));
    }
    acc += e;
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) part[threadIdx.x / PQ_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x / PQ_WAVE); ++w) v += part[w];
    energy[chain] = v / c.M;
  }
}

// CompositeAction::getBeadAction (CompositeAction.cc:57-67) on bead (i, t):
// SpringAction (SpringAction.cc:112-123: utau and forces, u = 0), SHOAction
// (SHOAction.cc:71-105), PairAction (PairAction.cc:509-599, no exchange) with
// uk0CalcDerivatives (:640-657), and with lr CoulombAction's long-range
// Ewald part on particle 0 (CoulombAction.cc:164-170: u += tau LR, utau +=
// LR, no force).  beads [M][N][D] with perm = the path permutation and its
// inverse ([2][N]): slices -1 and M wrap through them (SerialPaths.cc:68-117).
// out = u, utau, fm[D] (from the link into slice t), fp[D] (the link out);
// ulambda is 0 for every action here.
template <int D>
__device__ void bead_action(const DevCtx& x, const ChainP& c, const double* beads, const int* perm, int i, int t,
                            bool lr, double* out) {
  const int* invperm = perm + c.N;
  auto P = [&](int p, int tt) -> const double* {
    if (tt < 0) { tt += c.M; p = invperm[p]; }
    else if (tt >= c.M) { tt -= c.M; p = perm[p]; }
    return beads + ((int64_t)tt * c.N + p) * D;
  };
  double* fm = out + 2;
  double* fp = out + 2 + D;
  const int sp = species_of(c, i);
  const double* b1 = P(i, t);
  const double* b0 = P(i, t - 1);
  const double* b2 = P(i, t + 1);
  double dm[D], dp[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { dm[d] = b1[d] - b0[d]; dp[d] = b1[d] - b2[d]; fm[d] = 0; fp[d] = 0; }
  pbc<D>(c, dm);
  pbc<D>(c, dp);
  double u = 0, utau = 0;
  if (c.spring && !c.stat[sp]) {
    const double lam = c.lambda[sp], tau = c.tau;
    double dd = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) { fm[d] -= dm[d] / (2 * lam * tau); fp[d] -= dp[d] / (2 * lam * tau); dd += dm[d] * dm[d]; }
    utau += D / (2 * tau) - dd / (4.0 * lam * tau * tau);
  }
  for (int a = 0; a < c.nsho; ++a) {
    const int ssp = c.sho_sp[a];
    if (i < c.ifirst[ssp] || i >= c.ifirst[ssp] + c.count[ssp]) continue;
    const int ndim = c.sho_ndim[a] ? c.sho_ndim[a] : D;
    const double omega = c.sho_omega[a], m = c.mass[ssp], tau = c.tau;
    const double wt = omega * tau, sinhwt = sinh(wt), coshwt = cosh(wt);
    const double cschwt = 1.0 / sinhwt, cothwt = coshwt * cschwt, tin = 1.0 / tau;
    double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
    for (int d = D - ndim; d < D; ++d) {
      const double r1 = b1[d] - c.sho_center[a][d], r0 = b0[d] - c.sho_center[a][d], r2 = b2[d] - c.sho_center[a][d];
      r0r0 += r0 * r0; r0r1 += r0 * r1; r1r1 += r1 * r1; delta2 += dm[d] * dm[d];
      fm[d] -= m * omega * (r1 * coshwt - r0) * cschwt - m * dm[d] * tin;
      fp[d] -= m * omega * (r1 * coshwt - r2) * cschwt - m * dp[d] * tin;
    }
    utau += 0.5 * ndim * omega * cothwt
            + 0.5 * m * omega * omega * ((r1r1 + r0r0) * (1 - cothwt * cothwt) + 2.0 * r0r1 * coshwt * cschwt * cschwt)
            + 0.5 * m * delta2 * tin * tin - 0.5 * ndim * tin;
    u += 0.5 * ndim * log(2 * 3.14159265358979 * sinhwt / (m * omega)) +
         m * omega * ((r1r1 + r0r0) * coshwt - 2.0 * r0r1) * cschwt * 0.5 -
         0.5 * log(2.0 * 3.14159265358979 * tau / m) - 0.5 * m * delta2 / tau;
  }
  for (int pp = 0; pp < c.npair; ++pp) {
    const PairP& PP = c.pair[pp];
    int js;
    if (sp == PP.s1) js = PP.s2;
    else if (sp == PP.s2) js = PP.s1;
    else continue;
    const TableP T = x.tables[PP.table];
    const double* tu = x.tab_u + T.off;
    const double* ut = x.tab_ut + T.off;
    for (int j = c.ifirst[js]; j < c.ifirst[js] + c.count[js]; ++j) {
      if (j == i) continue;
      const double* a1 = P(j, t);
      const double* a0 = P(j, t - 1);
      const double* a2 = P(j, t + 1);
      double dl[D], pd[D], nd[D];
#pragma unroll
      for (int d = 0; d < D; ++d) { dl[d] = b1[d] - a1[d]; pd[d] = b0[d] - a0[d]; nd[d] = b2[d] - a2[d]; }
      pbc<D>(c, dl);
      pbc<D>(c, pd);
      pbc<D>(c, nd);
      double r = 0, pr = 0, nr = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) { r += dl[d] * dl[d]; pr += pd[d] * pd[d]; nr += nd[d] * nd[d]; }
      r = sqrt(r); pr = sqrt(pr); nr = sqrt(nr);
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const double* od = side ? nd : pd;
        const double q = 0.5 * (r + (side ? nr : pr));
        double sv[D], s2 = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) { sv[d] = dl[d] - od[d]; s2 += sv[d] * sv[d]; }
        s2 /= q * q;
        // uk0CalcDerivatives: v, vtau, dv/dq, dv/ds2
        const double qq = q * T.rgridinv;
        double xx = log(qq) * T.logrratioinv;
        int ig = (int)xx;
        if (qq < 1) { ig = 0; xx = 0; }
        else if (ig > T.ngpts - 2) { ig = T.ngpts - 2; xx = 1; }
        else xx -= ig;
        double v = 0, vt = 0, vq = 0, vs2 = 0;
        for (int k = T.K1 - 1; k >= 0; --k) {
          v *= s2; v += (1 - xx) * tu[ig * T.K1 + k] + xx * tu[(ig + 1) * T.K1 + k];
          vt *= s2; vt += (1 - xx) * ut[ig * T.K1 + k] + xx * ut[(ig + 1) * T.K1 + k];
          vq *= s2; vq += tu[(ig + 1) * T.K1 + k] - tu[ig * T.K1 + k];
        }
        vq *= T.logrratioinv * T.rgridinv / qq;
        for (int k = T.K1 - 1; k > 0; --k) {
          vs2 *= s2; vs2 += k * ((1 - xx) * tu[ig * T.K1 + k] + xx * tu[(ig + 1) * T.K1 + k]);
        }
        if (side == 0) { u += 0.5 * v; utau += 0.5 * vt; }
        double* f = side ? fp : fm;
#pragma unroll
        for (int d = 0; d < D; ++d) f[d] -= vq * dl[d] / (2 * r) + vs2 * (2 * sv[d] / (q * q) - s2 * dl[d] / (q * r));
      }
    }
  }
  if (lr && c.ew.nk && i == 0) {
    const double l = ewald_slice<D>(x.ewk + c.ew.off, c.ew, c, beads + (int64_t)t * c.N * D) / c.ceps;
    u += l * c.tau;
    utau += l;
  }
  out[0] = u;
  out[1] = utau;
}

// utau and the force fm + fp of bead (i, t) of the resident paths (the
// virial estimator adds the long-range Ewald utau in parallel itself)
template <int D>
__device__ double bead_force(const DevCtx& x, const ChainP& c, const double* beads, int i, int t, double* f) {
  double o[2 + 2 * D];
  bead_action<D>(x, c, beads, x.perm + c.perm_off, i, t, false, o);
#pragma unroll
  for (int d = 0; d < D; ++d) f[d] = o[2 + d] + o[2 + D + d];
  return o[1];
}

// VirialEnergyEstimator (VirialEnergyEstimator.cc:27-78) over sumOverLinks
// with firstSlice 0.  Its running sums restart at every window of nwindow
// links of a particle, so lanes take (particle, window) items; a window
// closes with F_av . r0 / 2tau.  FixedNodeAction adds zero force and utau;
// static particles add utau only (VirialEnergyEstimator.cc:36,51,60,66).
template <int D>
__global__ __launch_bounds__(256) void virial_energy_kernel(DevCtx x, int nwindow, double* energy) {
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
  const int* invperm = x.perm + c.perm_off + c.N;
  const int M = c.M, nwin = (M + nwindow - 1) / nwindow;
  const double tau = c.tau;
  double acc = 0;
  for (int it = threadIdx.x; it < c.N * nwin; it += blockDim.x) {
    const int i = it / nwin, w = it - i * nwin;
    const bool stat = c.stat[species_of(c, i)];
    const int t0 = w * nwindow, t1 = min(M, t0 + nwindow);
    const int W = (t0 / nwindow < (M - 1) / nwindow) ? nwindow : (M - 1) % nwindow + 1;
    double r0[D], rav[D], fav[D];
#pragma unroll
    for (int d = 0; d < D; ++d) { r0[d] = 0; rav[d] = 0; fav[d] = 0; }
    double e = 0;
    for (int t = t0; t < t1; ++t) {
      const double* end = beads + ((int64_t)t * c.N + i) * D;
      const double* start = (t > 0) ? beads + ((int64_t)(t - 1) * c.N + i) * D
                                    : beads + ((int64_t)(M - 1) * c.N + invperm[i]) * D;
      double r[D], f[D];
#pragma unroll
      for (int d = 0; d < D; ++d) r[d] = end[d] - start[d];
      pbc<D>(c, r);
#pragma unroll
      for (int d = 0; d < D; ++d) { rav[d] += r[d]; r0[d] += (W - t % nwindow) * r[d] / W; }
      const double utau = bead_force<D>(x, c, beads, i, t, f);
      double fr = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) { fav[d] += f[d]; fr += f[d] * rav[d]; }
      e += utau;
      if (!stat) e += -(1. - 1. / W) * (D / (2 * tau)) - (1 / (2 * tau)) * fr;
    }
    double fr0 = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) fr0 += fav[d] * r0[d];
    acc += e + (stat ? 0.0 : (1 / (2 * tau)) * fr0);
  }
  // CoulombAction's long-range utau of particle 0 on every slice (no force)
  if (c.ew.nk) acc += ewald_all_slices<D>(x.ewk + c.ew.off, c.ew, c, beads) / c.ceps;
  acc = wave_sum(acc);
  if (lane_id() == 0) part[threadIdx.x / PQ_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x / PQ_WAVE); ++w) v += part[w];
    energy[chain] = v / M;
  }
}

// getBeadAction of a set of beads of one chain: out[q] = u, utau, ulambda
// (0), fm[D], fp[D] of bead (qi[q], qt[q]) on the paths beads/perm
template <int D>
__global__ __launch_bounds__(256) void bead_action_kernel(DevCtx x, int chain, const double* beads, const int* perm,
                                                           int nq, const int* qi, const int* qt, double* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const ChainP& c = x.chains[chain];
  double o[2 + 2 * D];
  bead_action<D>(x, c, beads, perm, qi[q], qt[q], true, o);
  double* w = out + (int64_t)q * (3 + 2 * D);
  w[0] = o[0];
  w[1] = o[1];
  w[2] = 0.0;
#pragma unroll
  for (int d = 0; d < 2 * D; ++d) w[3 + d] = o[2 + d];
}

// EwaldCoulombEstimator with its own OptEwaldSum (EwaldCoulombEstimator.cc:
// 63-140): per link q_i q_j (1/(r eps) + (1-x) v(ig) + x v(ig+1)) over pairs
// j < i inside rcut, v(i) = -f(i dr)/eps (v(0) = -f(0)/eps) on 1001 points
// (one more stored: r just below rcut can round to ig = 1000),
// dr = rcut/1000, ig = (int)(r/dr) and x = r - ig dr as the reference weighs
// it, and the k-space sum on particle 0 of every slice; divided by M.  One
// workgroup per chain, ew[chain] its sum.
template <int D>
__global__ __launch_bounds__(256) void ewald_coulomb_energy_kernel(DevCtx x, const EwP* ew, const double* kpool,
                                                                    const int64_t* voff, const double* vtab,
                                                                    double eps, double* energy) {
  __shared__ ChainP c;
  __shared__ double part[4];
  const int chain = blockIdx.x;
  {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&x.chains[chain]);
    uint32_t* d = reinterpret_cast<uint32_t*>(&c);
    for (int k = threadIdx.x; k < (int)(sizeof(ChainP) / 4); k += blockDim.x) d[k] = s[k];
    __syncthreads();
  }
  const EwP e = ew[chain];
  const double* beads = x.beads + c.bead_off;
  const double dr = e.rcut / 1000, drinv = 1. / dr;
  // tradEwald: -f(i dr) tabulated on the host; optEwald: the fitted polynomial
  const int64_t vo = voff[chain];
  auto vgrid = [&](int ig) -> double {
    if (vo >= 0) return vtab[vo + ig] / eps;
    if (ig == 0) return -e.coef[0] / eps;
    const double r = ig * dr;
    if (r > e.rcut) return -(1. / r) / eps;
    double v = 0., r2 = r * r;
    for (int k = 7; k >= 0; --k) { v *= r2; v += e.coef[k]; }
    return -v / eps;
  };
  double acc = 0;
  for (int q = threadIdx.x; q < c.M * c.N; q += blockDim.x) {
    const int t = q / c.N, i = q - t * c.N;
    const double* ri = beads + ((int64_t)t * c.N + i) * D;
    const double qi = c.charge[species_of(c, i)];
    double en = 0;
    for (int j = 0; j < i; ++j) {
      const double* rj = beads + ((int64_t)t * c.N + j) * D;
      double delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = ri[d] - rj[d];
      pbc<D>(c, delta);
      double dd = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) dd += delta[d] * delta[d];
      const double r = sqrt(dd);
      if (r < e.rcut) {
        const int ig = (int)(r * drinv);
        const double xw = r - ig * dr;
        en += qi * c.charge[species_of(c, j)] * (1. / (r * eps) + (1 - xw) * vgrid(ig) + xw * vgrid(ig + 1));
      }
    }
    acc += en;
  }
  acc += ewald_all_slices<D>(kpool + e.off, e, c, beads) / eps;
  acc = wave_sum(acc);
  if (lane_id() == 0) part[threadIdx.x / PQ_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x / PQ_WAVE); ++w) v += part[w];
    energy[chain] = v / c.M;
  }
}

// ------------------------------------------------------------ launchers

#define PQ_DISPATCH_R(FN, D_, R_, ...)                                              \
  do {                                                                                \
    switch (NC) {                                                                     \
      case 4: return FN<D_, 4, R_>(__VA_ARGS__);                                      \
      case 8: return FN<D_, 8, R_>(__VA_ARGS__);                                      \
      case 10: return FN<D_, 10, R_>(__VA_ARGS__);                                    \
      case 12: return FN<D_, 12, R_>(__VA_ARGS__);                                    \
      case 16: return FN<D_, 16, R_>(__VA_ARGS__);                                    \
      default: return FN<D_, 32, R_>(__VA_ARGS__);                                    \
    }                                                                                 \
  } while (0)

#define PQ_DISPATCH(FN, ...)                                                          \
  do {                                                                                \
    if (mixed) {                                                                      \
      if (D == 2) PQ_DISPATCH_R(FN, 2, float, __VA_ARGS__);                           \
      PQ_DISPATCH_R(FN, 3, float, __VA_ARGS__);                                       \
    }                                                                                 \
    if (D == 2) PQ_DISPATCH_R(FN, 2, double, __VA_ARGS__);                            \
    PQ_DISPATCH_R(FN, 3, double, __VA_ARGS__);                                        \
  } while (0)

hipError_t launch_sweep_2f(int, bool, const DevCtx&, int, int, const SweepLayoutH&, hipStream_t);
hipError_t launch_sweep_2d(int, bool, const DevCtx&, int, int, const SweepLayoutH&, hipStream_t);
hipError_t launch_sweep_3f(int, bool, const DevCtx&, int, int, const SweepLayoutH&, hipStream_t);
hipError_t launch_sweep_3d(int, bool, const DevCtx&, int, int, const SweepLayoutH&, hipStream_t);

hipError_t launch_sweep(int D, int NC, bool mixed, bool multi, const DevCtx& x, int nchain, int nsweep,
                        const SweepLayoutH& L, hipStream_t st) {
  if (D == 2)
    return mixed ? launch_sweep_2f(NC, multi, x, nchain, nsweep, L, st) : launch_sweep_2d(NC, multi, x, nchain, nsweep, L, st);
  return mixed ? launch_sweep_3f(NC, multi, x, nchain, nsweep, L, st) : launch_sweep_3d(NC, multi, x, nchain, nsweep, L, st);
}

template <int D, int NC, class R>
static hipError_t launch_action_diff_t(const DevCtx& x, int nq, const CallBufsH& b, hipStream_t st) {
  CallBufs cb{b.qchain, b.qlevel, b.qsec, b.qmov, b.qidx, b.nmov, b.sec_stride, b.mov_stride, b.out};
  hipLaunchKernelGGL((action_diff_kernel<D, NC, R>), dim3(nq), dim3(PQ_WAVE), 0, st, x, cb);
  return hipGetLastError();
}

hipError_t launch_action_diff(int D, int NC, bool mixed, const DevCtx& x, int nq, const CallBufsH& b,
                              hipStream_t st) {
  (void)NC;
  if (mixed) { if (D == 2) return launch_action_diff_t<2, 4, float>(x, nq, b, st);
               return launch_action_diff_t<3, 4, float>(x, nq, b, st); }
  if (D == 2) return launch_action_diff_t<2, 4, double>(x, nq, b, st);
  return launch_action_diff_t<3, 4, double>(x, nq, b, st);
}

template <int D, int NC, class R>
static hipError_t launch_fn_t(const DevCtx& x, int chain, const FnBufsH& b, hipStream_t st) {
  FnBufs fb{b.sec1, b.sec2, b.mov, b.wdet, b.werr, b.wscale, b.notmine, b.op, b.level, b.idx, b.nmov, b.out};
  hipLaunchKernelGGL((fn_kernel<D, NC, R>), dim3(1), dim3(PQ_WAVE), 0, st, x, chain, fb);
  return hipGetLastError();
}

hipError_t launch_fn(int D, int NC, bool mixed, const DevCtx& x, int chain, const FnBufsH& b,
                     hipStream_t st) {
  PQ_DISPATCH(launch_fn_t, x, chain, b, st);
}

template <int D, int NC, class R>
static hipError_t launch_node_eval_t(const DevCtx& x, int chain, int fa, int nmat, double* r1,
                                     double* r2, double* inv, double* egr, double* det, int* err,
                                     double* dist, hipStream_t st) {
  hipLaunchKernelGGL((node_eval_kernel<D, NC, R>), dim3(1), dim3(PQ_WAVE), 0, st, x, chain, fa, nmat,
                     r1, r2, inv, egr, det, err, dist);
  return hipGetLastError();
}

hipError_t launch_node_eval(int D, int NC, bool mixed, const DevCtx& x, int chain, int fa, int nmat,
                            double* r1, double* r2, double* inv, double* egr, double* det, int* err,
                            double* dist, hipStream_t st) {
  PQ_DISPATCH(launch_node_eval_t, x, chain, fa, nmat, r1, r2, inv, egr, det, err, dist, st);
}

hipError_t launch_coulomb_energy(int D, const DevCtx& x, int nchain, double eps, double* energy, hipStream_t st) {
  if (D == 2) hipLaunchKernelGGL(coulomb_energy_kernel<2>, dim3(nchain), dim3(256), 0, st, x, eps, energy);
  else hipLaunchKernelGGL(coulomb_energy_kernel<3>, dim3(nchain), dim3(256), 0, st, x, eps, energy);
  return hipGetLastError();
}

hipError_t launch_ewald_coulomb_energy(int D, const DevCtx& x, int nchain, const EwP* ew, const double* kpool,
                                       const int64_t* voff, const double* vtab, double eps, double* energy,
                                       hipStream_t st) {
  if (D == 2)
    hipLaunchKernelGGL(ewald_coulomb_energy_kernel<2>, dim3(nchain), dim3(256), 0, st, x, ew, kpool, voff, vtab, eps, energy);
  else
    hipLaunchKernelGGL(ewald_coulomb_energy_kernel<3>, dim3(nchain), dim3(256), 0, st, x, ew, kpool, voff, vtab, eps, energy);
  return hipGetLastError();
}

hipError_t launch_virial_energy(int D, const DevCtx& x, int nchain, int nwindow, double* energy, hipStream_t st) {
  if (D == 2) hipLaunchKernelGGL(virial_energy_kernel<2>, dim3(nchain), dim3(256), 0, st, x, nwindow, energy);
  else hipLaunchKernelGGL(virial_energy_kernel<3>, dim3(nchain), dim3(256), 0, st, x, nwindow, energy);
  return hipGetLastError();
}

hipError_t launch_bead_action(int D, const DevCtx& x, int chain, const double* beads, const int* perm, int nq,
                              const int* qi, const int* qt, double* out, hipStream_t st) {
  const int nb = (nq + 255) / 256;
  if (D == 2) hipLaunchKernelGGL(bead_action_kernel<2>, dim3(nb), dim3(256), 0, st, x, chain, beads, perm, nq, qi, qt, out);
  else hipLaunchKernelGGL(bead_action_kernel<3>, dim3(nb), dim3(256), 0, st, x, chain, beads, perm, nq, qi, qt, out);
  return hipGetLastError();
}

}  // namespace pq
