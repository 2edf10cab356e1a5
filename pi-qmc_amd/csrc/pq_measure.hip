// pq_measure.hip -- the Measure step on the resident paths: thermodynamic
// energy (ThermoEnergyEstimator over CompositeAction::getBeadAction) and the
// pair-correlation histogram (PairCFEstimator<1> with PairRadial).
//
// Work is split into items of consecutive slices of one chain, sized so that
// every item costs about the same whatever the chain's N and M: a chain of
// 64 particles gives a few slices per item, a chain of 2 particles hundreds.
// Each workgroup takes a contiguous run of the item list; an item stages its slices (and the
// slice before the first, through the path permutation at slice 0) and the
// chain's utau tables in LDS, so every bead and table entry is read from HBM
// or L2 once per item.  Energy items write one partial sum each and a second
// kernel adds a chain's partials in item order: the result does not depend
// on scheduling.  Histogram items count in LDS and add their bins to the
// chain's global counts with integer atomics, also order-independent.
#include "pq_device.h"
#include "pq_launch.h"

namespace pq {

constexpr int kMeasThreads = 256;     // histogram workgroups
constexpr int kEnergyThreads = 512;   // energy workgroups: two per CU share their tables across eight waves
constexpr int kMeasPairLinks = 16;     // links per lane and pair in the energy items

// the CoulombAction pair table of species (si, sj), -1 when they do not interact
struct PairMap {
  int slot[PQ_MAX_SPECIES][PQ_MAX_SPECIES];
};

// Stage slices t0-1 .. t1-1 of chain c into sl[(t1-t0+1)][N][D]; slice -1 is
// slice M-1 of inversePermutation[i] (SerialPaths.cc:68-77).
template <int D>
__device__ void stage_slices(const DevCtx& x, const ChainP& c, int t0, int t1, double* sl, int ss) {
  const double* beads = x.beads + c.bead_off;
  const int* invperm = x.perm + c.perm_off + c.N;
  const int nd = c.N * D;
  const int n = (t1 - t0 + 1) * nd;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int k = q / nd, r = q - k * nd;
    const int t = t0 - 1 + k;
    double v;
    if (t < 0) {
      const int i = r / D, d = r - i * D;
      v = beads[((int64_t)(c.M - 1) * c.N + invperm[i]) * D + d];
    } else {
      v = beads[(int64_t)t * nd + r];
    }
    sl[k * ss + r] = v;
  }
}

// q / d and q % d for q < 2^16 and 0 < d < 2^16 by a float reciprocal:
// (q + 1/2)/d sits at least 1/(2d) from an integer, and the float product's
// error stays below q 2^-23 < 1/(2d) there, so the truncation is exact
__device__ __forceinline__ int div_small(int q, int d, float invd) { return (int)(((float)q + 0.5f) * invd); }

// unordered pair p of n particles, i < j: p < n ((n-1)/2) walks the circulant
// differences 1..(n-1)/2, the rest (even n only) the difference n/2
__device__ __forceinline__ void pair_of(int n, float invn, int p, int& i, int& j) {
  const int full = n * ((n - 1) / 2);
  int a, b;
  if (p < full) {
    const int dd = div_small(p, n, invn);
    a = p - dd * n;
    b = a + dd + 1;
    if (b >= n) b -= n;
  } else {
    a = p - full;
    b = a + n / 2;
  }
  i = min(a, b);
  j = max(a, b);
}

// the separation of the pair (i, j) on staged slice k: pbc(r_i - r_j) and its length
template <int D>
__device__ __forceinline__ double pair_sep(const ChainP& c, const double* sl, int ss, int k, int i, int j,
                                           double* dl) {
  const double* a = sl + k * ss + i * D;
  const double* b = sl + k * ss + j * D;
#pragma unroll
  for (int d = 0; d < D; ++d) dl[d] = a[d] - b[d];
  pbc<D>(c, dl);
  double r = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) r += dl[d] * dl[d];
  return fast_sqrt(r);
}

// ln x for the grid position of a link, x >= 1: fast_log's reduction with
// the atanh series to s^17 (relative truncation below (0.1716)^18/19 =
// 9e-16 of ln m, |ln m| < 0.35); the position x ln ratio^-1 then errs by
// ~1e-13 grid steps, far inside the estimator's 1e-12.
__device__ __forceinline__ double grid_log(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752) { m += m; --e; }
  const double num = m - 1.0, den = m + 1.0;
  double r = __builtin_amdgcn_rcp(den);
  r = fma(fma(-den, r, 1.0), r, r);
  r = fma(fma(-den, r, 1.0), r, r);
  double s = num * r;
  s = fma(fma(-den, s, num), r, s);
  const double s2 = s * s;
  double p = 1.0 / 17;
  p = fma(p, s2, 1.0 / 15);
  p = fma(p, s2, 1.0 / 13);
  p = fma(p, s2, 1.0 / 11);
  p = fma(p, s2, 1.0 / 9);
  p = fma(p, s2, 1.0 / 7);
  p = fma(p, s2, 1.0 / 5);
  p = fma(p, s2, 1.0 / 3);
  const double t = (s + s) * (s2 * p);
  const double de = (double)e;
  return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, t) + (s + s));
}

// vtau of a pair on one link from its separations dl (slice t, length r) and
// pd (slice t-1, length pr) (uk0CalcDerivatives' vtau, PairAction.cc:640-657):
// the Horner sums in s^2/q^2 of the two bracketing grid points' coefficients,
// A and B, interpolated as A + x (B - A).  K1 > 0 fixes the number of
// coefficients at compile time (the table row of both grid points is read at
// once); K1 = 0 reads tb.K1.  Coefficient k of grid point g is u[g sg + k sk]:
// rows [g][k] in global memory (sg = K1, sk = 1), planes [k][g] in LDS (sg =
// 1, sk = ngpts), where a lane's two grid points are adjacent doubles and
// random rows of different lanes spread over the banks.
template <int D, int K1>
__device__ __forceinline__ double link_vt(const TableP& tb, const double* u, int sg, int sk, const double* dl,
                                          double r, const double* pd, double pr) {
  const double qm = 0.5 * (r + pr);
  double s2 = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) { const double sv = dl[d] - pd[d]; s2 += sv * sv; }
  s2 *= fast_rcp(qm * qm);
  const double qq = qm * tb.rgridinv;
  double xx = 0;
  int ig = 0;
  if (qq >= 1) {
    xx = grid_log(qq) * tb.logrratioinv;
    ig = (int)xx;
    if (ig > tb.ngpts - 2) { ig = tb.ngpts - 2; xx = 1; }
    else xx -= ig;
  }
  const int k1 = K1 > 0 ? K1 : tb.K1;
  const double* u0 = u + ig * sg;
  double a = u0[(k1 - 1) * sk], b = u0[(k1 - 1) * sk + sg];
  if constexpr (K1 > 0) {
#pragma unroll
    for (int kk = K1 - 2; kk >= 0; --kk) { a = fma(a, s2, u0[kk * sk]); b = fma(b, s2, u0[kk * sk + sg]); }
  } else {
    for (int kk = k1 - 2; kk >= 0; --kk) { a = fma(a, s2, u0[kk * sk]); b = fma(b, s2, u0[kk * sk + sg]); }
  }
  return fma(xx, b - a, a);
}

// The links k0..k1-1 of the pair (i, j): a lane takes each link's earlier
// separation from the previous link (one pbc and one sqrt per link).
template <int D, int K1>
__device__ __forceinline__ double pair_links(const ChainP& c, const TableP& tb, const double* u, int sg, int sk,
                                             const double* sl, int ss, int k0, int k1, int i, int j) {
  double acc = 0;
  double pd[D], dl[D];
  double pr = pair_sep<D>(c, sl, ss, k0, i, j, pd);
  for (int k = k0; k < k1; ++k) {
    const double r = pair_sep<D>(c, sl, ss, k + 1, i, j, dl);
    acc += link_vt<D, K1>(tb, u, sg, sk, dl, r, pd, pr);
#pragma unroll
    for (int d = 0; d < D; ++d) pd[d] = dl[d];
    pr = r;
  }
  return acc;
}

// utau of the links (t-1, t) of slices t0..t1-1 of one chain, this thread's
// share: SpringAction (SpringAction.cc:112-123), SHOAction (SHOAction.cc:
// 71-105), PairAction (PairAction.cc:509-599, uk0CalcDerivatives :640-657)
// counted once per unordered pair -- each bead of the pair takes half of the
// same vtau -- and CoulombAction's long-range utau of the slices
// (CoulombAction.cc:164-170).  Items hold fewer than 2^16 (slice, pair)
// entries, so indices split by div_small.
template <int D>
__device__ __forceinline__ double item_utau(const DevCtx& x, const ChainP& c, const PairMap& pm, const double* sl,
                                            const double* tab, bool planes, const int* utoff,
                                            const TableP* tbs, int t0, int t1) {
  const int N = c.N, T = t1 - t0;
  const int ss = meas_slice_stride(N * D);
  const float invN = 1.0f / N;
  double acc = 0;
  // spring and SHO: lanes over (slice, particle)
  if (c.spring || c.nsho) {
    for (int q = threadIdx.x; q < T * N; q += blockDim.x) {
      const int k = div_small(q, N, invN), i = q - k * N;
      const double* b1 = sl + (k + 1) * ss + i * D;
      const double* b0 = sl + k * ss + i * D;
      const int sp = species_of(c, i);
      double delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = b1[d] - b0[d];
      pbc<D>(c, delta);
      if (c.spring && !c.stat[sp]) {
        double dd = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) dd += delta[d] * delta[d];
        acc += D / (2 * c.tau) - dd / (4.0 * c.lambda[sp] * c.tau * c.tau);
      }
      for (int a = 0; a < c.nsho; ++a) {
        const int ssp = c.sho_sp[a];
        if (i < c.ifirst[ssp] || i >= c.ifirst[ssp] + c.count[ssp]) continue;
        const int ndim = c.sho_ndim[a] ? c.sho_ndim[a] : D;
        const double omega = c.sho_omega[a], m = c.mass[ssp];
        // cosh(omega tau) and 1/(2 sinh(omega tau)) of the chain block
        const double coshwt = c.sho_cosh[a][0], cschwt = 2.0 * c.sho_div1[a][0];
        const double cothwt = coshwt * cschwt, tin = 1.0 / c.tau;
        double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
        for (int d = D - ndim; d < D; ++d) {
          const double r1 = b1[d] - c.sho_center[a][d], r0 = b0[d] - c.sho_center[a][d];
          r0r0 += r0 * r0; r0r1 += r0 * r1; r1r1 += r1 * r1; delta2 += delta[d] * delta[d];
        }
        acc += 0.5 * ndim * omega * cothwt
               + 0.5 * m * omega * omega * ((r1r1 + r0r0) * (1 - cothwt * cothwt) + 2.0 * r0r1 * coshwt * cschwt * cschwt)
               + 0.5 * m * delta2 * tin * tin - 0.5 * ndim * tin;
      }
    }
  }
  // pair actions: lanes over (pair, group of up to kMeasPairLinks links),
  // pair-major, so that the lanes of one pair walk consecutive slices and
  // look up the same or neighbouring table rows (random rows across a wave
  // make the gathers conflict in the LDS banks); a lane takes each link's
  // earlier separation from the previous link
  if (c.npair) {
    const int P = N * (N - 1) / 2;
    const int H = (T + kMeasPairLinks - 1) / kMeasPairLinks;
    const float invH = 1.0f / H;
    for (int q = threadIdx.x; q < H * P; q += blockDim.x) {
      const int p = div_small(q, H, invH), h = q - p * H;
      int i, j;
      pair_of(N, invN, p, i, j);
      const int slot = pm.slot[species_of(c, i)][species_of(c, j)];
      if (slot < 0) continue;
      const TableP tb = tbs[slot];
      const double* u = tab + utoff[slot];
      const int k0 = (h * T) / H, k1 = ((h + 1) * T) / H;   // groups of equal length
      const int sg = planes ? 1 : tb.K1, sk = planes ? tb.ngpts : 1;
      if (tb.K1 == 4) acc += pair_links<D, 4>(c, tb, u, sg, sk, sl, ss, k0, k1, i, j);
      else if (tb.K1 == 1) acc += pair_links<D, 1>(c, tb, u, sg, sk, sl, ss, k0, k1, i, j);
      else acc += pair_links<D, 0>(c, tb, u, sg, sk, sl, ss, k0, k1, i, j);
    }
  }
  // Ewald k-space utau of every slice (particle 0 carries it): lanes over (slice, k)
  if (c.ew.nk) {
    const double* kp = x.ewk + c.ew.off;
    double lr = 0;
    for (int q = threadIdx.x; q < T * c.ew.nk; q += blockDim.x) {
      const int k = q / c.ew.nk;
      const double* kv = kp + 4 * (q - k * c.ew.nk);
      const double* s = sl + (k + 1) * ss;
      double rr = 0, ri = 0;
      for (int jj = 0; jj < N; ++jj) {
        double co, si;
        ewald_phase<D>(kv, s + jj * D, co, si);
        const double qj = c.charge[species_of(c, jj)];
        rr += qj * co;
        ri += qj * si;
      }
      lr += kv[3] * (rr * rr + ri * ri);
    }
    if (threadIdx.x == 0) lr += T * c.ew.self;
    acc += lr / c.ceps;
  }
  return acc;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if (lane_id() == 0) red[threadIdx.x / PQ_WAVE] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / PQ_WAVE); ++w) s += red[w];
  __syncthreads();
  return s;
}

// LDS budget of the utau tables staged per item; larger table sets are read
// from global memory (L2)
constexpr int kTabLdsDoubles = 4096;

template <int D>
__global__ __launch_bounds__(kEnergyThreads, 4) void energy_items_kernel(DevCtx x, const MeasItem* items, int nitems,
                                                                     double* part) {
  extern __shared__ double lds[];
  __shared__ ChainP c;
  __shared__ PairMap pm;
  __shared__ int utoff[PQ_MAX_PAIRS];
  __shared__ TableP tbs[PQ_MAX_PAIRS];
  __shared__ int in_lds;
  __shared__ double red[kEnergyThreads / PQ_WAVE];
  int cur = -1;
  double* tab = lds;
  double* sl = lds + kTabLdsDoubles;
  // contiguous runs of items per workgroup: a workgroup meets few chains and
  // stages each chain's parameters and tables once
  const int run = (nitems + gridDim.x - 1) / gridDim.x;
  const int it_end = min(nitems, (int)(blockIdx.x + 1) * run);
  for (int it = blockIdx.x * run; it < it_end; ++it) {
    const MeasItem m = items[it];
    if (m.chain != cur) {
      __syncthreads();
      copy_chain(&c, &x.chains[m.chain]);
      if (threadIdx.x == 0) {
        for (int a = 0; a < PQ_MAX_SPECIES; ++a)
          for (int b = 0; b < PQ_MAX_SPECIES; ++b) pm.slot[a][b] = -1;
        // the chain's distinct utau tables, packed; staged in LDS when they fit
        int used = 0;
        for (int pp = 0; pp < c.npair; ++pp) {
          pm.slot[c.pair[pp].s1][c.pair[pp].s2] = pp;
          pm.slot[c.pair[pp].s2][c.pair[pp].s1] = pp;
          tbs[pp] = x.tables[c.pair[pp].table];
          int q = 0;
          while (q < pp && c.pair[q].table != c.pair[pp].table) ++q;
          if (q < pp) { utoff[pp] = utoff[q]; continue; }
          const TableP& tb = x.tables[c.pair[pp].table];
          utoff[pp] = used;
          used += tb.ngpts * tb.K1;
        }
        in_lds = used <= kTabLdsDoubles;
        if (!in_lds)
          for (int pp = 0; pp < c.npair; ++pp) utoff[pp] = (int)x.tables[c.pair[pp].table].off;
      }
      __syncthreads();
      if (in_lds)
        for (int pp = 0; pp < c.npair; ++pp) {
          int q = 0;
          while (q < pp && c.pair[q].table != c.pair[pp].table) ++q;
          if (q < pp) continue;
          const TableP& tb = x.tables[c.pair[pp].table];
          const double* src = x.tab_ut + tb.off;
          // rows [g][k] to planes [k][g]
          for (int e = threadIdx.x; e < tb.ngpts * tb.K1; e += blockDim.x) {
            const int g = e / tb.K1, k = e - g * tb.K1;
            tab[utoff[pp] + k * tb.ngpts + g] = src[e];
          }
        }
      cur = m.chain;
    }
    __syncthreads();
    stage_slices<D>(x, c, m.t0, m.t1, sl, meas_slice_stride(c.N * D));
    __syncthreads();
    const double v = in_lds ? item_utau<D>(x, c, pm, sl, tab, true, utoff, tbs, m.t0, m.t1)
                            : item_utau<D>(x, c, pm, sl, x.tab_ut, false, utoff, tbs, m.t0, m.t1);
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) part[it] = s;
  }
}

// E = (1/M) sum over a chain's items (SimpleScalarAccumulator's storeValue
// divides by the slice count): one wave per chain, lanes over the items in a
// fixed order and a fixed butterfly, so the sum does not depend on scheduling
__global__ void energy_reduce_kernel(const DevCtx x, const int* first_item, int nchain, const double* part,
                                     double* energy) {
  const int c = blockIdx.x * (blockDim.x / PQ_WAVE) + threadIdx.x / PQ_WAVE;
  if (c >= nchain) return;
  double s = 0;
  for (int it = first_item[c] + lane_id(); it < first_item[c + 1]; it += PQ_WAVE) s += part[it];
  s = wave_sum(s);
  if (lane_id() == 0) energy[c] = s / x.chains[c].M;
}

// SuperCell::pbc without contraction: the bins of the histogram are those of
// the FP64 expression evaluated operation by operation
template <int D>
__device__ __forceinline__ void pbc_exact(const ChainP& c, double* v) {
#pragma clang fp contract(off)
  double r2 = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) r2 += v[d] * v[d];
  if (r2 > c.rcut2) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double i = floor(c.b[d] * v[d] + 1000.5) - 1000;
      v[d] -= c.a[d] * i;
    }
  }
}

// PairCFEstimator<1>::handleLink with PairRadial (PairCFEstimator.h:72-93,
// PairDistance.h:22-39) on slices t0..t1-1: particle i of species sp[0]
// against every j != i of each partner species sp[1..nsp-1] in list order.
template <int D>
__global__ __launch_bounds__(kMeasThreads) void pair_cf_items_kernel(DevCtx x, const MeasItem* items, int nitems,
                                                                      PairCfArgs g, uint64_t* counts) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  __shared__ ChainP c;
  __shared__ int jfirst[PQ_MAX_CF_SPECIES], jcount[PQ_MAX_CF_SPECIES];
  __shared__ int njtot, if1, n1, valid;
  unsigned int* hist = reinterpret_cast<unsigned int*>(lds);
  double* sl = lds + (g.nbin + 1) / 2;
  int cur = -1;
  // contiguous runs of items per workgroup: a workgroup meets few chains and
  // stages each chain's parameters and tables once
  const int run = (nitems + gridDim.x - 1) / gridDim.x;
  const int it_end = min(nitems, (int)(blockIdx.x + 1) * run);
  for (int it = blockIdx.x * run; it < it_end; ++it) {
    const MeasItem m = items[it];
    if (m.chain != cur) {
      __syncthreads();
      copy_chain(&c, &x.chains[m.chain]);
      if (threadIdx.x == 0) {
        valid = g.sp[0] < c.nsp;
        int n = 0;
        for (int k = 1; k < g.nsp; ++k) {
          const bool ok = g.sp[k] < c.nsp;
          valid &= ok;
          jfirst[k] = ok ? c.ifirst[g.sp[k]] : 0;
          jcount[k] = ok ? c.count[g.sp[k]] : 0;
          n += jcount[k];
        }
        njtot = n;
        if1 = valid ? c.ifirst[g.sp[0]] : 0;
        n1 = valid ? c.count[g.sp[0]] : 0;
      }
      __syncthreads();
      cur = m.chain;
    }
    if (!valid || njtot == 0) continue;
    for (int b = threadIdx.x; b < g.nbin; b += blockDim.x) hist[b] = 0;
    stage_slices<D>(x, c, m.t0, m.t1, sl, c.N * D);
    __syncthreads();
    const int nd = c.N * D, per = n1 * njtot, T = m.t1 - m.t0;
    const float invnj = 1.0f / njtot;
    // a lane owns an (i, j) pair over all the item's slices when there are
    // enough pairs for every lane, else lanes over (slice, pair)
    const bool pair_major = per >= (int)blockDim.x / 2;
    const int count = pair_major ? per : T * per;
    for (int q = threadIdx.x; q < count; q += blockDim.x) {
      int k0 = 0, r = q;
      if (!pair_major) {
        k0 = T * per < 65536 ? div_small(q, per, 1.0f / per) : q / per;
        r = q - k0 * per;
      }
      const int ii = per < 65536 ? div_small(r, njtot, invnj) : r / njtot;
      const int i = if1 + ii;
      int rr = r - ii * njtot, kk = 1;
      while (rr >= jcount[kk]) { rr -= jcount[kk]; ++kk; }
      const int j = jfirst[kk] + rr;
      if (i == j) continue;
      const int k1 = pair_major ? T : k0 + 1;
      for (int k = k0; k < k1; ++k) {
        const double* a = sl + (k + 1) * nd + i * D;
        const double* b = sl + (k + 1) * nd + j * D;
        double delta[D];
#pragma unroll
        for (int d = 0; d < D; ++d) delta[d] = a[d] - b[d];
        pbc_exact<D>(c, delta);
        double r2 = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) r2 += delta[d] * delta[d] * 1.0;
        const double dd = sqrt(r2);
        const int ibin = (int)((dd - g.rmin) * g.deltaInv);
        if (dd < g.rmin || ibin > g.nbin - 1) continue;
        atomicAdd(&hist[ibin], 1u);
      }
    }
    __syncthreads();
    uint64_t* out = counts + (int64_t)m.chain * g.nbin;
    for (int b = threadIdx.x; b < g.nbin; b += blockDim.x)
      if (hist[b]) atomicAdd(reinterpret_cast<unsigned long long*>(out + b), (unsigned long long)hist[b]);
    __syncthreads();
  }
}

__global__ void fill_u64_kernel(uint64_t* p, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0;
}

// grids of 32 rounds over 256 CUs of six histogram workgroups (~23 KB of
// LDS, 62 VGPRs each) or two energy workgroups (under 80 KB of LDS each)
// per CU, each workgroup taking a contiguous run of items: short runs keep
// the end of the launch even (4 rounds: 19.0 + 5.8 ms on the 4096-chain 3D
// batch, 32 rounds: 17.2 + 4.9 ms), long enough that a workgroup stages a
// chain's tables for several items
static int meas_grid(int nitems, int per_cu) {
  const int g = 32 * 256 * per_cu;
  return nitems < g ? (nitems > 0 ? nitems : 1) : g;
}

hipError_t launch_energy_items(int D, const DevCtx& x, const MeasItem* items, int nitems, const int* first_item,
                               int nchain, int max_slice_doubles, double* part, double* energy, hipStream_t st) {
  const size_t lds = sizeof(double) * ((size_t)kTabLdsDoubles + max_slice_doubles);
  const int grid = meas_grid(nitems, 2);
  if (D == 2)
    hipLaunchKernelGGL(energy_items_kernel<2>, dim3(grid), dim3(kEnergyThreads), lds, st, x, items, nitems, part);
  else
    hipLaunchKernelGGL(energy_items_kernel<3>, dim3(grid), dim3(kEnergyThreads), lds, st, x, items, nitems, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(energy_reduce_kernel, dim3((nchain + 3) / 4), dim3(4 * PQ_WAVE), 0, st, x, first_item, nchain,
                     part, energy);
  return hipGetLastError();
}

hipError_t launch_pair_cf_items(int D, const DevCtx& x, const MeasItem* items, int nitems, int max_slice_doubles,
                                const PairCfArgs& g, uint64_t* counts, hipStream_t st) {
  const size_t lds = sizeof(double) * ((size_t)(g.nbin + 1) / 2 + max_slice_doubles);
  const int grid = meas_grid(nitems, 6);
  if (D == 2)
    hipLaunchKernelGGL(pair_cf_items_kernel<2>, dim3(grid), dim3(kMeasThreads), lds, st, x, items, nitems, g, counts);
  else
    hipLaunchKernelGGL(pair_cf_items_kernel<3>, dim3(grid), dim3(kMeasThreads), lds, st, x, items, nitems, g, counts);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, int64_t n, hipStream_t st) {
  const int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)(nb < 4096 ? (nb > 0 ? nb : 1) : 4096)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace pq
