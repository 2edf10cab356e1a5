// pq_sweep.h -- the device-resident sampler kernel, included by one
// translation unit per (dimension, arithmetic type) so the instantiations
// compile in parallel.
#pragma once
#include "pq_engine.h"
#include "pq_launch.h"

namespace pq {

// LDS carve-up of the sampler workgroup
struct SweepLayout {
  int off_sec1, off_sec2, off_mov, off_fst, off_wdet, off_werr, off_wscale, off_scale, off_notmine;
  int off_red, off_cnt, off_inv;   // off_inv < 0: inverse cache in global memory
  int off_gj;
  int off_tab, off_tabofs;         // off_tab < 0: pair tables read from global memory
  int bytes;
};

// two waves per SIMD whatever the team: up to 256 threads two workgroups share
// a CU, a 512-thread team has one to itself
template <int D, int NC, class R, int NT>
__global__ __launch_bounds__(NT, NT <= 256 ? 2 : 1) void sweep_kernel(DevCtx x, int nsweep, SweepLayout lay) {
  extern __shared__ double smem[];
  const int chain = x.order[blockIdx.x];
  ChainP& c = *reinterpret_cast<ChainP*>(smem);
  copy_chain(&c, &x.chains[chain]);
  char* base = reinterpret_cast<char*>(smem);
  Eng<D, NC, R, NT> e(x, c, chain);
  e.sec1 = reinterpret_cast<double*>(base + lay.off_sec1);
  e.sec2 = reinterpret_cast<double*>(base + lay.off_sec2);
  e.mov = reinterpret_cast<double*>(base + lay.off_mov);
  e.fst = (lay.off_fst >= 0) ? reinterpret_cast<double*>(base + lay.off_fst) : x.fnsweep + c.fn_off;
  e.wdet = reinterpret_cast<double*>(base + lay.off_wdet);
  e.werr = reinterpret_cast<int*>(base + lay.off_werr);
  e.wscale = reinterpret_cast<double*>(base + lay.off_wscale);
  e.scale = reinterpret_cast<double*>(base + lay.off_scale);
  e.notmine = reinterpret_cast<int*>(base + lay.off_notmine);
  e.red = reinterpret_cast<double*>(base + lay.off_red);
  e.gjrow = reinterpret_cast<double*>(base + lay.off_gj);
  e.inv = (lay.off_inv >= 0) ? reinterpret_cast<double*>(base + lay.off_inv) : x.inv + c.inv_off;
  e.egr = x.egr + c.egr_off;
  if (c.fnc_off >= 0) e.fcache = x.fncache + c.fnc_off;
  e.top_known = true;
  e.lasterr = e.notmine + PQ_MAX_NODES;
  // per-level try/accept counters, bead updates, accepted, node errors
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base + lay.off_cnt);
  const int tid = e.tid;
  if (tid < 19) cnt[tid] = 0;
  if (tid < c.nfn) e.scale[tid] = x.scale[(int64_t)chain * PQ_MAX_NODES + tid];
  if (lay.off_tab >= 0) {
    // this chain's distinct pair tables (u only) staged in LDS once
    int* tabofs = reinterpret_cast<int*>(base + lay.off_tabofs);
    R* tab = reinterpret_cast<R*>(base + lay.off_tab);
    if (tid == 0) {
      int used = 0;
      for (int pp = 0; pp < c.npair; ++pp) {
        int q = 0;
        while (q < pp && c.pair[q].table != c.pair[pp].table) ++q;
        if (q < pp) { tabofs[pp] = tabofs[q]; continue; }
        tabofs[pp] = used;
        const TableP& T = x.tables[c.pair[pp].table];
        used += T.ngpts * T.K1;
      }
    }
    __syncthreads();
    for (int pp = 0; pp < c.npair; ++pp) {
      int q = 0;
      while (q < pp && c.pair[q].table != c.pair[pp].table) ++q;
      if (q < pp) continue;
      const TableP& T = x.tables[c.pair[pp].table];
      const R* src = table_ptr<R>(x, T);
      for (int k = tid; k < T.ngpts * T.K1; k += NT) tab[tabofs[pp] + k] = src[k];
    }
    e.tab = tab;
    e.tabofs = tabofs;
  }
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
        e.sec_t0 = t0;
        const int ne = e.fn_initialize();
        if (tid == 0) cnt[18] += ne;
      }

      for (int smp = 0; smp < c.nsample; ++smp) {
        int i;
        if (!e.choose(smp, k, i)) continue;
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

template <int D, int NC, class R, int NT>
static hipError_t launch_sweep_nt(const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                  hipStream_t st) {
  SweepLayout lay;
  lay.off_sec1 = L.off_sec1; lay.off_sec2 = L.off_sec2; lay.off_mov = L.off_mov;
  lay.off_fst = L.off_fst; lay.off_wdet = L.off_wdet; lay.off_werr = L.off_werr;
  lay.off_wscale = L.off_wscale; lay.off_scale = L.off_scale; lay.off_notmine = L.off_notmine;
  lay.off_red = L.off_red; lay.off_cnt = L.off_cnt; lay.off_inv = L.off_inv;
  lay.off_gj = L.off_gj;
  lay.off_tab = L.off_tab;
  lay.off_tabofs = L.off_tabofs;
  lay.bytes = L.bytes;
  auto k = sweep_kernel<D, NC, R, NT>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, L.bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(nchain), dim3(NT), L.bytes, st, x, nsweep, lay);
  return hipGetLastError();
}

template <int D, int NC, class R>
static hipError_t launch_sweep_t(const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                 hipStream_t st) {
  if (L.team == 512) return launch_sweep_nt<D, NC, R, 512>(x, nchain, nsweep, L, st);
  if (L.team == 256) return launch_sweep_nt<D, NC, R, 256>(x, nchain, nsweep, L, st);
  if (L.team == 128) return launch_sweep_nt<D, NC, R, 128>(x, nchain, nsweep, L, st);
  return launch_sweep_nt<D, NC, R, PQ_WAVE>(x, nchain, nsweep, L, st);
}

// one translation unit per (D, R) instantiates every node-matrix width
template <int D, class R>
static hipError_t launch_sweep_dr(int NC, const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                  hipStream_t st) {
  switch (NC) {
    case 4: return launch_sweep_t<D, 4, R>(x, nchain, nsweep, L, st);
    case 8: return launch_sweep_t<D, 8, R>(x, nchain, nsweep, L, st);
    case 10: return launch_sweep_t<D, 10, R>(x, nchain, nsweep, L, st);
    case 12: return launch_sweep_t<D, 12, R>(x, nchain, nsweep, L, st);
    case 16: return launch_sweep_t<D, 16, R>(x, nchain, nsweep, L, st);
    default: return launch_sweep_t<D, 32, R>(x, nchain, nsweep, L, st);
  }
}


}  // namespace pq
