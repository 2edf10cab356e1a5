// pq_sweep.h -- the device-resident sampler kernel, included by one
// translation unit per (dimension, arithmetic type) so the instantiations
// compile in parallel.
#pragma once
#include "pq_engine.h"
#include "pq_launch.h"

namespace pq {

// the LDS carve-up of the sampler workgroup is SweepLayoutH (pq_launch.h),
// passed to the kernel as it is

// SerialPaths::putBeads (SerialPaths.cc:144-185) of section slices k0..k1-1:
// slices before 0 are skipped, slices from M on go to particle perm[i].
template <int D, int NT>
__device__ __forceinline__ void put_slices(double* beads, const double* sec, const ChainP& c, int t0, int k0, int k1,
                           const int* perm) {
  const int slice = c.N * D;
  for (int q = k0 * slice + (int)threadIdx.x; q < k1 * slice; q += NT) {
    const int kk = q / slice;
    const int rem = q - kk * slice;
    const int t = kk + t0;
    if (t < 0) continue;
    if (t < c.M) {
      beads[(int64_t)t * slice + rem] = sec[q];
    } else {
      const int i = rem / D, d = rem - i * D;
      beads[(int64_t)(t - c.M) * slice + perm[i] * D + d] = sec[q];
    }
  }
}

// the rest of putBeads for a non-identity section permutation sp: slices
// ifollowing..M-1 are relabelled, bead(i,t) <- bead(temp[i],t), with temp = sp,
// or perm o sp o perm^-1 when the section ran past M (Permutation::prepend /
// append, Permutation.cc:23-43), staged through buf (S slices at a time); then
// perm <- perm o temp and its inverse.  tmp and tmp2 are N-int scratch.
template <int D, int NT>
__device__ __forceinline__ void relabel_following(double* beads, double* buf, const ChainP& c, int t0, const int* sp,
                                  int* perm, int* invperm, int* tmp, int* tmp2) {
  const int N = c.N, M = c.M, S = c.S;
  const int slice = N * D;
  int ifol = t0 + S;
  const bool wrapped = ifol > M;
  if (wrapped) ifol -= M;
  if (threadIdx.x == 0) {
    for (int i = 0; i < N; ++i) tmp[i] = wrapped ? perm[sp[invperm[i]]] : sp[i];
  }
  __syncthreads();
  for (int tc = ifol; tc < M; tc += S) {
    const int nsl = min(S, M - tc);
    for (int q = threadIdx.x; q < nsl * slice; q += NT) {
      const int tt = q / slice;
      const int rem = q - tt * slice;
      const int i = rem / D, d = rem - i * D;
      buf[q] = beads[(int64_t)(tc + tt) * slice + tmp[i] * D + d];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nsl * slice; q += NT) beads[(int64_t)tc * slice + q] = buf[q];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < N; ++i) tmp2[i] = perm[tmp[i]];
    for (int i = 0; i < N; ++i) { perm[i] = tmp2[i]; invperm[tmp2[i]] = i; }
  }
  __syncthreads();
}

// two waves per SIMD whatever the team: up to 256 threads two workgroups share
// a CU, a 512-thread team has one to itself.  MM bounds the moving particles
// per Sample: MM = 1 never permutes and carries no relabelling code.
// All sweeps of one chain by the calling workgroup.
template <int D, int NC, class R, int NT, int MM>
__device__ __forceinline__ void sweep_chain(const DevCtx& x, int chain, int nsweep, const SweepLayoutH& lay) {
  extern __shared__ double smem[];
  const uint64_t t_begin = wall_clock64();
  ChainP& c = *reinterpret_cast<ChainP*>(smem);
  copy_chain(&c, &x.chains[chain]);
  char* base = reinterpret_cast<char*>(smem);
  Eng<D, NC, R, NT, MM> e(x, c, chain);
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
  static_assert(sizeof(typename Eng<D, NC, R, NT, MM>::ChooseBuf) <= 512, "chooser scratch outgrew its LDS slot");
  e.cbuf = reinterpret_cast<typename Eng<D, NC, R, NT, MM>::ChooseBuf*>(base + lay.off_choose);
  if (lay.off_walk >= 0) e.wrows = reinterpret_cast<double*>(base + lay.off_walk);
  // path permutation and inverse, section permutation, scratch
  int* perm = reinterpret_cast<int*>(base + lay.off_perm);
  int* invperm = perm + c.N;
  int* secperm = invperm + c.N;
  int* ptmp = secperm + c.N;
  int* gperm = x.perm + c.perm_off;
  for (int q = threadIdx.x; q < 2 * c.N; q += NT) perm[q] = gperm[q];
  bool perm_changed = false;
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

  const int N = c.N, M = c.M;
  const int64_t slice = (int64_t)N * D;
  double* beads = x.beads + c.bead_off;
  uint64_t k = x.counter[chain];

  const int nsw = x.plan ? nsweep * x.plan[chain] : nsweep;
  for (int sw = 0; sw < nsw; ++sw)
  for (int g = 0; g < c.ngroup; ++g) {
    // the sweep's g-th <ChooseSection nlevel> with its Samples: its level
    // count, section length and Samples take the chain parameters' place
    if (c.ngroup > 1) {
      __syncthreads();
      if (tid == 0) {
        const GroupP& G = c.grp[g];
        c.L = G.L;
        c.S = (1 << G.L) + 1;
        c.nsection = G.nsection;
        c.nsample = G.nsample;
        for (int q = 0; q < G.nsample; ++q) {
          c.sample_sp[q] = G.sample_sp[q];
          c.sample_ch[q] = G.sample_ch[q];
          c.sample_nm[q] = G.sample_nm[q];
          c.sample_rep[q] = G.sample_rep[q];
          if constexpr (MM > 1) {
            c.sample_dr[q] = G.sample_dr[q];
            c.sample_fac[q] = G.sample_fac[q];
            c.sample_nfac[q] = G.sample_nfac[q];
          }
        }
      }
      __syncthreads();
    }
    const int S = c.S, L = c.L;
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
      // SerialPaths::getBeads (SerialPaths.cc:119-132): slices before 0 through
      // the inverse permutation, slices from M on through the permutation.
      // Sections inside [0, M) are contiguous runs of S slices: copied with
      // four loads of each section in flight per thread before their stores
      // (a load-store loop waits out one memory latency per element).
      const int nsec = S * (int)slice;
      if (t0 >= 0 && t0 + S <= M && (!c.nfn || (t0b >= 0 && t0b + S <= M))) {
        const double* s1 = beads + (int64_t)t0 * slice;
        const double* s2 = beads + (int64_t)t0b * slice;
        for (int q0 = tid; q0 < nsec; q0 += 4 * NT) {
          double v1[4], v2[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * NT;
            v1[u] = q < nsec ? s1[q] : 0.0;
            v2[u] = (c.nfn && q < nsec) ? s2[q] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * NT;
            if (q < nsec) {
              e.sec1[q] = v1[u];
              if (c.nfn) e.sec2[q] = v2[u];
            }
          }
        }
      } else
      for (int q = tid; q < nsec; q += NT) {
        const int kk = q / (int)slice;
        const int rem = q - kk * (int)slice;
        const int i = rem / D, d = rem - i * D;
        int t = kk + t0, src = i;
        if (t < 0) { t += M; src = invperm[i]; }
        else if (t >= M) { t -= M; src = perm[i]; }
        e.sec1[q] = beads[t * slice + src * D + d];
        if (c.nfn) {
          int tb = kk + t0b, srcb = i;
          if (tb < 0) { tb += M; srcb = invperm[i]; }
          else if (tb >= M) { tb -= M; srcb = perm[i]; }
          e.sec2[q] = beads[tb * slice + srcb * D + d];
        }
      }
      for (int q = tid; q < N; q += NT) secperm[q] = q;
      __syncthreads();
      bool sec_permuted = false;
      if (c.nfn) {
        e.sec_t0 = t0;
        bool pid = true;
        for (int q = 0; q < N; ++q) pid &= perm[q] == q;
        e.rec_on = pid || (t0 >= 0 && t0 + S <= M && t0b >= 0 && t0b + S <= M);
        const int ne = e.fn_initialize();
        if (tid == 0) cnt[18] += ne;
      }

      // every Sample tries nrepeat moves on the section (MultiLevelSampler::run,
      // MultiLevelSampler.cc:68-98); the choosers read the current section
      for (int smp = 0; smp < c.nsample; ++smp)
      for (int rep = 0; rep < c.sample_rep[smp]; ++rep) {
        double lnTP;
        if (!e.choose(smp, k, lnTP)) continue;
        const int nmov = e.nmov();
        const int* midx = e.midx;
        const int* mcyc = e.cbuf->cyc;
        // the moving beads start on the chosen particles' first beads and end
        // on the permuted particles' last beads (MultiLevelSampler.cc:74-87)
        bool permuted = false;
        if constexpr (MM > 1)
          for (int m = 0; m < nmov; ++m) permuted |= mcyc[m] != m;
        e.top_known = !permuted;
        for (int q = tid; q < nmov * D; q += NT) {
          const int m = q / D, d = q - m * D;
          e.mov[m * D + d] = e.sb(e.sec1, 0, e.mv(m))[d];
          e.mov[((S - 1) * nmov + m) * D + d] = e.sb(e.sec1, S - 1, e.mv(MM > 1 ? mcyc[m] : m))[d];
        }
        __syncthreads();
        // MultiLevelSampler::tryMove (MultiLevelSampler.cc:100-177).  Every
        // level's beads are proposed first: a level's mover only reads the
        // coarser levels' new beads, and its draws start after the draws and
        // the accept draw of the levels above, wherever the move stops.  The
        // node determinants of every changed slice then take one pass.
        // The general instances (MM > 1) also run the movers' sigma factors
        // and delayed rejection (MultiLevelSampler.cc:114-143,
        // DoubleMLSampler.cc:150-206): the beads B rejected at level L-1 are
        // kept and the move is tried once more, C at level L-1 with the
        // second sigma factor, its draws after B's accept draw, and the
        // levels below proposed again after C's accept draw.
        double old = lnTP;
        bool accepted = true;
        if constexpr (MM == 1) {
          {
            // the first wave proposes every level; the team synchronises once
            if (e.wave == 0) e.mover_all(k, e.cbuf->lnT);
            __syncthreads();
            if (c.nfn) e.fn_eval_moved();
          }
          for (int l = L; l >= 0; --l) {
            if (tid == 0) {
              cnt[l]++;
              if (l < L) cnt[16] += (unsigned long long)nmov << (L - 1 - l);
            }
            k += e.mover_draws(l);
            double dA = e.action_diff(l);
            if (c.nfn) dA += e.fn_diff_moved(l);
            const double piRatio = -dA + old;
            const double acceptProb = exp(e.cbuf->lnT[l] + piRatio);
            if (e.U(k++) > acceptProb) { accepted = false; break; }
            old = dA;
            if (tid == 0) cnt[8 + l]++;
          }
        } else {
          const bool dr = c.sample_dr[smp] && L >= 1;
          int lfrom = L;     // below L: the delayed second proposal from level lfrom
          for (;;) {
            {
              uint64_t kk = lfrom < L ? e.cbuf->kC : k;
              for (int l = lfrom; l >= 0; --l) {
                int nd;
                e.fac = (lfrom < L && l == lfrom) ? c.sample_nfac[smp] : c.sample_fac[smp];
                const double lnT = e.mover(l, kk, &nd, dr && lfrom == L && l == L - 1 ? &e.cbuf->transA2B : nullptr);
                kk += nd + 1;
                if (tid == 0) e.cbuf->lnT[l] = lnT;
                __syncthreads();
              }
              e.fac = c.sample_fac[smp];
              if (c.nfn) e.fn_eval_moved();
            }
            const int lstart = lfrom;
            for (int l = lstart; l >= 0; --l) {
              const bool second = lstart < L && l == lstart;
              if (tid == 0) {
                if (!second) cnt[l]++;
                if (l < L) cnt[16] += (unsigned long long)nmov << (L - 1 - l);
              }
              k += e.mover_draws(l);
              double dA = e.action_diff(l);
              if (c.nfn) dA += e.fn_diff_moved(l);
              if (second) {
                const double transA2B = e.cbuf->transA2B, accA2B = e.cbuf->accA2B;
                const double piCoA = -dA + old;
                const double accA2C = exp(e.cbuf->lnT[l] + piCoA);
                double transC2B;
                const double lnTC2B = e.delayed_mover(l, e.cbuf->rej, &transC2B);
                double accC2B = exp(lnTC2B - piCoA + e.cbuf->piBoA);
                accC2B = (accC2B >= 1) ? 1 : accC2B;
                const double accA2B2C = accA2C * exp(transC2B + transA2B) * (1 - accC2B) / (1 - accA2B);
                dA = dA - transC2B - transA2B - log((1 - accC2B) / (1 - accA2B));
                if (e.U(k++) > accA2B2C) { accepted = false; break; }
              } else {
                const double piRatio = -dA + old;
                const double acceptProb = exp(e.cbuf->lnT[l] + piRatio);
                if (e.U(k++) > acceptProb) {
                  if (dr && l == L - 1) {
                    const int tm = 1 << l;
                    for (int q = tid; q < nmov * D; q += NT) e.cbuf->rej[q] = e.mov[tm * nmov * D + q];
                    if (tid == 0) {
                      e.cbuf->accA2B = acceptProb;
                      e.cbuf->piBoA = piRatio;
                      e.cbuf->kC = k;
                    }
                    __syncthreads();
                    lfrom = l;
                  } else {
                    accepted = false;
                  }
                  break;
                }
              }
              old = dA;
              if (tid == 0) cnt[8 + l]++;
            }
            if (!accepted || lfrom == lstart) break;
          }
        }
        if (accepted) {
          if (c.nfn) e.fn_accept(permuted);
          for (int q = tid; q < S * nmov * D; q += NT) {
            const int t = q / (nmov * D), rem = q - t * nmov * D;
            const int m = rem / D, d = rem - m * D;
            e.sb(e.sec1, t, e.mv(m))[d] = e.mov[q];
          }
          if (tid == 0) {
            cnt[17] += (unsigned long long)nmov * ((1ull << L) - 1);
            // append the permutation to the section permutation (MultiLevelSampler.cc:168-176)
            if (MM > 1 && permuted) {
              for (int m = 0; m < nmov; ++m) ptmp[m] = secperm[midx[m]];
              for (int m = 0; m < nmov; ++m) secperm[midx[m]] = ptmp[mcyc[m]];
            }
          }
          sec_permuted |= permuted;
          __syncthreads();
        }
      }
      // SerialPaths::putBeads (SerialPaths.cc:144-185) / putDoubleBeads (:187-206).
      // Slice 0 is never moved; the last slice only under a permutation.
      bool ident = true;
      if (MM > 1 && sec_permuted)
        for (int q = 0; q < N; ++q) ident &= secperm[q] == q;
      put_slices<D, NT>(beads, e.sec1, c, t0, 1, ident ? S - 1 : S, perm);
      __syncthreads();
      if constexpr (MM > 1) if (!ident) {
        relabel_following<D, NT>(beads, e.sec1, c, t0, secperm, perm, invperm, ptmp, ptmp + N);
        if (c.nfn && t0 < M / 2 - 1) {
          // section 2 comes later in the path: its beads take the section
          // permutation (Beads::permute, Beads.h:254-261) and go back with
          // the identity permutation (p1 p2 p1^-1 with p2 the identity)
          for (int q = tid; q < S * (int)slice; q += NT) {
            const int kk = q / (int)slice;
            const int rem = q - kk * (int)slice;
            const int i = rem / D, d = rem - i * D;
            e.sec1[q] = e.sec2[kk * (int)slice + secperm[i] * D + d];
          }
          __syncthreads();
          put_slices<D, NT>(beads, e.sec1, c, t0b, 0, S, perm);
          __syncthreads();
        }
        // the slices after the section now carry other labels: every node
        // record of the chain is stale
        if (e.fcache) {
          double* rec = e.fcache;
          for (int fa = 0; fa < c.nfn; ++fa) {
            const int rd = fn_record_doubles(c.fn[fa].ns);
            for (int q = tid; q < M / 2; q += NT) rec[(int64_t)q * rd + 2] = 0.0;
            rec += (int64_t)(M / 2) * rd;
          }
        }
        perm_changed = true;
        __syncthreads();
      }
    }
  }
  __syncthreads();
  if (MM > 1 && perm_changed)
    for (int q = threadIdx.x; q < 2 * c.N; q += NT) gperm[q] = perm[q];
  if (tid == 0) {
    x.counter[chain] = k;
    pq_chain_stats& st = x.stats[chain];
    st.draws = k;
    st.bead_updates += cnt[16];
    st.accepted_bead_updates += cnt[17];
    st.node_errors += cnt[18];
    for (int l = 0; l < 8; ++l) { st.tries[l] += cnt[l]; st.accepts[l] += cnt[8 + l]; }
    st.sweeps += nsw;
    st.t_start = t_begin;
    st.t_end = wall_clock64();
  }
  if (tid < c.nfn) x.scale[(int64_t)chain * PQ_MAX_NODES + tid] = e.scale[tid];
}

// One chain per workgroup; or (Q) a grid of at most the resident
// workgroups in which a workgroup that finishes a chain takes the next one of
// the launch's order (heaviest first) from the queue counter lay.queue
// (zeroed before the launch) until the lay.nwork chains are taken, so the
// launch keeps its slots until its queue is empty.  Q has its own instances:
// the loop around the chain's sweeps costs registers.
template <int D, int NC, class R, int NT, int MM, bool Q = false>
__global__ __launch_bounds__(NT, NT <= 256 ? 2 : 1) void sweep_kernel(DevCtx x, int nsweep, SweepLayoutH lay) {
  if constexpr (!Q) {
    sweep_chain<D, NC, R, NT, MM>(x, x.order[blockIdx.x], nsweep, lay);
    return;
  }
  __shared__ int next;
  int wi = blockIdx.x;
  for (;;) {
    sweep_chain<D, NC, R, NT, MM>(x, x.order[wi], nsweep, lay);
    __syncthreads();
    if (threadIdx.x == 0) next = (int)gridDim.x + atomicAdd(lay.queue, 1);
    __syncthreads();
    wi = next;
    if (wi >= lay.nwork) return;
  }
}

template <int D, int NC, class R, int NT, int MM, bool Q = false>
static hipError_t launch_sweep_nt(const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                  hipStream_t st) {
  auto k = sweep_kernel<D, NC, R, NT, MM, Q>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, L.bytes);
  if (e != hipSuccess) return e;
  int grid = nchain;
  if (Q) {
    // a queued launch: as many workgroups as can be resident at once
    int per_cu = 0, dev = 0, ncu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, NT, L.bytes);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (per_cu > 0 && ncu > 0 && per_cu * ncu < nchain) grid = per_cu * ncu;
  }
  SweepLayoutH Lq = L;
  Lq.nwork = nchain;
  hipLaunchKernelGGL(k, dim3(grid), dim3(NT), L.bytes, st, x, nsweep, Lq);
  return hipGetLastError();
}

// batches whose Samples all move one particle take the MM = 1 instances at
// every team size; batches with cycles of several particles take the general
// instances, built for one- and two-wave teams
template <int D, int NC, class R>
static hipError_t launch_sweep_t(bool multi, const DevCtx& x, int nchain, int nsweep, const SweepLayoutH& L,
                                 hipStream_t st) {
  if (multi) {
    if (L.team == 128) return launch_sweep_nt<D, NC, R, 128, PQ_MAX_MOVING>(x, nchain, nsweep, L, st);
    if (L.team == PQ_WAVE) return launch_sweep_nt<D, NC, R, PQ_WAVE, PQ_MAX_MOVING>(x, nchain, nsweep, L, st);
    return hipErrorInvalidValue;
  }
  // queued launches: the heavy chains' teams of 256 or 512 threads
  if (L.queue && L.team == 512) return launch_sweep_nt<D, NC, R, 512, 1, true>(x, nchain, nsweep, L, st);
  if (L.queue && L.team == 256) return launch_sweep_nt<D, NC, R, 256, 1, true>(x, nchain, nsweep, L, st);
  if (L.team == 512) return launch_sweep_nt<D, NC, R, 512, 1>(x, nchain, nsweep, L, st);
  if (L.team == 256) return launch_sweep_nt<D, NC, R, 256, 1>(x, nchain, nsweep, L, st);
  if (L.team == 128) return launch_sweep_nt<D, NC, R, 128, 1>(x, nchain, nsweep, L, st);
  return launch_sweep_nt<D, NC, R, PQ_WAVE, 1>(x, nchain, nsweep, L, st);
}

// one translation unit per (D, R) instantiates every node-matrix width
template <int D, class R>
static hipError_t launch_sweep_dr(int NC, bool multi, const DevCtx& x, int nchain, int nsweep,
                                  const SweepLayoutH& L, hipStream_t st) {
  switch (NC) {
    case 4: return launch_sweep_t<D, 4, R>(multi, x, nchain, nsweep, L, st);
    case 8: return launch_sweep_t<D, 8, R>(multi, x, nchain, nsweep, L, st);
    case 10: return launch_sweep_t<D, 10, R>(multi, x, nchain, nsweep, L, st);
    case 12: return launch_sweep_t<D, 12, R>(multi, x, nchain, nsweep, L, st);
    case 16: return launch_sweep_t<D, 16, R>(multi, x, nchain, nsweep, L, st);
    default: return launch_sweep_t<D, 32, R>(multi, x, nchain, nsweep, L, st);
  }
}


}  // namespace pq
