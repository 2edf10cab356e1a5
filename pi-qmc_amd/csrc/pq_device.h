// pq_device.h -- device building blocks of the bisection-move hot path.
//
// Execution model: one workgroup (a team of 64 or 256 threads) owns one
// Markov chain.  Every routine is called by all threads of the team with
// team-uniform arguments and returns team-uniform results; threads split the
// independent work (moved beads, links, partner x link pairs, node matrices,
// matrix rows) and combine it with cross-lane and LDS reductions.  Section beads live in LDS for the
// fused sampler and in global memory for the call-level kernels; the routines
// only see plain pointers.
//
// Arithmetic type R is double (FP64 mode) or float (mixed mode).  In mixed
// mode the spring, SHO and pair-table arithmetic and the node-matrix fill run
// in FP32 with compensated per-lane sums; coordinate differences and minimum
// image, the mover, the node LU/inverse/determinant, the sign test and the
// Metropolis exponent stay in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pq_params.h"
#include "../../include/piqmc_hip.h"

namespace pq {

struct DevCtx {
  const ChainP* chains;
  const TableP* tables;
  const double* tab_u;
  const double* tab_ut;
  const float* tab_uf;
  double* beads;
  double* inv;            // per-chain [S][ns_max][ns_max] inverse cache
  double* egr;            // per-chain [S][ns_max][ns_max][D] density-matrix gradient factors
  double* upd;            // per-chain MatrixUpdate blocks of the node actions with useUpdates
  double* fnstate;        // per-chain fixed-node state (call-level API)
  double* fncache;        // per-chain node records of every slice pair (sampler)
  double* fnsweep;        // per-chain fixed-node state of the sampler when not in LDS
  double* scale;          // [nchain][PQ_MAX_NODES]
  uint64_t* counter;      // uniforms consumed per chain
  pq_chain_stats* stats;
  const double* inj;
  int64_t inj_per_chain;
  int rng_mode;
  uint64_t seed;
  uint32_t first_stream;
  const int* order;       // workgroup -> chain
  int* perm;              // per-chain path permutation then its inverse, [2][N]
  const double* ewk;      // Ewald k-vector pool, [nk][4] per sum (EwP)
  const int* plan;        // sweeps per launch of each chain times nsweep, null: 1
};

// ------------------------------------------------------------ wave ops

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// compensated accumulator: plain sum in FP64, Neumaier sum in FP32
template <class R> struct Acc {
  R s = 0;
  __device__ __forceinline__ void add(R v) { s += v; }
  __device__ __forceinline__ double value() const { return (double)s; }
};
template <> struct Acc<float> {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// --------------------------------------------------------------- RNG

// Philox4x32-10; draw k of stream s: counter (k>>2, k>>34, s, 0), key (seed),
// word k&3, u = w * 2^-32.
__device__ __forceinline__ uint32_t philox_word(uint64_t seed, uint32_t stream, uint64_t k) {
  uint32_t c0 = (uint32_t)(k >> 2), c1 = (uint32_t)(k >> 34), c2 = stream, c3 = 0u;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  uint32_t w = c0;
  int sel = (int)(k & 3);
  w = (sel == 1) ? c1 : w;
  w = (sel == 2) ? c2 : w;
  w = (sel == 3) ? c3 : w;
  return w;
}

__device__ __forceinline__ double uniform(const DevCtx& x, int chain, uint64_t k) {
  if (x.rng_mode == PQ_RNG_PHILOX)
    return (double)philox_word(x.seed, x.first_stream + (uint32_t)chain, k) * 2.3283064365386963e-10;
  // past the end of an injected stream: a placeholder; the host reports the
  // overrun as an error once the stream has drained (pq_sync, pq_get_stats)
  return (k < (uint64_t)x.inj_per_chain) ? x.inj[(int64_t)chain * x.inj_per_chain + (int64_t)k] : 0.5;
}

// ------------------------------------------------- SuperCell and PG

// SuperCell::pbc (src/util/SuperCell.cc:25-31)
template <int D>
__device__ __forceinline__ void pbc(const ChainP& c, double* v) {
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

struct PG {
  double nome, pref, k;
  int nmax;
};

__device__ __forceinline__ PG pg_make(double alpha, double length) {
  const double PI = 3.141592653589793;
  PG p;
  p.nome = exp(-PI * PI / (alpha * length * length));
  p.nmax = 1 + (int)(2.0 * sqrt(alpha * length * length));
  p.pref = sqrt(PI / (alpha * length * length));
  p.k = 2 * PI / length;
  return p;
}

// PeriodicGaussian::evaluate (src/util/PeriodicGaussian.cc:20-40): value and gradient
__device__ __forceinline__ double pg_eval(double nome, double pref, double k, int nmax,
                                          double x, double* grad) {
  double sx, cx;
  sincos(k * x, &sx, &cx);
  double s0r = 0, s1r = 0, tr = 1.0, ti = 0.0, nomeN = 1.0;
  for (int n = 1; n < nmax; ++n) {
    double f = nomeN * nomeN * nome;
    double fwr = f * cx, fwi = f * sx;
    double nr = tr * fwr - ti * fwi, ni = tr * fwi + ti * fwr;
    tr = nr; ti = ni;
    s0r += tr;
    s1r += (-ti * k) * (double)n;
    nomeN *= nome;
  }
  if (grad) *grad = pref * 2 * s1r;
  return pref * (1 + 2 * s0r);
}

__device__ __forceinline__ double pg_eval(const PG& p, double x) {
  return pg_eval(p.nome, p.pref, p.k, p.nmax, x, nullptr);
}

// Natural logarithm of a positive normal double to a few ulp: x = m 2^e with
// m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with s = (m - 1)/(m + 1),
// |s| < 0.1716, by the odd series 2 s (1 + s^2/3 + ... + s^20/21), whose
// truncation is below 1e-18 relative.  About a third of the instructions of
// the library's log, which carries a double-double result.
__device__ __forceinline__ double fast_log(double x) {
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
  double p = 1.0 / 21;
  p = fma(p, s2, 1.0 / 19);
  p = fma(p, s2, 1.0 / 17);
  p = fma(p, s2, 1.0 / 15);
  p = fma(p, s2, 1.0 / 13);
  p = fma(p, s2, 1.0 / 11);
  p = fma(p, s2, 1.0 / 9);
  p = fma(p, s2, 1.0 / 7);
  p = fma(p, s2, 1.0 / 5);
  p = fma(p, s2, 1.0 / 3);
  const double t = (s + s) * (s2 * p);   // 2 atanh(s) - 2 s
  const double de = (double)e;
  // ln 2 split so that e ln2_hi is exact for |e| < 2^11
  return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, t) + (s + s));
}

// 1/d for a positive normal double to about an ulp: the hardware estimate and
// two Newton steps
__device__ __forceinline__ double fast_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  return fma(fma(-d, r, 1.0), r, r);
}

// sqrt of a non-negative double to about an ulp (0 gives 0) for arguments
// in the normal range: the hardware reciprocal square root refined by one
// Goldschmidt step and one Newton correction
__device__ __forceinline__ double fast_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  g = fma(fma(-g, g, x), h, g);
  return x > 0.0 ? g : 0.0;
}

// fast transcendental overloads for the arithmetic type
__device__ __forceinline__ double r_log(double v) { return fast_log(v); }
__device__ __forceinline__ float r_log(float v) { return __logf(v); }
__device__ __forceinline__ double r_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float r_sqrt(float v) { return __fsqrt_rn(v); }
__device__ __forceinline__ double r_exp(double v) { return exp(v); }
__device__ __forceinline__ float r_exp(float v) { return __expf(v); }

// exp(-alpha x^2) summed over the three nearest images x, x-L, x+L, with its
// derivative; equal to the periodic Gaussian's theta series (Poisson summation)
// to rounding when alpha L^2 >= 20 and |x| <= L/2.
template <class R>
__device__ __forceinline__ R pg_images(R alpha, R len, R x, R* grad) {
  R v = 0, g = 0;
#pragma unroll
  for (int m = -1; m <= 1; ++m) {
    const R y = x - (R)m * len;
    const R e = r_exp(-alpha * y * y);
    v += e;
    g += (R)-2 * alpha * y * e;
  }
  if (grad) *grad = g;
  return v;
}

// g / (v + tiny) of FreeParticleNodes.cc:201 (tiny = 1e-300); FP32 uses its
// smallest normal number, the nearest representable guard
template <class R> __device__ __forceinline__ R ratio(R g, R v);
template <> __device__ __forceinline__ double ratio<double>(double g, double v) { return g / (v + 1e-300); }
template <> __device__ __forceinline__ float ratio<float>(float g, float v) {
  return g * __frcp_rn(v + 1.17549435e-38f);
}

// copy a chain's parameter block into LDS
__device__ __forceinline__ void copy_chain(ChainP* dst, const ChainP* src) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (int k = threadIdx.x; k < (int)(sizeof(ChainP) / 4); k += blockDim.x) d[k] = s[k];
  __syncthreads();
}

__device__ __forceinline__ int species_of(const ChainP& c, int i) {
  int s = 0;
  for (int k = 1; k < c.nsp; ++k) s = (i >= c.ifirst[k]) ? k : s;
  return s;
}

// (cos, sin) of the phase k.r of one particle for the k-vector kv
template <int D>
__device__ __forceinline__ void ewald_phase(const double* kv, const double* r, double& co, double& si) {
  double ph = kv[0] * r[0] + kv[1] * r[1];
  if constexpr (D == 3) ph += kv[2] * r[2];
  sincos(ph, &si, &co);
}

// EwaldSum::evalLongRange (src/util/EwaldSum.cc:114-178) of one slice
// sl [N][D] by one thread: sum over the half-space k-vectors of
// 2 f(k)/(2V) |sum_j q_j e^{i k.r_j}|^2, plus the self energy
template <int D>
__device__ double ewald_slice(const double* kp, const EwP& ew, const ChainP& c, const double* sl) {
  double sum = 0;
  for (int k = 0; k < ew.nk; ++k) {
    const double* kv = kp + 4 * k;
    double rr = 0, ri = 0;
    for (int j = 0; j < c.N; ++j) {
      double co, si;
      ewald_phase<D>(kv, sl + j * D, co, si);
      const double q = c.charge[species_of(c, j)];
      rr += q * co;
      ri += q * si;
    }
    sum += kv[3] * (rr * rr + ri * ri);
  }
  return sum + ew.self;
}

template <class R> __device__ __forceinline__ const R* table_ptr(const DevCtx& x, const TableP& t);
template <> __device__ __forceinline__ const double* table_ptr<double>(const DevCtx& x, const TableP& t) {
  return x.tab_u + t.off;
}
template <> __device__ __forceinline__ const float* table_ptr<float>(const DevCtx& x, const TableP& t) {
  return x.tab_uf + t.off;
}

// PairAction::u00 / uk0 (src/action/PairAction.cc:601-621); K1=1 or the
// diagonal-only branch give u00.
template <class R>
__device__ __forceinline__ R pair_u(const TableP& t, const R* u, R q, R s2, bool offdiag) {
  q *= (R)t.rgridinv;
  R xx = r_log(q) * (R)t.logrratioinv;
  int i = (int)xx;
  if (q < (R)1) { i = 0; xx = 0; }
  else if (i > t.ngpts - 2) { i = t.ngpts - 2; xx = 1; }
  else xx -= (R)i;
  const R* u0 = u + (int64_t)i * t.K1;
  const R* u1 = u0 + t.K1;
  if (!offdiag) return ((R)1 - xx) * u0[0] + xx * u1[0];
  // both grid points' coefficients loaded at once (the tables may be in
  // global memory), then the Horner sum in the same order
  constexpr int KM = PQ_MAX_NORDER + 1;
  R c0[KM], c1[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    c0[k] = k < t.K1 ? u0[k] : (R)0;
    c1[k] = k < t.K1 ? u1[k] : (R)0;
  }
  R a = 0;
#pragma unroll
  for (int k = KM - 1; k >= 0; --k)
    if (k < t.K1) {
      a *= s2;
      a += ((R)1 - xx) * c0[k] + xx * c1[k];
    }
  return a;
}

}  // namespace pq
