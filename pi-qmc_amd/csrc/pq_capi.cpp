// pq_capi.cpp -- the C ABI of include/piqmc_hip.h: chain setup, Coulomb pair
// tables, device memory, and the launches of the gfx950 kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/piqmc_hip.h"
#include "pq_ewald.h"
#include "pq_launch.h"
#include "pq_params.h"

using namespace pq;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(PQ_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                          \
  } while (0)

extern "C" const char* pq_last_error(void) { return g_err.c_str(); }

// failures of the host-only parts in other files (the XML reader)
void pq_set_last_error(const char* msg) { g_err = msg; }

// ------------------------------------------------------- Coulomb tables
// Coulomb1DLinkAction (src/action/coulomb/Coulomb1DLinkAction.cpp:4-133) and
// Coulomb3DLinkAction (src/action/coulomb/Coulomb3DLinkAction.cpp:13-90):
// the published rational fits of the 1D Coulomb link action and the 3D
// orders built from them.  The coefficients are data and must match the
// reference exactly.
namespace {
struct Fit1D {
  double s, m1, m2, m3, m4, m5;
  explicit Fit1D(double stau) : s(stau) {
    m1 = 1.0 / stau; m2 = m1 * m1; m3 = m2 * m1; m4 = m3 * m1; m5 = m4 * m1;
  }
  double origin() const {
    return s * (1.772453851 + s * (-0.074137740 + s * (0.005834805 + s * (-0.000382686
           + s * (0.000008738 + s * 0.000002138)))));
  }
  double u0(double uo, double r) const {
    const double a = 0.25300593 * m1 + 0.01432126, b = 0.07936898 * m2 - 0.01634421 / s;
    const double c = 0.07263383 * m3, d = 0.00940013 * m4, e = 0.03160181 * m5;
    const double f = 0.18976335 * m1, g = 0.00343053 * m2;
    double v = uo + r * ((uo * a - 1.) + r * (f + r * (g + r * e * s * s)));
    return v / (1. + r * (a + r * (b + r * (c + r * (d + r * e)))));
  }
  double u1(double r) const {
    const double a = 0.07776906 * m1, b = -0.04695226 * m2, c = 0.01291994 * m3;
    const double d = 0.60942687 * m1, e = -1.09588064 * m2, f = 1.15591565 * m3;
    const double g = -0.57485278 * m4, h = 0.15212765 * m5;
    return r * r * (a + r * (b + r * c)) / (1 + r * (d + r * (e + r * (f + r * (g + r * h)))));
  }
  double u2(double r) const {
    const double a = 9.52130652e-04 * m3, b = -6.39946071e-04 * m4, c = 1.14359558e-04 * m5;
    const double d = -7.59575215e-01 * m3, e = 5.53798980e-01 * m4, f = -7.15626226e-02 * m5;
    const double g = -3.49081618e-02 * m5 * m1, h = 8.40084881e-03 * m5 * m2;
    return std::pow(r, 4) * (a + r * (b + r * c))
           / (1 + std::pow(r, 3) * (d + r * (e + r * (f + r * (g + r * h)))));
  }
  double u3(double r) const {
    const double a = -2.06984256e-1 * m5, b = -9.51423947e-3 * m5 * m1, c = 8.97532561e-2 * m5 * m2;
    const double d = 6.10876797e+4 * m3, e = -1.19016292e+5 * m4, f = 9.29836599e+4 * m5;
    const double g = -3.49919130e+4 * m5 * m1, h = 6.42490539e+3 * m5 * m2;
    const double i = -6.02729064e+2 * m5 * m3, j = 6.01285550e+1 * m5 * m4;
    return std::pow(r, 6) * (a + r * (b + r * c))
           / (1 + std::pow(r, 3) * (d + r * (e + r * (f + r * (g + r * (h + r * (i + r * j)))))));
  }
  double u4(double r) const {
    const double a = -2.84102296e-6 * m5, b = 6.26961672e-7 * m5 * m2, c = -2.20166473e-1 * m3;
    const double d = 1.07981903e-1 * m5, e = -1.94671111e-2 * m5 * m2, f = 1.56217930e-3 * m5 * m4;
    return std::pow(r, 6) * (a + b * r * r)
           / (1 + r * r * r * (c + r * r * (d + r * r * (e + r * r * f))));
  }
};

double coulomb3d(const Fit1D& F, int order, double r) {
  const double s = F.s;
  switch (order) {
    case 0: {
      const double a = 0.5 * r * r / (s * s), b = std::exp(-2. * a);
      const double c = 1.0 / (1.0 + 2 * (1.0 - b) * F.u1(r) / (a + 1e-200));
      return F.u0(F.origin(), r) + std::log(c);
    }
    case 1: {
      const double w1 = F.u1(r), w2 = F.u2(r);
      const double a = 0.5 * r * r / (s * s), b = std::exp(-2. * a);
      const double c = 1.0 / (1.0 + 2 * (1.0 - b) * w1 / (a + 1e-200));
      return w1 + c * (b * w1 - 4 * (1 - b) * w2 / (a + 1e-200));
    }
    case 2: {
      const double w1 = F.u1(r), w2 = F.u2(r), w3 = F.u3(r);
      const double a = 2. * s * s / (r * r + 1e-50), b = std::exp(-2. / (a + 1e-200));
      const double c = 1. / (1 + 2 * a * (1 - b) * w1);
      return w2 + 0.25 * c / (a + 1e-200)
             * (c * (b * w1 * (1 + 2 * a * w1) + 8 * a * b * w2 + 32 * a * a * a * (1 - b) * (1 - b) * w2 * w2)
                - 24 * a * a * (1 - b) * w3);
    }
    case 3: {
      const double u1 = F.u1(r), u2 = F.u2(r), u3 = F.u3(r), u4 = F.u4(r);
      const double a = 2. * s * s / (r * r + 1e-50), b = std::exp(-2. / (a + 1e-200));
      const double c = 1. / (1 + 2 * a * (1 - b) * u1);
      const double A = 1 + 2 * a * u1;
      return u3 + c * c * c / (24. * a * a + 1e-200)
        * (128 * a * a * a * a * a * b * b * b * (4 * u2 * u2 * u2 - 9 * u1 * u2 * u3 + 6 * u1 * u1 * u4)
           + 64 * a * a * a * (a * u2 * (-8 * a * u2 * u2 + 9 * A * u3) - 3 * A * A * u4)
           + 2 * a * b * b * (2 * a * u1 * u1 * u1
                              + 96 * a * a * u2 * (u2 - 8 * a * a * u2 * u2 + 3 * a * u3)
                              + 12 * a * u1 * (u2 - 6 * a * u3 + 144 * a * a * a * u2 * u3 - 32 * a * a * u4)
                              + u1 * u1 * (1 - 1152 * a * a * a * a * u4))
           + b * (4 * a * a * u1 * u1 * u1
                  + 12 * a * (u2 - 16 * a * a * u2 * u2 + 128 * a * a * a * a * u2 * u2 * u2 + 6 * a * u3
                              - 96 * a * a * a * u2 * u3 + 16 * a * a * u4)
                  + 4 * u1 * u1 * (a + 576 * a * a * a * a * a * u4)
                  + u1 * (1 + 24 * a * a * (u2 + 6 * a * u3 - 144 * a * a * a * u2 * u3 + 64 * a * a * u4))));
    }
    default:
      return F.u4(r);
  }
}

// CoulombAction::u and ::utau (src/action/CoulombAction.cc:209-259); with an
// Ewald sum u_0 gives up tau q1 q2 f(r), which the k-space sum restores
double coulomb_u(double r, int order, double mu, double q1q2, double tau, const EwaldFit* ew) {
  r = std::sqrt(r * r + 0.0);
  const double reff = 2.0 * mu * q1q2 * r;
  Fit1D F(q1q2 * std::sqrt(2.0 * mu * tau));
  double u = coulomb3d(F, order, reff);
  if (order == 0 && ew) u -= tau * q1q2 * ew->fr(r);
  return u;
}

double coulomb_utau(double r, int order, double mu, double q1q2, double tau, const EwaldFit* ew) {
  const double eps = 1e-7;
  const double up = coulomb_u(r, order, mu, q1q2, tau * (1.0 + eps), ew);
  const double um = coulomb_u(r, order, mu, q1q2, tau * (1.0 - eps), ew);
  return (up - um) / (2.0 * eps * tau);
}

int pg_nterms(double alpha, double length) { return 1 + (int)(2.0 * std::sqrt(alpha * length * length)); }
}  // namespace

// ------------------------------------------------------------- the batch

// requested sampler schedule: -1 automatic, else team size or 0/1 placement
struct Sched {
  int team = -1, inv = -1, tab = -1, fst = -1;
  int heavy = 0;          // the heaviest chains launched apart with a larger team
  int heavy_team = 256;
  int ncsplit = 0;        // split at the node-width boundary instead: the chains with
                          // 17-32-particle node species apart, the rest on a narrower
                          // Gauss-Jordan instance with a layout of their own
};

struct pq_batch {
  int device = 0;
  int precision = 0;
  int nchain = 0;
  int D = 3;
  int NC = 4;     // node-matrix width class of the batch: 4, 8, 10, 12, 16 or 32
  int NC_bulk = 0; // with 32-wide chains among narrower ones: the narrower chains' class
  int wide = 0;    // ... and the number of 32-wide chains, first in the launch order
  bool multi = false;   // some Sample moves several particles (cycles, relabelling)
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // the heavy chains' launch, forked from and joined to stream
  hipEvent_t evf = nullptr, evj = nullptr;
  std::vector<pq_chain_desc> descs;
  std::vector<ChainP> hc;
  std::vector<TableP> ht;
  std::vector<double> hu, hut;
  // Ewald sums of the chains' CoulombActions: fits and the k pool [nk][4]
  std::vector<EwaldFit> fits;
  std::map<std::tuple<int, double, double, double, double, double, int, double>, int> fit_of;
  std::vector<int64_t> fit_off;
  std::vector<double> hk;
  double* dewk = nullptr;
  // the EwaldCoulombEstimator's sums, kept for repeated (rcut, kcut, epsilon)
  double est_rcut = -1, est_kcut = -1, est_kappa = -1;
  int est_type = -1;
  EwP* dew_est = nullptr;
  double* dewk_est = nullptr;
  int64_t* dvoff_est = nullptr;   // tradEwald: offset of each chain's -f(i dr) grid, -1 for optEwald
  double* dvgrid_est = nullptr;
  std::vector<int> order;
  int64_t nbeads = 0;
  SweepLayoutH lay{};
  SweepLayoutH lay_heavy{};
  Sched sched{};   // requested schedule, all automatic
  // device
  ChainP* dchains = nullptr;
  TableP* dtables = nullptr;
  double *du = nullptr, *dut = nullptr;
  float* duf = nullptr;
  double* dbeads = nullptr;
  double* dinv = nullptr;
  double* degr = nullptr;
  double* dupd = nullptr;   // MatrixUpdate blocks
  double* dfn = nullptr;
  double* dfncache = nullptr;   // slice-pair node records of the sampler
  double* dfnsweep = nullptr;   // sampler fixed-node state when it does not fit LDS
  int64_t nfncache = 0;
  double* dscale = nullptr;
  uint64_t* dcounter = nullptr;
  pq_chain_stats* dstats = nullptr;
  double* dinj = nullptr;
  int64_t inj_per_chain = 0;
  int* dorder = nullptr;
  int* dplan = nullptr;         // sweeps per launch of each chain, null: one
  int* dperm = nullptr;         // per-chain path permutation and its inverse
  int64_t nperm = 0;
  int rng_mode = PQ_RNG_PHILOX;
  uint64_t seed = 0;
  uint32_t first_stream = 0;
  // call-level scratch
  double* dscr = nullptr;
  uint64_t* dqueue = nullptr;   // queue counter of the heavy chains' launch
  size_t scr_bytes = 0;
  double* dfnscr = nullptr;   // fixed-node protocol buffers, kept between calls
  size_t fnscr_bytes = 0;
  double* dbead_stage = nullptr; // caller paths of pq_bead_action, then their permutation
  size_t bead_stage_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // measured-sweep graph: nblock x (sweeps, energy) captured once, replayed
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;        // what the graph was captured with
  double* dmeas = nullptr;       // [nblock][nchain] energies written by the graph
  size_t nmeas = 0;
  std::vector<hipEvent_t> gev;   // per block: before and after the sweep graph's launch (timed runs)
  // the sweeps of one block alone, captured for timed runs
  hipGraph_t sgraph = nullptr;
  hipGraphExec_t sexec = nullptr;
  std::vector<char> skey;
  // Measure work items (slices of one chain): energy items with their
  // per-chain ranges and partial sums, and the histogram's coarser items
  std::vector<MeasItem> items_e, items_g;
  std::vector<int> first_e;
  MeasItem* ditems_e = nullptr;
  MeasItem* ditems_g = nullptr;
  int* dfirst_e = nullptr;
  double* dpart = nullptr;
  int slice_e = 0, slice_g = 0;  // largest staged slice block of an item, doubles
  // the pair-correlation estimator of every pq_run_measured Measure
  PairCfArgs cf{};
  uint64_t* dcf = nullptr;       // [nchain][nbin] counts since the last collect
};

static DevCtx make_ctx(const pq_batch* b) {
  DevCtx x;
  x.chains = b->dchains;
  x.tables = b->dtables;
  x.tab_u = b->du;
  x.tab_ut = b->dut;
  x.tab_uf = b->duf;
  x.beads = b->dbeads;
  x.inv = b->dinv;
  x.egr = b->degr;
  x.upd = b->dupd;
  x.fnstate = b->dfn;
  x.fncache = b->dfncache;
  x.fnsweep = b->dfnsweep;
  x.scale = b->dscale;
  x.counter = b->dcounter;
  x.stats = b->dstats;
  x.inj = b->dinj;
  x.inj_per_chain = b->inj_per_chain;
  x.rng_mode = b->rng_mode;
  x.seed = b->seed;
  x.first_stream = b->first_stream;
  x.order = b->dorder;
  x.perm = b->dperm;
  x.ewk = b->dewk;
  x.plan = b->dplan;
  return x;
}

// The OptEwaldSum of a cell, rcut and kcut, built once per batch: its index
// in b->fits (its k-vectors at b->fit_off[index] of the k pool), -1 when the
// fit is not finite.
static int ewald_fit(pq_batch* b, int D, const double* a, double rcut, double kcut, int type, double kappa) {
  const auto key = std::make_tuple(D, a[0], a[1], D == 3 ? a[2] : 0.0, rcut, kcut, type, kappa);
  auto it = b->fit_of.find(key);
  if (it != b->fit_of.end()) return it->second;
  EwaldFit e;
  if (type ? !ewald_build_trad(e, D, a, rcut, kcut, kappa) : !ewald_build(e, D, a, rcut, kcut)) return -1;
  const int id = (int)b->fits.size();
  b->fit_off.push_back((int64_t)b->hk.size());
  for (int k = 0; k < e.nk(); ++k) {
    for (int d = 0; d < 3; ++d) b->hk.push_back(e.kvec[3 * k + d]);
    b->hk.push_back(2 * e.vk[k] * e.oneOver2V);
  }
  b->fits.push_back(std::move(e));
  b->fit_of[key] = id;
  return id;
}

static void ewald_params(const pq_batch* b, int id, const pq_chain_desc& d, int N, EwP& p) {
  const EwaldFit& e = b->fits[id];
  std::vector<double> q;
  for (int s = 0; s < d.nspecies; ++s)
    for (int i = 0; i < d.species[s].count; ++i) q.push_back(d.species[s].charge);
  p.nk = e.nk();
  p.off = b->fit_off[id];
  p.self = e.self_energy(q.data(), N);
  p.rcut = e.rcut;
  for (int i = 0; i < kEwaldPoly; ++i) p.coef[i] = e.coef[i];
}

static int scratch_buffer(pq_batch* b, size_t bytes, void** p) {
  if (bytes > b->scr_bytes) {
    if (b->dscr) (void)hipFree(b->dscr);
    b->dscr = nullptr;
    b->scr_bytes = 0;
    HIPCHK(hipMalloc(&b->dscr, bytes));
    b->scr_bytes = bytes;
  }
  *p = b->dscr;
  return PQ_OK;
}

static int build_chain(pq_batch* b, const pq_chain_desc& d, ChainP& c,
                       std::map<std::tuple<double, double, double, double, double, int, int, int>, int>& tabs) {
  std::memset(&c, 0, sizeof c);
  if (d.ndim != 2 && d.ndim != 3) return fail(PQ_ERR_ARG, "ndim must be 2 or 3 (got %d)", d.ndim);
  if (d.nspecies < 1 || d.nspecies > PQ_MAX_SPECIES) return fail(PQ_ERR_ARG, "nspecies out of range");
  if (d.nlevel < 0 || d.nlevel > 7) return fail(PQ_ERR_ARG, "nlevel must be in 0..7");
  if (!(d.tau > 0) || !std::isfinite(d.tau)) return fail(PQ_ERR_ARG, "tau must be positive and finite");
  // a chain needs at least one slice (a zero-length path has no bead to
  // start a section from) and a finite, positive SuperCell side per dimension
  if (d.nslice < 1) return fail(PQ_ERR_ARG, "nslice must be at least 1 (got %d)", d.nslice);
  for (int k = 0; k < d.ndim; ++k)
    if (!(d.cell[k] > 0) || !std::isfinite(d.cell[k]))
      return fail(PQ_ERR_ARG, "cell side %d must be positive and finite (got %g)", k, d.cell[k]);
  c.D = d.ndim; c.M = d.nslice; c.L = d.nlevel; c.S = (1 << d.nlevel) + 1; c.nsp = d.nspecies;
  c.tau = d.tau;
  if (d.ngroup < 0 || d.ngroup > PQ_MAX_GROUPS - 1) return fail(PQ_ERR_ARG, "ngroup must be in 0..%d", PQ_MAX_GROUPS - 1);
  c.ngroup = 1 + d.ngroup;
  c.Lmax = c.L;
  for (int g = 0; g < d.ngroup; ++g) {
    if (d.group[g].nlevel < 0 || d.group[g].nlevel > 7) return fail(PQ_ERR_ARG, "group %d: nlevel must be in 0..7", g + 1);
    c.Lmax = std::max(c.Lmax, d.group[g].nlevel);
  }
  c.Smax = (1 << c.Lmax) + 1;
  // a section may span the whole path (2^L = M, the reference's own SHO
  // system test): its moved slices 1..2^L-1 stay distinct and both ends are
  // bead (i, t0); shorter paths would move a slice twice
  if (c.M < c.Smax - 1) return fail(PQ_ERR_ARG, "nslice %d shorter than a section's 2^L = %d links", c.M, c.Smax - 1);
  int n = 0;
  for (int s = 0; s < c.nsp; ++s) {
    if (d.species[s].count < 1 || d.species[s].mass <= 0) return fail(PQ_ERR_ARG, "bad species %d", s);
    c.ifirst[s] = n;
    c.count[s] = d.species[s].count;
    c.mass[s] = d.species[s].mass;
    c.charge[s] = d.species[s].charge;
    c.stat[s] = d.species[s].is_static ? 1 : 0;
    c.lambda[s] = 0.5 / d.species[s].mass;
    n += d.species[s].count;
  }
  c.N = n;
  // SuperCell (src/util/SuperCell.cc:14-23)
  double l2 = d.cell[0] * d.cell[0];
  c.rcut2 = l2;
  for (int k = 0; k < 3; ++k) { c.a[k] = d.cell[k]; c.b[k] = (k < c.D) ? 1. / d.cell[k] : 0; }
  for (int k = 1; k < c.D; ++k) { l2 = d.cell[k] * d.cell[k]; c.rcut2 = (c.rcut2 > l2) ? l2 : c.rcut2; }
  c.rcut2 *= 0.25;
  // periodic Gaussians of SpringAction (SpringAction.cc:31-49) and the movers
  // (FreeMover.cc:37-51, FreeMoverPBC.cc:37-51), used when fewer than 16 terms
  for (int l = 0; l < PQ_PG_LEVELS; ++l)
    for (int s = 0; s < c.nsp; ++s)
      for (int k = 0; k < c.D; ++k) {
        c.spg_on[l][s][k] = (l <= c.Lmax && pg_nterms(c.mass[s] / (2 * c.tau * (1 << l)), c.a[k]) < 16) ? 1 : 0;
        c.mpg_on[l][s][k] = (l <= c.Lmax + 1 && pg_nterms(c.mass[s] / (c.tau * (1 << l)), c.a[k]) < 16) ? 1 : 0;
      }
  // per-level constants of SpringAction (SpringAction.cc:71) and SHOAction (SHOAction.cc:28-33)
  for (int l = 0; l < PQ_PG_LEVELS; ++l) {
    c.div2[l] = 1.0 / (2.0 * c.tau * (1 << l));
    for (int s = 0; s < c.nsp; ++s) c.spring_c[l][s] = 0.25 / (c.lambda[s] * c.tau * (1 << l));
  }
  c.spring = d.spring;
  if (d.nsho < 0 || d.nsho > PQ_MAX_SHO) return fail(PQ_ERR_ARG, "nsho out of range");
  c.nsho = d.nsho;
  for (int a = 0; a < d.nsho; ++a) {
    if (d.sho[a].species < 0 || d.sho[a].species >= c.nsp) return fail(PQ_ERR_ARG, "SHO species");
    c.sho_sp[a] = d.sho[a].species;
    c.sho_ndim[a] = d.sho[a].ndim;
    c.sho_omega[a] = d.sho[a].omega == 0 ? 1.0 : d.sho[a].omega;
    for (int k = 0; k < 3; ++k) c.sho_center[a][k] = d.sho[a].center[k];
    for (int l = 0; l < PQ_PG_LEVELS; ++l) {
      const double wt = c.sho_omega[a] * c.tau * (1 << l);
      c.sho_cosh[a][l] = std::cosh(wt);
      c.sho_div1[a][l] = 1.0 / (2.0 * std::sinh(wt));
    }
  }
  // CoulombAction constructor (src/action/CoulombAction.cc:45-90): rmin, rmax
  // and ngpts defaults set by the first interacting pair stick for the rest.
  if (d.coulomb.enabled) {
    if (d.coulomb.norder < 0 || d.coulomb.norder > PQ_MAX_NORDER) return fail(PQ_ERR_ARG, "norder out of range");
    const double eps = d.coulomb.epsilon == 0 ? 1 : d.coulomb.epsilon;
    c.ceps = eps;
    // the Ewald sum comes first: the tables subtract its f(r)
    // (CoulombAction.cc:28-44, ewaldType optEwald with ewaldNDim = NDIM)
    int fit = -1;
    if (d.coulomb.use_ewald) {
      const double rc = d.coulomb.ewald_rcut == 0. ? d.cell[0] / 2. : d.coulomb.ewald_rcut;
      if (!(rc > 0) || !(d.coulomb.ewald_kcut > 0)) return fail(PQ_ERR_ARG, "Ewald sum needs rcut > 0 and kcut > 0");
      if (d.coulomb.ewald_type != 0 && d.coulomb.ewald_type != 1) return fail(PQ_ERR_ARG, "unknown ewald_type");
      double kappa = d.coulomb.ewald_kappa;
      if (d.coulomb.ewald_type == 1 && kappa == 0) {
        // the parser's default (ActionParser.cc:137-146)
        double lside = 1000000000000000.0;
        for (int k = 0; k < c.D; ++k) lside = std::min(lside, d.cell[k]);
        kappa = std::sqrt(std::pow(3.6 * c.N, 1.0 / 6.0) * std::sqrt(3.1415926535897931) / lside);
      }
      fit = ewald_fit(b, c.D, d.cell, rc, d.coulomb.ewald_kcut, d.coulomb.ewald_type, kappa);
      if (fit < 0) return fail(PQ_ERR_ARG, "the Ewald sum for rcut %g, kcut %g is not finite", rc, d.coulomb.ewald_kcut);
      ewald_params(b, fit, d, c.N, c.ew);
    }
    const EwaldFit* ewp = fit >= 0 ? &b->fits[fit] : nullptr;
    double rmin = d.coulomb.rmin, rmax = d.coulomb.rmax;
    int ngpts = d.coulomb.ngpts;
    for (int i = 0; i < c.nsp; ++i)
      for (int j = i; j < c.nsp; ++j) {
        const double q1q2 = d.species[i].charge * d.species[j].charge / eps;
        if (!(q1q2 != 0 && (i != j || d.species[i].count > 1))) continue;
        const double mu = 1. / (1. / d.species[i].mass + 1. / d.species[j].mass);
        if (rmin == 0) rmin = 0.005 / (2 * ((mu > 500) ? 1 : mu) * std::fabs(q1q2));
        if (rmax == 0) {
          double a2 = 0;
          for (int k = 0; k < c.D; ++k) a2 += d.cell[k] * d.cell[k];
          rmax = std::sqrt(a2);
        }
        if (ngpts == 0) ngpts = 500;
        if (ngpts < 2) return fail(PQ_ERR_ARG, "ngridPoints must be at least 2");
        const int norder = (mu > 500) ? 0 : d.coulomb.norder;
        if (c.npair >= PQ_MAX_PAIRS) return fail(PQ_ERR_ARG, "too many pair actions");
        auto key = std::make_tuple(mu, q1q2, c.tau, rmin, rmax, ngpts, norder, fit);
        auto it = tabs.find(key);
        int tid;
        if (it == tabs.end()) {
          // PairAction(EmpiricalPairAction) grid (src/action/PairAction.cc:161-185)
          TableP t;
          t.ngpts = ngpts;
          t.K1 = norder + 1;
          t.rgridinv = 1.0 / rmin;
          t.logrratioinv = (ngpts - 1) / std::log(rmax / rmin);
          t.off = (int64_t)b->hu.size();
          b->hu.resize(b->hu.size() + (size_t)ngpts * t.K1);
          b->hut.resize(b->hu.size());
          for (int ig = 0; ig < ngpts; ++ig) {
            const double r = (1. / t.rgridinv) * std::exp(ig / t.logrratioinv);
            for (int k = 0; k < t.K1; ++k) {
              b->hu[t.off + (int64_t)ig * t.K1 + k] = coulomb_u(r, k, mu, q1q2, c.tau, ewp);
              b->hut[t.off + (int64_t)ig * t.K1 + k] = coulomb_utau(r, k, mu, q1q2, c.tau, ewp);
            }
          }
          tid = (int)b->ht.size();
          b->ht.push_back(t);
          tabs[key] = tid;
        } else {
          tid = it->second;
        }
        c.pair[c.npair].s1 = i;
        c.pair[c.npair].s2 = j;
        c.pair[c.npair].table = tid;
        c.pair[c.npair].tp = b->ht[tid];
        c.npair++;
      }
  }
  if (d.nnodes < 0 || d.nnodes > PQ_MAX_NODES) return fail(PQ_ERR_ARG, "nnodes out of range");
  c.nfn = d.nnodes;
  const double T = 1.0 / (c.tau * c.M);
  for (int a = 0; a < d.nnodes; ++a) {
    const pq_node_desc& nd = d.nodes[a];
    FnP& f = c.fn[a];
    if (nd.species < 0 || nd.species >= c.nsp) return fail(PQ_ERR_ARG, "node species");
    f.sp = nd.species;
    f.model = nd.model;
    f.ifirst = c.ifirst[nd.species];
    f.ns = c.count[nd.species];
    if (f.ns > PQ_MAX_NODE_PART) return fail(PQ_ERR_ARG, "node species of %d particles exceeds %d", f.ns, PQ_MAX_NODE_PART);
    f.nodal = nd.nodal_action;
    f.manybody = nd.many_body_distance ? 1 : 0;
    f.updates = nd.use_updates ? 1 : 0;
    if (f.updates && nd.model != PQ_NODES_FREE)
      return fail(PQ_ERR_ARG, "use_updates is a FreeParticleNodes option (FreeParticleNodes.cc:60)");
    f.mass = c.mass[nd.species];
    const double Tn = nd.temperature == 0 ? T : nd.temperature;
    // SHONodes constructor (src/fixednode/SHONodes.cc:22-35)
    f.coshwt = std::cosh(nd.omega * 0.5 / Tn);
    const double sinhwt = std::sinh(nd.omega * 0.5 / Tn);
    f.c = f.mass * 0.5 * nd.omega / sinhwt;
    // FreeParticleNodes periodic Gaussians, alpha = m T (FreeParticleNodes.cc:54-58)
    const double PI = 3.141592653589793;
    for (int k = 0; k < c.D; ++k) {
      const double alpha = f.mass * Tn, len = d.cell[k];
      f.pg_nome[k] = std::exp(-PI * PI / (alpha * len * len));
      f.pg_nmax[k] = pg_nterms(alpha, len);
      f.pg_pref[k] = std::sqrt(PI / (alpha * len * len));
      f.pg_k[k] = 2 * PI / len;
      f.pg_alpha[k] = alpha;
      f.pg_len[k] = len;
      // with alpha L^2 >= 20 the images beyond the nearest three weigh less than
      // e^-40 of the leading one, so their sum equals the theta series to rounding
      f.pg_img[k] = (alpha * len * len >= 20.0) ? 1 : 0;
    }
    if (nd.model != PQ_NODES_SHO && nd.model != PQ_NODES_FREE) return fail(PQ_ERR_ARG, "unknown node model");
    c.ns_max = std::max(c.ns_max, f.ns);
  }
  if (c.nfn && 2 * c.Smax > c.M)
    return fail(PQ_ERR_ARG, "the two sections beta/2 apart overlap: need 2(2^L+1) <= M");
  if (d.mover != PQ_MOVER_FREE && d.mover != PQ_MOVER_FREE_PBC) return fail(PQ_ERR_ARG, "unknown mover");
  c.mover = d.mover;
  // the section groups of a sweep: group 0 from the chain's own fields
  for (int g = 0; g < c.ngroup; ++g) {
    const int L = g ? d.group[g - 1].nlevel : d.nlevel;
    const int nsample = g ? d.group[g - 1].nsample : d.nsample;
    const pq_sample_desc* smp = g ? d.group[g - 1].sample : d.sample;
    GroupP& G = c.grp[g];
    G.L = L;
    G.nsection = g ? d.group[g - 1].nsection : d.nsection;
    if (G.nsection < 0) return fail(PQ_ERR_ARG, "group %d: nsection must not be negative", g);
    if (nsample < 1 || nsample > PQ_MAX_SAMPLES) return fail(PQ_ERR_ARG, "group %d: nsample out of range", g);
    G.nsample = nsample;
    for (int k = 0; k < nsample; ++k) {
      const int ch = smp[k].chooser;
      if (ch < PQ_CHOOSER_SPECIES || ch > PQ_CHOOSER_ALL_RANDOM) return fail(PQ_ERR_ARG, "unknown chooser");
      const bool all = ch == PQ_CHOOSER_ALL || ch == PQ_CHOOSER_ALL_RANDOM;
      if (!all && (smp[k].species < 0 || smp[k].species >= c.nsp)) return fail(PQ_ERR_ARG, "sample species");
      // the reference's choosers redraw until they hold nmoving distinct
      // particles: more than the pool would never end
      const int pool = all ? c.N : c.count[smp[k].species];
      if (smp[k].nmoving < 1 || smp[k].nmoving > PQ_MAX_MOVING || smp[k].nmoving > pool)
        return fail(PQ_ERR_ARG, "sample %d: nmoving %d not in 1..min(%d, %d)", k, smp[k].nmoving, PQ_MAX_MOVING, pool);
      G.sample_sp[k] = all ? 0 : smp[k].species;
      G.sample_ch[k] = ch;
      G.sample_nm[k] = smp[k].nmoving;
      if (smp[k].nrepeat < 0) return fail(PQ_ERR_ARG, "sample %d: nrepeat must not be negative", k);
      G.sample_rep[k] = smp[k].nrepeat == 0 ? 1 : smp[k].nrepeat;
      // PIMCParser.cc:312-322: defaultSigmaFactor 0 means 1, newSigmaFactor 0
      // means 0.75 under delayedRejection
      if (!(smp[k].sigma_factor >= 0) || !(smp[k].new_sigma_factor >= 0) ||
          !std::isfinite(smp[k].sigma_factor) || !std::isfinite(smp[k].new_sigma_factor))
        return fail(PQ_ERR_ARG, "sample %d: sigma factors must be finite and not negative", k);
      G.sample_dr[k] = smp[k].delayed_rejection ? 1 : 0;
      G.sample_fac[k] = smp[k].sigma_factor == 0 ? 1.0 : smp[k].sigma_factor;
      G.sample_nfac[k] = smp[k].new_sigma_factor == 0 ? (G.sample_dr[k] ? 0.75 : 0.0) : smp[k].new_sigma_factor;
      if (G.sample_dr[k] && d.mover != PQ_MOVER_FREE)
        return fail(PQ_ERR_ARG, "sample %d: delayed rejection needs the Free mover", k);
      // a permuted section relabels the slices after it, which a section of
      // the whole path (2^L = M) would overlap
      const bool permuting = (ch == PQ_CHOOSER_WALKING || ch == PQ_CHOOSER_ALL_RANDOM) && smp[k].nmoving > 1;
      if (permuting && c.M < (1 << L) + 1)
        return fail(PQ_ERR_ARG, "sample %d permutes particles: nslice %d must exceed 2^L = %d", k, c.M, 1 << L);
      // the reference sizes MatrixUpdate for 3 moving particles of the
      // species (maxMovers, ActionParser.cc:809-811)
      for (int a = 0; a < c.nfn; ++a)
        if (c.fn[a].updates && (all || smp[k].species == c.fn[a].sp) && smp[k].nmoving > 3)
          return fail(PQ_ERR_ARG, "sample %d: nmoving %d exceeds the 3 moving particles of use_updates", k,
                      smp[k].nmoving);
    }
  }
  c.nsample = c.grp[0].nsample;
  for (int k = 0; k < c.nsample; ++k) {
    c.sample_sp[k] = c.grp[0].sample_sp[k];
    c.sample_ch[k] = c.grp[0].sample_ch[k];
    c.sample_nm[k] = c.grp[0].sample_nm[k];
    c.sample_rep[k] = c.grp[0].sample_rep[k];
    c.sample_dr[k] = c.grp[0].sample_dr[k];
    c.sample_fac[k] = c.grp[0].sample_fac[k];
    c.sample_nfac[k] = c.grp[0].sample_nfac[k];
  }
  c.nsection = d.nsection;
  return PQ_OK;
}

static int64_t align8(int64_t v) { return (v + 7) & ~(int64_t)7; }

// LDS carve-up of the sampler workgroup, sized by the batch's largest chain.
// Each request is -1 (automatic), 0 or 1: team = threads per chain (64 or
// 128), tab = pair tables staged in LDS, fst = fixed-node state in LDS,
// inv = node inverse cache in LDS.  Automatic placement fills a 40 KB budget
// (four chains per CU) in the order tables, fixed-node state, inverses;
// requested placements may use the whole 160 KB.
static SweepLayoutH layout_for(const pq_batch* b, const Sched& req, int lo = 0, int hi = -1, int NC = 0) {
  SweepLayoutH lay{};
  int64_t mS = 0, mSND = 0, mSD = 0, mF = 0, mInv = 0, mTab = 0, mN = 0, mWalk = 0;
  bool nodes = false;
  const int64_t w = b->precision == PQ_PREC_MIXED ? (int64_t)sizeof(float) : (int64_t)sizeof(double);
  if (hi < 0) hi = b->nchain;
  if (NC == 0) NC = b->NC;
  // the chains of this launch: order[lo..hi)
  for (int q = lo; q < hi; ++q) {
    const ChainP& ch = b->hc[b->order.empty() ? q : b->order[q]];
    mS = std::max<int64_t>(mS, ch.Smax);
    mSND = std::max<int64_t>(mSND, (int64_t)ch.Smax * ch.N * ch.D);
    int nm = 1;
    for (int g = 0; g < ch.ngroup; ++g)
      for (int k = 0; k < ch.grp[g].nsample; ++k) {
        nm = std::max(nm, ch.grp[g].sample_nm[k]);
        if (ch.grp[g].sample_ch[k] == PQ_CHOOSER_WALKING)
          mWalk = std::max<int64_t>(mWalk, 2 * (int64_t)ch.grp[g].sample_nm[k] * ch.count[ch.grp[g].sample_sp[k]]);
      }
    mSD = std::max<int64_t>(mSD, (int64_t)ch.Smax * nm * ch.D);
    mN = std::max<int64_t>(mN, ch.N);
    mF = std::max<int64_t>(mF, fn_state_doubles(ch));
    if (ch.nfn) { nodes = true; mInv = std::max<int64_t>(mInv, (int64_t)ch.Smax * ch.ns_max * ch.ns_max); }
    int64_t tabs = 0;
    for (int pp = 0; pp < ch.npair; ++pp) {
      int q = 0;
      while (q < pp && ch.pair[q].table != ch.pair[pp].table) ++q;
      if (q == pp) tabs += (int64_t)b->ht[ch.pair[pp].table].ngpts * b->ht[ch.pair[pp].table].K1;
    }
    mTab = std::max<int64_t>(mTab, tabs);
  }
  // team: two waves per chain when node matrices dominate, else one; either
  // way two waves fit per SIMD.  Four-wave teams serve batches of few chains.
  int team = nodes ? 128 : PQ_WAVE;
  if (req.team == 64 || req.team == 128 || req.team == 256 || req.team == 512) team = req.team;
  // the sampler instances for several moving particles have one- and two-wave teams
  if (b->multi && team > 128) team = 128;
  lay.team = team;
  int64_t o = align8(sizeof(ChainP));
  lay.off_sec1 = (int)o; o += mSND * 8;
  lay.off_sec2 = (int)o; o += mSND * 8;
  lay.off_mov = (int)o; o += align8(mSD * 8);
  // per-matrix determinants and error flags: one pass of node matrices holds
  // up to team / 4 of them (four-lane groups), a section up to Smax
  const int64_t nwd = std::max<int64_t>(PQ_WAVE, std::max<int64_t>(mS, team / 4));
  lay.off_wdet = (int)o; o += nwd * 8;
  lay.off_wscale = (int)o; o += align8(mS * 8);
  lay.off_scale = (int)o; o += PQ_MAX_NODES * 8;
  lay.off_red = (int)o; o += 16 * 8;   // two halves of 8 waves
  lay.off_cnt = (int)o; o += 24 * 8;
  lay.off_werr = (int)o; o += nwd * 4;
  lay.off_notmine = (int)o; o += 2 * PQ_MAX_NODES * 4;   // notmine, then last-slice error flags
  lay.off_tabofs = (int)o; o += PQ_MAX_PAIRS * 4;
  o = align8(o);
  lay.off_choose = (int)o; o += 512;   // Eng::ChooseBuf
  lay.off_perm = (int)o; o += align8(5 * mN * 4);   // perm, inverse, section perm, two scratch
  lay.off_walk = -1;
  if (mWalk) { lay.off_walk = (int)o; o += mWalk * 8; }
  // Gauss-Jordan scratch: pivot rows (2 doubles per thread), or for 32-wide
  // node matrices the MFMA scratch of every wave of the team
  lay.off_gj = (int)o;
  o += NC == 32 ? (int64_t)(team / PQ_WAVE) * PQ_GJ32_SCRATCH_DOUBLES * 8 : (int64_t)2 * 8 * team;
  const int64_t budget = 40 * 1024, whole = 160 * 1024;
  auto place = [&](int r, int64_t bytes, bool wanted) -> int {
    if (!wanted || bytes <= 0 || r == 0) return -1;
    if (o + bytes > (r == 1 ? whole : budget)) return -1;
    const int at = (int)o;
    o = align8(o + bytes);
    return at;
  };
  lay.off_tab = place(req.tab, mTab * w, mTab > 0);
  lay.off_fst = place(req.fst, mF * 8, nodes);
  lay.off_inv = place(req.inv, mInv * 8, nodes);
  lay.bytes = (int)align8(o);
  return lay;
}

static void make_layout(pq_batch* b, const Sched& req) {
  if (req.ncsplit && b->wide > 0 && b->wide < b->nchain) {
    Sched h = req;
    h.team = req.heavy_team;
    b->lay_heavy = layout_for(b, h, 0, b->wide, b->NC);
    b->lay = layout_for(b, req, b->wide, b->nchain, b->NC_bulk);
    return;
  }
  b->lay = layout_for(b, req);
  b->lay_heavy = b->lay;
  if (req.heavy > 0) {
    Sched h = req;
    h.team = req.heavy_team;
    b->lay_heavy = layout_for(b, h);
  }
}

static int layout_bytes(const pq_batch* b) { return std::max(b->lay.bytes, b->lay_heavy.bytes); }

// one sampler launch over the batch: the heaviest chains (first in the cost
// order) on the second stream with their own team size, the rest on the
// batch stream; the batch stream waits for both
static hipError_t run_launch(pq_batch* b, const DevCtx& x, int nsweep) {
  const bool mixed = b->precision == PQ_PREC_MIXED;
  const bool ncs = b->sched.ncsplit && b->wide > 0 && b->wide < b->nchain;
  const int h = ncs ? b->wide : b->sched.heavy;
  const int nc_rest = ncs ? b->NC_bulk : b->NC;
  if (h <= 0 || h >= b->nchain) return launch_sweep(b->D, b->NC, mixed, b->multi, x, b->nchain, nsweep, b->lay, b->stream);
  hipError_t e = hipEventRecord(b->evf, b->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->stream2, b->evf, 0);
  // the heavy chains' launch takes its chains from a queue (zeroed by a fill
  // kernel, which graph capture records as a kernel node): its workgroups
  // keep their slots until every heavy chain is taken, and the other
  // launch fills the slots they leave after that
  SweepLayoutH lh = b->lay_heavy;
  if (!b->multi && lh.team >= 256) lh.queue = reinterpret_cast<int*>(b->dqueue);
  if (e == hipSuccess && lh.queue) e = launch_fill_u64(b->dqueue, 1, b->stream2);
  if (e == hipSuccess) e = launch_sweep(b->D, b->NC, mixed, b->multi, x, h, nsweep, lh, b->stream2);
  DevCtx xl = x;
  xl.order = x.order + h;
  if (e == hipSuccess) e = launch_sweep(b->D, nc_rest, mixed, b->multi, xl, b->nchain - h, nsweep, b->lay, b->stream);
  if (e == hipSuccess) e = hipEventRecord(b->evj, b->stream2);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->evj, 0);
  return e;
}

// Measure work items.  An energy item costs about N spring/SHO links plus
// N(N-1)/2 pair links per slice; items hold up to ~32k of those (64 per
// thread of a 512-thread workgroup, and always fewer than 2^16 (slice,
// pair) entries) and at most 3840 staged doubles: with the 32 KB of utau
// tables and ~6 KB of fixed LDS an energy workgroup stays under half of
// the CU's 160 KB, so two run per CU.  Histogram items are
// coarser (~64k distances, at most 2048 staged doubles, six workgroups per
// CU) so that each adds its bins to the chain's global counts once.
static void build_measure_items(pq_batch* b) {
  b->items_e.clear();
  b->items_g.clear();
  b->first_e.assign(1, 0);
  b->slice_e = b->slice_g = 0;
  for (int c = 0; c < b->nchain; ++c) {
    const ChainP& ch = b->hc[c];
    const int64_t nd = (int64_t)ch.N * ch.D;
    const int64_t per = ch.N + (ch.npair ? (int64_t)ch.N * (ch.N - 1) / 2 : 0) + (ch.ew.nk ? (int64_t)ch.ew.nk * ch.N / 4 : 0);
    int64_t te = std::max<int64_t>(1, 32768 / std::max<int64_t>(1, per));
    const int64_t ss = meas_slice_stride((int)nd);
    te = std::min<int64_t>(te, std::max<int64_t>(1, 3840 / ss - 1));
    te = std::min<int64_t>(te, ch.M);
    for (int t = 0; t < ch.M; t += (int)te) b->items_e.push_back(MeasItem{c, t, (int)std::min<int64_t>(ch.M, t + te)});
    b->first_e.push_back((int)b->items_e.size());
    b->slice_e = std::max<int>(b->slice_e, (int)((te + 1) * ss));
    int64_t tg = std::max<int64_t>(1, 65536 / std::max<int64_t>(1, (int64_t)ch.N * ch.N));
    tg = std::min<int64_t>(tg, std::max<int64_t>(1, 2048 / nd - 1));
    tg = std::min<int64_t>(tg, ch.M);
    for (int t = 0; t < ch.M; t += (int)tg) b->items_g.push_back(MeasItem{c, t, (int)std::min<int64_t>(ch.M, t + tg)});
    b->slice_g = std::max<int>(b->slice_g, (int)((tg + 1) * nd));
  }
}

static int static_scratch_setup(int device);

extern "C" int pq_batch_create(const pq_chain_desc* chains, int nchain, int precision, int device,
                               pq_batch** out) {
  if (!chains || nchain < 1 || !out) return fail(PQ_ERR_ARG, "pq_batch_create: bad arguments");
  if (precision != PQ_PREC_FP64 && precision != PQ_PREC_MIXED) return fail(PQ_ERR_ARG, "unknown precision");
  pq_batch* b = new pq_batch;
  b->device = device;
  b->precision = precision;
  b->nchain = nchain;
  b->descs.assign(chains, chains + nchain);
  b->hc.resize(nchain);
  b->D = chains[0].ndim;
  std::map<std::tuple<double, double, double, double, double, int, int, int>, int> tabs;
  int64_t beads = 0, inv = 0, fnst = 0, egr = 0, fnc = 0, upd = 0;
  int ns_max = 1;
  for (int c = 0; c < nchain; ++c) {
    if (chains[c].ndim != b->D) { delete b; return fail(PQ_ERR_ARG, "all chains of a batch share ndim"); }
    int rc = build_chain(b, chains[c], b->hc[c], tabs);
    if (rc) { delete b; return rc; }
    ns_max = std::max(ns_max, b->hc[c].ns_max);
  }
  // the descriptors are checked before any device call, so a bad argument
  // is PQ_ERR_ARG on any host
  {
    int ndev = 0;
    const hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) { delete b; return fail(PQ_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e)); }
    if (device < 0 || device >= ndev) { delete b; return fail(PQ_ERR_ARG, "device %d not present (%d devices)", device, ndev); }
    if (hipSetDevice(device) != hipSuccess) { delete b; return fail(PQ_ERR_HIP, "hipSetDevice(%d) failed", device); }
    if (int rc = static_scratch_setup(device)) { delete b; return rc; }
  }
  for (int c = 0; c < nchain; ++c)
    for (int g = 0; g < b->hc[c].ngroup; ++g)
      for (int k = 0; k < b->hc[c].grp[g].nsample; ++k) {
        // the general instances also carry sigma factors and delayed rejection
        const GroupP& G = b->hc[c].grp[g];
        b->multi |= G.sample_nm[k] > 1 || G.sample_dr[k] || G.sample_fac[k] != 1.0;
      }
  // and FreeParticleNodes::MatrixUpdate
  for (int c = 0; c < nchain; ++c) b->multi |= chain_has_updates(b->hc[c]);
  // node-matrix row width: the instantiated widths, the smallest that fits
  b->NC = ns_max <= 4 ? 4 : ns_max <= 8 ? 8 : ns_max <= 10 ? 10 : ns_max <= 12 ? 12 : ns_max <= 16 ? 16 : 32;
  for (int c = 0; c < nchain; ++c) {
    ChainP& ch = b->hc[c];
    ch.bead_off = beads;
    beads += (int64_t)ch.M * ch.N * ch.D;
    ch.inv_off = inv;
    if (ch.nfn) inv += (int64_t)ch.Smax * ch.ns_max * ch.ns_max;
    ch.egr_off = egr;
    if (ch.nfn) egr += (int64_t)ch.Smax * ch.ns_max * ch.ns_max * ch.D;
    ch.upd_off = -1;
    if (chain_has_updates(ch)) {
      ch.upd_off = upd;
      upd += ch.nfn * upd_block_doubles(ch);
    }
    ch.fn_off = fnst;
    fnst += fn_state_doubles(ch);
    const int64_t nrec = fn_cache_doubles(ch);
    ch.fnc_off = nrec ? fnc : -1;
    fnc += nrec;
    ch.perm_off = b->nperm;
    b->nperm += 2 * (int64_t)ch.N;
  }
  build_measure_items(b);
  make_layout(b, b->sched);
  if (layout_bytes(b) > 160 * 1024) {
    int n = layout_bytes(b);
    delete b;
    return fail(PQ_ERR_ARG, "chain too large for one workgroup's LDS (%d bytes)", n);
  }
  b->nbeads = beads;
  // heaviest chains first so the longest workgroups start earliest
  b->order.resize(nchain);
  for (int c = 0; c < nchain; ++c) b->order[c] = c;
  // estimated sweep time: moves x levels x (a fixed cost + one per particle).
  // Measured per-chain sweep times of the uneven batches follow this to a
  // correlation of 0.97 (the node matrices' n^3 work hides under the
  // latency of their n pivot steps and of the level chain).
  std::vector<double> cost(nchain);
  for (int c = 0; c < nchain; ++c) {
    const ChainP& ch = b->hc[c];
    cost[c] = 0;
    for (int g = 0; g < ch.ngroup; ++g)
      cost[c] += (double)ch.grp[g].nsection * ch.grp[g].nsample * (ch.grp[g].L + 1) * (40.0 + ch.N);
  }
  std::stable_sort(b->order.begin(), b->order.end(), [&](int p, int q) { return cost[p] > cost[q]; });
  // with 32-wide node matrices beside narrower ones, the 32-wide chains go
  // first (each class in cost order): a node-width split launches them apart
  // and the rest with the narrower Gauss-Jordan instance
  if (b->NC == 32) {
    int ns_rest = 0;
    for (int c = 0; c < nchain; ++c) {
      if (b->hc[c].ns_max > 16) ++b->wide;
      else ns_rest = std::max(ns_rest, b->hc[c].ns_max);
    }
    if (b->wide > 0 && b->wide < nchain) {
      std::stable_sort(b->order.begin(), b->order.end(),
                       [&](int p, int q) { return (b->hc[p].ns_max > 16) > (b->hc[q].ns_max > 16); });
      b->NC_bulk = ns_rest <= 4 ? 4 : ns_rest <= 8 ? 8 : ns_rest <= 10 ? 10 : ns_rest <= 12 ? 12 : 16;
    } else {
      b->wide = 0;
    }
  }

  auto cleanup_fail = [&](int rc) { pq_batch_destroy(b); return rc; };
  if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->stream2) != hipSuccess)
    return cleanup_fail(fail(PQ_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&b->evf, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->evj, hipEventDisableTiming) != hipSuccess)
    return cleanup_fail(fail(PQ_ERR_HIP, "hipEventCreate failed"));
  auto dalloc = [&](void** p, size_t bytes) -> bool {
    if (bytes == 0) bytes = 8;
    return hipMalloc(p, bytes) == hipSuccess;
  };
  bool ok = dalloc((void**)&b->dchains, sizeof(ChainP) * nchain) &&
            dalloc((void**)&b->dqueue, sizeof(uint64_t)) &&
            dalloc((void**)&b->dtables, sizeof(TableP) * std::max<size_t>(1, b->ht.size())) &&
            dalloc((void**)&b->du, sizeof(double) * b->hu.size()) &&
            dalloc((void**)&b->dut, sizeof(double) * b->hut.size()) &&
            dalloc((void**)&b->duf, sizeof(float) * b->hu.size()) &&
            dalloc((void**)&b->dbeads, sizeof(double) * beads) &&
            dalloc((void**)&b->dinv, sizeof(double) * inv) &&
            dalloc((void**)&b->degr, sizeof(double) * egr) &&
            dalloc((void**)&b->dupd, sizeof(double) * upd) &&
            dalloc((void**)&b->dfn, sizeof(double) * fnst) &&
            dalloc((void**)&b->dfncache, sizeof(double) * fnc) &&
            dalloc((void**)&b->dfnsweep, sizeof(double) * fnst) &&
            dalloc((void**)&b->dscale, sizeof(double) * nchain * PQ_MAX_NODES) &&
            dalloc((void**)&b->dcounter, sizeof(uint64_t) * nchain) &&
            dalloc((void**)&b->dstats, sizeof(pq_chain_stats) * nchain) &&
            dalloc((void**)&b->dorder, sizeof(int) * nchain) &&
            dalloc((void**)&b->dperm, sizeof(int) * b->nperm) &&
            dalloc((void**)&b->dewk, sizeof(double) * b->hk.size()) &&
            dalloc((void**)&b->ditems_e, sizeof(MeasItem) * b->items_e.size()) &&
            dalloc((void**)&b->ditems_g, sizeof(MeasItem) * b->items_g.size()) &&
            dalloc((void**)&b->dfirst_e, sizeof(int) * b->first_e.size()) &&
            dalloc((void**)&b->dpart, sizeof(double) * b->items_e.size());
  if (!ok) return cleanup_fail(fail(PQ_ERR_NOMEM, "device allocation failed"));
  b->nfncache = fnc;
  std::vector<float> uf(b->hu.begin(), b->hu.end());
  std::vector<double> ones((size_t)nchain * PQ_MAX_NODES, 1.0);
  std::vector<int> ident((size_t)b->nperm);
  for (int c = 0; c < nchain; ++c)
    for (int i = 0; i < b->hc[c].N; ++i) ident[b->hc[c].perm_off + i] = ident[b->hc[c].perm_off + b->hc[c].N + i] = i;
  if (hipMemcpy(b->dchains, b->hc.data(), sizeof(ChainP) * nchain, hipMemcpyHostToDevice) != hipSuccess ||
      (b->ht.size() && hipMemcpy(b->dtables, b->ht.data(), sizeof(TableP) * b->ht.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (b->hu.size() && hipMemcpy(b->du, b->hu.data(), sizeof(double) * b->hu.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (b->hut.size() && hipMemcpy(b->dut, b->hut.data(), sizeof(double) * b->hut.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (uf.size() && hipMemcpy(b->duf, uf.data(), sizeof(float) * uf.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (b->hk.size() && hipMemcpy(b->dewk, b->hk.data(), sizeof(double) * b->hk.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(b->ditems_e, b->items_e.data(), sizeof(MeasItem) * b->items_e.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->ditems_g, b->items_g.data(), sizeof(MeasItem) * b->items_g.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dfirst_e, b->first_e.data(), sizeof(int) * b->first_e.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dscale, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dorder, b->order.data(), sizeof(int) * nchain, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dperm, ident.data(), sizeof(int) * ident.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->dcounter, 0, sizeof(uint64_t) * nchain) != hipSuccess ||
      hipMemset(b->dstats, 0, sizeof(pq_chain_stats) * nchain) != hipSuccess ||
      hipMemset(b->dbeads, 0, sizeof(double) * std::max<int64_t>(1, beads)) != hipSuccess ||
      hipMemset(b->dfn, 0, sizeof(double) * std::max<int64_t>(1, fnst)) != hipSuccess ||
      hipMemset(b->dfncache, 0, sizeof(double) * std::max<int64_t>(1, fnc)) != hipSuccess ||
      hipMemset(b->dupd, 0, sizeof(double) * std::max<int64_t>(1, upd)) != hipSuccess)
    return cleanup_fail(fail(PQ_ERR_HIP, "device upload failed"));
  *out = b;
  return PQ_OK;
}

static void drop_graph(pq_batch* b);

extern "C" void pq_batch_destroy(pq_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  void* ptrs[] = {b->dchains, b->dtables, b->du, b->dut, b->duf, b->dbeads, b->dinv, b->dfn, b->dfncache, b->dfnsweep,
                  b->dscale, b->dcounter, b->dstats, b->dinj, b->dorder, b->dscr, b->dfnscr, b->degr, b->dupd,
                  b->dperm, b->dewk, b->dew_est, b->dewk_est, b->dvoff_est, b->dvgrid_est,
                  b->ditems_e, b->ditems_g, b->dfirst_e, b->dpart, b->dcf, b->dplan, b->dbead_stage, b->dqueue};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  drop_graph(b);
  if (b->dmeas) (void)hipFree(b->dmeas);
  for (hipEvent_t e : b->gev) (void)hipEventDestroy(e);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  if (b->evf) (void)hipEventDestroy(b->evf);
  if (b->evj) (void)hipEventDestroy(b->evj);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

static int check_chain(const pq_batch* b, int c) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (c < 0 || c >= b->nchain) return fail(PQ_ERR_ARG, "chain %d out of range", c);
  return PQ_OK;
}

extern "C" int pq_batch_info(const pq_batch* b, int c, int* npart, int* nslice, int* ndim, int64_t* off) {
  if (int rc = check_chain(b, c)) return rc;
  const ChainP& ch = b->hc[c];
  if (npart) *npart = ch.N;
  if (nslice) *nslice = ch.M;
  if (ndim) *ndim = ch.D;
  if (off) *off = ch.bead_off;
  return PQ_OK;
}

extern "C" int pq_set_paths(pq_batch* b, int c, const double* beads) {
  if (int rc = check_chain(b, c)) return rc;
  if (!beads) return fail(PQ_ERR_ARG, "null beads");
  const ChainP& ch = b->hc[c];
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(b->dbeads + ch.bead_off, beads, sizeof(double) * ch.M * ch.N * ch.D, hipMemcpyHostToDevice));
  // new paths: the chain's node records no longer describe them, and the
  // permutation starts from the identity (PathReader::run clears it)
  if (ch.fnc_off >= 0) HIPCHK(hipMemset(b->dfncache + ch.fnc_off, 0, sizeof(double) * fn_cache_doubles(ch)));
  std::vector<int> ident(2 * (size_t)ch.N);
  for (int i = 0; i < ch.N; ++i) ident[i] = ident[ch.N + i] = i;
  HIPCHK(hipMemcpy(b->dperm + ch.perm_off, ident.data(), sizeof(int) * ident.size(), hipMemcpyHostToDevice));
  return PQ_OK;
}

extern "C" int pq_set_permutation(pq_batch* b, int c, const int* perm) {
  if (int rc = check_chain(b, c)) return rc;
  if (!perm) return fail(PQ_ERR_ARG, "null permutation");
  const ChainP& ch = b->hc[c];
  std::vector<int> pi(2 * (size_t)ch.N, -1);
  for (int i = 0; i < ch.N; ++i) {
    if (perm[i] < 0 || perm[i] >= ch.N || pi[ch.N + perm[i]] >= 0)
      return fail(PQ_ERR_ARG, "chain %d: not a permutation of 0..%d", c, ch.N - 1);
    pi[i] = perm[i];
    pi[ch.N + perm[i]] = i;
  }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(b->dperm + ch.perm_off, pi.data(), sizeof(int) * pi.size(), hipMemcpyHostToDevice));
  if (ch.fnc_off >= 0) HIPCHK(hipMemset(b->dfncache + ch.fnc_off, 0, sizeof(double) * fn_cache_doubles(ch)));
  return PQ_OK;
}

extern "C" int pq_get_permutation(pq_batch* b, int c, int* perm) {
  if (int rc = check_chain(b, c)) return rc;
  if (!perm) return fail(PQ_ERR_ARG, "null permutation");
  const ChainP& ch = b->hc[c];
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(perm, b->dperm + ch.perm_off, sizeof(int) * ch.N, hipMemcpyDeviceToHost));
  return PQ_OK;
}

extern "C" int pq_get_paths(pq_batch* b, int c, double* beads) {
  if (int rc = check_chain(b, c)) return rc;
  if (!beads) return fail(PQ_ERR_ARG, "null beads");
  const ChainP& ch = b->hc[c];
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(beads, b->dbeads + ch.bead_off, sizeof(double) * ch.M * ch.N * ch.D, hipMemcpyDeviceToHost));
  return PQ_OK;
}

extern "C" int pq_set_all_paths(pq_batch* b, const double* beads, int64_t count) {
  if (!b || !beads || count != b->nbeads) return fail(PQ_ERR_ARG, "pq_set_all_paths: expected %lld values", (long long)(b ? b->nbeads : 0));
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(b->dbeads, beads, sizeof(double) * count, hipMemcpyHostToDevice));
  if (b->nfncache) HIPCHK(hipMemset(b->dfncache, 0, sizeof(double) * b->nfncache));
  std::vector<int> ident((size_t)b->nperm);
  for (const ChainP& ch : b->hc)
    for (int i = 0; i < ch.N; ++i) ident[ch.perm_off + i] = ident[ch.perm_off + ch.N + i] = i;
  HIPCHK(hipMemcpy(b->dperm, ident.data(), sizeof(int) * ident.size(), hipMemcpyHostToDevice));
  return PQ_OK;
}

extern "C" int pq_get_all_paths(pq_batch* b, double* beads, int64_t count) {
  if (!b || !beads || count != b->nbeads) return fail(PQ_ERR_ARG, "pq_get_all_paths: expected %lld values", (long long)(b ? b->nbeads : 0));
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(beads, b->dbeads, sizeof(double) * count, hipMemcpyDeviceToHost));
  return PQ_OK;
}

// ------------------------------------------------------------- path files

// WritePaths::run (src/algorithm/WritePaths.cc:36-112) for one clone on one
// worker: "#Permutations " and the permutation (the identity when perm is
// null), each entry followed by a space; "#Path coordinates: N particles in
// D-d"; then every slice on one line, each coordinate followed by a space,
// at the stream's precision 15 (printf's %.15g).
extern "C" int pq_write_paths_file(const char* filename, const double* beads, const int* perm, int nslice,
                                   int npart, int ndim) {
  if (!filename || !beads || nslice < 1 || npart < 1 || ndim < 1) return fail(PQ_ERR_ARG, "pq_write_paths_file: bad arguments");
  FILE* f = std::fopen(filename, "w");
  if (!f) return fail(PQ_ERR_ARG, "cannot open %s for writing", filename);
  std::fprintf(f, "#Permutations ");
  for (int i = 0; i < npart; ++i) std::fprintf(f, "%d ", perm ? perm[i] : i);
  std::fprintf(f, "\n#Path coordinates: %d particles in %d-d\n", npart, ndim);
  for (int t = 0; t < nslice; ++t) {
    for (int i = 0; i < npart; ++i)
      for (int d = 0; d < ndim; ++d) std::fprintf(f, "%.15g ", beads[((size_t)t * npart + i) * ndim + d]);
    std::fprintf(f, "\n");
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? PQ_OK : fail(PQ_ERR_ARG, "writing %s failed", filename);
}

// PathReader::run (src/algorithm/PathReader.cc:17-132) for one clone on one
// worker: an optional "#Permutations" line (npart integers), the header line,
// then nslice/bfactor slices of npart*ndim numbers; slice s of the file fills
// slices s + k nslice/bfactor, k < bfactor.  A file without the permutation
// line carries one more slice, whose match against the first slice
// (|difference| summed over dimensions < 1e-5) gives the permutation p.
// The last slice of each of the bfactor copies is put back with p
// (PathReader.cc:113-117), last copy first: copy k ends up relabelled k times,
// bead(i, t) = file bead(p^k[i], t), and the path permutation is p^bfactor.
// With perm_out null a non-identity permutation is refused.
extern "C" int pq_read_paths_file(const char* filename, double* beads, int* perm_out, int nslice, int npart,
                                  int ndim, int bfactor) {
  if (!filename || !beads || nslice < 1 || npart < 1 || ndim < 1) return fail(PQ_ERR_ARG, "pq_read_paths_file: bad arguments");
  if (bfactor <= 0) bfactor = 1;
  if (nslice % bfactor) return fail(PQ_ERR_ARG, "bfactor %d does not divide %d slices", bfactor, nslice);
  FILE* f = std::fopen(filename, "r");
  if (!f) return fail(PQ_ERR_ARG, "cannot open %s", filename);
  auto bad = [&](const char* what) { std::fclose(f); return fail(PQ_ERR_ARG, "%s: %s", filename, what); };
  char tok[256];
  if (std::fscanf(f, "%255s", tok) != 1) return bad("empty file");
  std::vector<int> perm(npart);
  bool have_perm = false;
  if (std::strcmp(tok, "#Permutations") == 0) {
    have_perm = true;
    for (int i = 0; i < npart; ++i)
      if (std::fscanf(f, "%d", &perm[i]) != 1) return bad("short permutation line");
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') {}
  }
  // the rest of the header line, model state included (the token read first
  // opens it when there is no permutation line)
  {
    if (have_perm && std::fscanf(f, "%255s", tok) != 1) return bad("no header line");
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') {}
  }
  const int nf = nslice / bfactor;
  std::vector<double> fs((size_t)nf * npart * ndim);
  for (size_t q = 0; q < fs.size(); ++q)
    if (std::fscanf(f, "%lf", &fs[q]) != 1) return bad("fewer coordinates than slices x particles x dimensions");
  if (!have_perm) {
    std::vector<double> wrap((size_t)npart * ndim);
    for (size_t q = 0; q < wrap.size(); ++q)
      if (std::fscanf(f, "%lf", &wrap[q]) != 1) return bad("no wrap slice and no #Permutations line");
    for (int i = 0; i < npart; ++i) {
      perm[i] = i;
      for (int j = 0; j < npart; ++j) {
        double sum = 0;
        for (int d = 0; d < ndim; ++d) sum += std::fabs(wrap[(size_t)i * ndim + d] - fs[(size_t)j * ndim + d]);
        if (sum < 10e-6) perm[i] = j;
      }
    }
  }
  std::fclose(f);
  {
    std::vector<int> seen(npart, 0);
    for (int i = 0; i < npart; ++i) {
      if (perm[i] < 0 || perm[i] >= npart || seen[perm[i]]) return fail(PQ_ERR_ARG, "%s: not a permutation", filename);
      seen[perm[i]] = 1;
    }
  }
  if (!perm_out)
    for (int i = 0; i < npart; ++i)
      if (perm[i] != i) return fail(PQ_ERR_ARG, "%s: permuted paths need a permutation output (particle %d -> %d)", filename, i, perm[i]);
  // pk = p^k, the relabelling of copy k
  std::vector<int> pk(npart), tmp(npart);
  for (int i = 0; i < npart; ++i) pk[i] = i;
  for (int k = 0; k < bfactor; ++k) {
    double* dst = beads + (size_t)k * nf * npart * ndim;
    for (int t = 0; t < nf; ++t)
      for (int i = 0; i < npart; ++i)
        std::memcpy(dst + ((size_t)t * npart + i) * ndim, &fs[((size_t)t * npart + pk[i]) * ndim], sizeof(double) * ndim);
    for (int i = 0; i < npart; ++i) tmp[i] = pk[perm[i]];
    pk = tmp;
  }
  if (perm_out)
    for (int i = 0; i < npart; ++i) perm_out[i] = pk[i];
  return PQ_OK;
}

// AsciiReportBuilder (src/stats/ascii/AsciiReportBuilder.cc:28-43) with
// AsciiScalarReportWriter (AsciiScalarReportWriter.cpp:12-35): "#", then for
// each scalar estimator a quoted name with " (unit)" when it has a unit, each
// followed by a space; then one line per block, each value (value + shift) *
// scale followed by a space at the stream's default precision (%g).  Array and
// acceptance estimators write nothing to this file (their writers are null).
extern "C" int pq_write_ascii_report(const char* filename, int nscalar, const char* const* names,
                                     const char* const* units, const double* shift, const double* scale,
                                     int nblock, const double* values) {
  if (!filename || nscalar < 1 || !names || nblock < 0 || (nblock > 0 && !values))
    return fail(PQ_ERR_ARG, "pq_write_ascii_report: bad arguments");
  for (int k = 0; k < nscalar; ++k)
    if (!names[k]) return fail(PQ_ERR_ARG, "pq_write_ascii_report: estimator %d has no name", k);
  FILE* f = std::fopen(filename, "w");
  if (!f) return fail(PQ_ERR_ARG, "cannot open %s for writing", filename);
  std::fprintf(f, "#");
  for (int k = 0; k < nscalar; ++k) {
    std::fprintf(f, "\"%s", names[k]);
    if (units && units[k] && units[k][0]) std::fprintf(f, " (%s)", units[k]);
    std::fprintf(f, "\" ");
  }
  std::fprintf(f, "\n");
  for (int blk = 0; blk < nblock; ++blk) {
    for (int k = 0; k < nscalar; ++k) {
      const double v = (values[(size_t)blk * nscalar + k] + (shift ? shift[k] : 0.0)) * (scale ? scale[k] : 1.0);
      std::fprintf(f, "%g ", v);
    }
    std::fprintf(f, "\n");
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? PQ_OK : fail(PQ_ERR_ARG, "writing %s failed", filename);
}

static std::string clone_file(const char* base, int c, int nchain) {
  return nchain > 1 ? std::string(base) + std::to_string(c) : std::string(base);
}

extern "C" int pq_write_all_paths(pq_batch* b, const char* filename) {
  if (!b || !filename) return fail(PQ_ERR_ARG, "pq_write_all_paths: bad arguments");
  for (int c = 0; c < b->nchain; ++c) {
    const ChainP& ch = b->hc[c];
    std::vector<double> host((size_t)ch.M * ch.N * ch.D);
    std::vector<int> perm(ch.N);
    if (int rc = pq_get_paths(b, c, host.data())) return rc;
    if (int rc = pq_get_permutation(b, c, perm.data())) return rc;
    if (int rc = pq_write_paths_file(clone_file(filename, c, b->nchain).c_str(), host.data(), perm.data(), ch.M, ch.N, ch.D))
      return rc;
  }
  return PQ_OK;
}

extern "C" int pq_read_all_paths(pq_batch* b, const char* filename, int bfactor) {
  if (!b || !filename) return fail(PQ_ERR_ARG, "pq_read_all_paths: bad arguments");
  for (int c = 0; c < b->nchain; ++c) {
    const ChainP& ch = b->hc[c];
    std::vector<double> host((size_t)ch.M * ch.N * ch.D);
    std::vector<int> perm(ch.N);
    if (int rc = pq_read_paths_file(clone_file(filename, c, b->nchain).c_str(), host.data(), perm.data(), ch.M, ch.N,
                                    ch.D, bfactor))
      return rc;
    if (int rc = pq_set_paths(b, c, host.data())) return rc;
    if (int rc = pq_set_permutation(b, c, perm.data())) return rc;
  }
  return PQ_OK;
}

extern "C" int pq_seed(pq_batch* b, uint64_t seed, uint32_t first_stream) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->rng_mode = PQ_RNG_PHILOX;
  b->seed = seed;
  b->first_stream = first_stream;
  HIPCHK(hipMemset(b->dcounter, 0, sizeof(uint64_t) * b->nchain));
  return PQ_OK;
}

extern "C" int pq_set_uniforms(pq_batch* b, const double* u, int64_t per_chain) {
  if (!b || !u || per_chain < 1) return fail(PQ_ERR_ARG, "pq_set_uniforms: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->dinj) (void)hipFree(b->dinj);
  b->dinj = nullptr;
  HIPCHK(hipMalloc(&b->dinj, sizeof(double) * per_chain * b->nchain));
  HIPCHK(hipMemcpy(b->dinj, u, sizeof(double) * per_chain * b->nchain, hipMemcpyHostToDevice));
  b->inj_per_chain = per_chain;
  b->rng_mode = PQ_RNG_INJECTED;
  HIPCHK(hipMemset(b->dcounter, 0, sizeof(uint64_t) * b->nchain));
  return PQ_OK;
}

static int apply_schedule(pq_batch* b, const Sched& s) {
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  const Sched keep = b->sched;
  const SweepLayoutH keep_lay = b->lay, keep_heavy = b->lay_heavy;
  b->sched = s;
  make_layout(b, s);
  if (layout_bytes(b) > 160 * 1024) {
    const int n = layout_bytes(b);
    b->sched = keep;
    b->lay = keep_lay;
    b->lay_heavy = keep_heavy;
    return fail(PQ_ERR_ARG, "schedule needs %d bytes of LDS", n);
  }
  return PQ_OK;
}

extern "C" int pq_set_schedule(pq_batch* b, int team, int inv_lds) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (team != -1 && team != 64 && team != 128 && team != 256 && team != 512)
    return fail(PQ_ERR_ARG, "team must be 64, 128, 256, 512 or -1");
  if (inv_lds < -1 || inv_lds > 1) return fail(PQ_ERR_ARG, "inv_lds must be -1, 0 or 1");
  Sched s = b->sched;
  s.team = team;
  s.inv = inv_lds;
  return apply_schedule(b, s);
}

extern "C" int pq_get_schedule(const pq_batch* b, int* team, int* inv_lds, int* lds_bytes) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (team) *team = b->lay.team;
  if (inv_lds) *inv_lds = b->lay.off_inv >= 0;
  if (lds_bytes) *lds_bytes = b->lay.bytes;
  return PQ_OK;
}

extern "C" int pq_set_staging(pq_batch* b, int tables_lds, int state_lds) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (tables_lds < -1 || tables_lds > 1 || state_lds < -1 || state_lds > 1)
    return fail(PQ_ERR_ARG, "staging requests must be -1, 0 or 1");
  Sched s = b->sched;
  s.tab = tables_lds;
  s.fst = state_lds;
  return apply_schedule(b, s);
}

extern "C" int pq_get_staging(const pq_batch* b, int* tables_lds, int* state_lds) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (tables_lds) *tables_lds = b->lay.off_tab >= 0;
  if (state_lds) *state_lds = b->lay.off_fst >= 0;
  return PQ_OK;
}

extern "C" int pq_set_split(pq_batch* b, int heavy_chains, int heavy_team) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (heavy_chains < -1 || heavy_chains >= b->nchain) return fail(PQ_ERR_ARG, "heavy_chains must be in [-1, nchain)");
  if (heavy_chains == -1 && b->wide == 0)
    return fail(PQ_ERR_ARG, "the node-width split needs 32-wide node species beside narrower ones");
  if (heavy_team != 128 && heavy_team != 256 && heavy_team != 512) return fail(PQ_ERR_ARG, "heavy_team must be 128, 256 or 512");
  if (b->multi && heavy_chains != 0 && heavy_team > 128)
    return fail(PQ_ERR_ARG, "batches with several moving particles per Sample run teams of at most 128");
  Sched s = b->sched;
  s.ncsplit = heavy_chains == -1;
  s.heavy = s.ncsplit ? b->wide : heavy_chains;
  s.heavy_team = heavy_team;
  return apply_schedule(b, s);
}

extern "C" int pq_get_split(const pq_batch* b, int* heavy_chains, int* heavy_team) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (heavy_chains) *heavy_chains = b->sched.ncsplit ? -1 : b->sched.heavy;
  if (heavy_team) *heavy_team = b->sched.heavy_team;
  return PQ_OK;
}

extern "C" int pq_autotune(pq_batch* b, int nsweep, float* best_ms, int* ntimed) {
  if (!b || nsweep < 1) return fail(PQ_ERR_ARG, "pq_autotune: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  std::vector<std::vector<char>> seen;
  std::vector<std::pair<float, Sched>> bases, all;
  float best = -1;
  Sched best_s = b->sched;
  int timed = 0;
  // one untimed launch first: a fresh batch has no node records yet, and its
  // first sweep evaluates every section from scratch
  {
    DevCtx x = make_ctx(b);
    HIPCHK(run_launch(b, x, nsweep));
    HIPCHK(hipStreamSynchronize(b->stream));
    ++timed;
  }
  // time one candidate unless an identical launch configuration was timed
  // one launch of nsweep sweeps under schedule s; ms per sweep
  auto time_one = [&](const Sched& s, float& ms) -> int {
    b->sched = s;
    make_layout(b, s);
    DevCtx x = make_ctx(b);
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(run_launch(b, x, nsweep));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    ms /= nsweep;
    ++timed;
    return PQ_OK;
  };
  auto trial = [&](const Sched& s) -> int {
    b->sched = s;
    make_layout(b, s);
    if (layout_bytes(b) > 160 * 1024) return PQ_OK;
    std::vector<char> key(2 * sizeof(SweepLayoutH) + 2 * sizeof(int));
    std::memcpy(key.data(), &b->lay, sizeof(SweepLayoutH));
    std::memcpy(key.data() + sizeof(SweepLayoutH), &b->lay_heavy, sizeof(SweepLayoutH));
    std::memcpy(key.data() + 2 * sizeof(SweepLayoutH), &s.heavy, sizeof(int));
    std::memcpy(key.data() + 2 * sizeof(SweepLayoutH) + sizeof(int), &s.ncsplit, sizeof(int));
    for (const auto& k : seen)
      if (k == key) return PQ_OK;
    seen.push_back(key);
    float ms = 0;
    if (int rc = time_one(s, ms)) return rc;
    if (s.heavy == 0) bases.emplace_back(ms, s);
    all.emplace_back(ms, s);
    return PQ_OK;
  };
  // every team size and placement in one launch
  for (int team : {64, 128, 256, 512})
    for (int tab : {0, 1})
      if (!b->multi || team <= 128)
      for (int fst : {0, 1})
        for (int inv : {0, 1}) {
          Sched s{};
          s.team = team; s.inv = inv; s.tab = tab; s.fst = fst;
          if (int rc = trial(s)) return rc;
        }
  // then the heaviest chains apart with a larger team, on the eight best
  // placements: with the heavy chains gone the bulk's best placement is often
  // another one (tables in global memory rather than LDS on the uneven batch)
  std::stable_sort(bases.begin(), bases.end(),
                   [](const std::pair<float, Sched>& u, const std::pair<float, Sched>& v) { return u.first < v.first; });
  for (size_t q = 0; q < bases.size() && q < 8 && !b->multi; ++q)
    for (int frac : {16, 8, 4})
      for (int ht : {256, 512}) {
        Sched s = bases[q].second;
        s.ncsplit = 0;
        s.heavy = b->nchain / frac;
        s.heavy_team = ht;
        if (s.heavy < 1 || s.heavy >= b->nchain) continue;
        if (int rc = trial(s)) return rc;
      }
  // and the node-width split: 32-wide chains apart, the rest on the narrower
  // instance with a layout sized to them, at every team size of the rest
  if (b->wide > 0 && !b->multi)
    for (int team : {64, 128, 256})
      for (int ht : {256, 512})
        for (size_t q = 0; q < bases.size() && q < 2; ++q) {
          Sched s = bases[q].second;
          s.team = team;
          s.ncsplit = 1;
          s.heavy = b->wide;
          s.heavy_team = ht;
          if (int rc = trial(s)) return rc;
        }
  // a chain's sweep time drifts with its state by about as much as the
  // schedules differ, so the two fastest single launches and the three
  // fastest splits are timed again, three rounds in turn, and the lowest
  // total wins
  std::stable_sort(all.begin(), all.end(),
                   [](const std::pair<float, Sched>& u, const std::pair<float, Sched>& v) { return u.first < v.first; });
  std::vector<Sched> fin;
  int nsingle = 0, nsplit = 0;
  for (const auto& t : all) {
    if (t.second.heavy == 0 && nsingle < 2) { fin.push_back(t.second); ++nsingle; }
    else if (t.second.heavy > 0 && nsplit < 3) { fin.push_back(t.second); ++nsplit; }
  }
  std::vector<float> tot(fin.size(), 0.f);
  for (int round = 0; round < 3; ++round)
    for (size_t q = 0; q < fin.size(); ++q) {
      float ms = 0;
      if (int rc = time_one(fin[q], ms)) return rc;
      tot[q] += ms;
    }
  for (size_t q = 0; q < fin.size(); ++q)
    if (best < 0 || tot[q] / 3 < best) { best = tot[q] / 3; best_s = fin[q]; }
  b->sched = best_s;
  make_layout(b, best_s);
  if (best_ms) *best_ms = best;
  if (ntimed) *ntimed = timed;
  return PQ_OK;
}

extern "C" int pq_run_sweeps(pq_batch* b, int nsweep) {
  if (!b || nsweep < 0) return fail(PQ_ERR_ARG, "pq_run_sweeps: bad arguments");
  if (nsweep == 0) return PQ_OK;
  HIPCHK(hipSetDevice(b->device));
  DevCtx x = make_ctx(b);
  HIPCHK(run_launch(b, x, nsweep));
  return PQ_OK;
}

extern "C" int pq_run_sweeps_timed(pq_batch* b, int nsweep, float* ms) {
  if (!b || nsweep < 1 || !ms) return fail(PQ_ERR_ARG, "pq_run_sweeps_timed: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  DevCtx x = make_ctx(b);
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  HIPCHK(run_launch(b, x, nsweep));
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return PQ_OK;
}

// After the batch stream has drained: with caller-supplied uniforms, a chain
// that drew past the end of its stream read placeholder values (0.5), so its
// decisions no longer follow the stream; that is an error, not a result.
static int check_stream(pq_batch* b) {
  if (b->rng_mode != PQ_RNG_INJECTED) return PQ_OK;
  std::vector<uint64_t> cnt(b->nchain);
  HIPCHK(hipMemcpy(cnt.data(), b->dcounter, sizeof(uint64_t) * b->nchain, hipMemcpyDeviceToHost));
  for (int c = 0; c < b->nchain; ++c)
    if (cnt[c] > (uint64_t)b->inj_per_chain)
      return fail(PQ_ERR_STATE, "chain %d drew %llu uniforms from an injected stream of %lld", c,
                  (unsigned long long)cnt[c], (long long)b->inj_per_chain);
  return PQ_OK;
}

extern "C" int pq_sync(pq_batch* b) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  return check_stream(b);
}

extern "C" int pq_set_sweep_plan(pq_batch* b, const int* sweeps) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  if (sweeps)
    for (int c = 0; c < b->nchain; ++c)
      if (sweeps[c] < 1 || sweeps[c] > 4096) return fail(PQ_ERR_ARG, "sweep plan of chain %d must be in 1..4096", c);
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (!sweeps) {
    // the captured graphs hold the plan's address
    if (b->dplan) { drop_graph(b); (void)hipFree(b->dplan); }
    b->dplan = nullptr;
    return PQ_OK;
  }
  if (!b->dplan) {
    drop_graph(b);
    HIPCHK(hipMalloc(&b->dplan, sizeof(int) * b->nchain));
  }
  // the kernels read the plan at run time: captured graphs stay valid
  HIPCHK(hipMemcpy(b->dplan, sweeps, sizeof(int) * b->nchain, hipMemcpyHostToDevice));
  return PQ_OK;
}

extern "C" int pq_get_sweep_plan(const pq_batch* b, int* sweeps) {
  if (!b || !sweeps) return fail(PQ_ERR_ARG, "pq_get_sweep_plan: bad arguments");
  if (!b->dplan) {
    for (int c = 0; c < b->nchain; ++c) sweeps[c] = 1;
    return PQ_OK;
  }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(sweeps, b->dplan, sizeof(int) * b->nchain, hipMemcpyDeviceToHost));
  return PQ_OK;
}

extern "C" int pq_clock_hz(const pq_batch* b, double* hz) {
  if (!b || !hz) return fail(PQ_ERR_ARG, "pq_clock_hz: bad arguments");
  int khz = 0;
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
  *hz = 1e3 * khz;
  return PQ_OK;
}

extern "C" int pq_get_stats(pq_batch* b, int c, pq_chain_stats* out) {
  if (int rc = check_chain(b, c)) return rc;
  if (!out) return fail(PQ_ERR_ARG, "null stats");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(out, b->dstats + c, sizeof(pq_chain_stats), hipMemcpyDeviceToHost));
  return check_stream(b);
}

extern "C" int pq_get_all_stats(pq_batch* b, pq_chain_stats* out) {
  if (!b || !out) return fail(PQ_ERR_ARG, "pq_get_all_stats: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(out, b->dstats, sizeof(pq_chain_stats) * b->nchain, hipMemcpyDeviceToHost));
  return check_stream(b);
}

extern "C" int pq_reset_stats(pq_batch* b) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<uint64_t> cnt(b->nchain);
  HIPCHK(hipMemcpy(cnt.data(), b->dcounter, sizeof(uint64_t) * b->nchain, hipMemcpyDeviceToHost));
  std::vector<pq_chain_stats> st(b->nchain);
  std::memset(st.data(), 0, sizeof(pq_chain_stats) * b->nchain);
  for (int c = 0; c < b->nchain; ++c) st[c].draws = cnt[c];
  HIPCHK(hipMemcpy(b->dstats, st.data(), sizeof(pq_chain_stats) * b->nchain, hipMemcpyHostToDevice));
  return PQ_OK;
}

extern "C" int pq_get_node_scale(pq_batch* b, int c, double* scale) {
  if (int rc = check_chain(b, c)) return rc;
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(scale, b->dscale + (int64_t)c * PQ_MAX_NODES, sizeof(double) * b->hc[c].nfn, hipMemcpyDeviceToHost));
  return PQ_OK;
}

extern "C" int pq_measure_energy(pq_batch* b, double* energy) {
  if (!b || !energy) return fail(PQ_ERR_ARG, "pq_measure_energy: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  void* p;
  if (int rc = scratch_buffer(b, sizeof(double) * b->nchain, &p)) return rc;
  DevCtx x = make_ctx(b);
  HIPCHK(launch_energy_items(b->D, x, b->ditems_e, (int)b->items_e.size(), b->dfirst_e, b->nchain, b->slice_e,
                             b->dpart, (double*)p, b->stream));
  HIPCHK(hipMemcpyAsync(energy, p, sizeof(double) * b->nchain, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_measure_coulomb_energy(pq_batch* b, double epsilon, double* energy) {
  if (!b || !energy) return fail(PQ_ERR_ARG, "pq_measure_coulomb_energy: bad arguments");
  if (epsilon == 0) epsilon = 1;
  HIPCHK(hipSetDevice(b->device));
  void* p;
  if (int rc = scratch_buffer(b, sizeof(double) * b->nchain, &p)) return rc;
  DevCtx x = make_ctx(b);
  HIPCHK(launch_coulomb_energy(b->D, x, b->nchain, epsilon, (double*)p, b->stream));
  HIPCHK(hipMemcpyAsync(energy, p, sizeof(double) * b->nchain, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

// EwaldCoulombEstimator (EwaldCoulombEstimator.cc:17-140) with an
// OptEwaldSum (trad false) or a TradEwaldSum of kappa, 0 meaning the parser's
// default sqrt((3.6 N)^(1/6) sqrt(pi) / Lside) (EstimatorParser.cc:192-203).
// One fit per distinct cell (and, for tradEwald, kappa) of the batch, charges
// per chain; tradEwald's 1001-point grid of -f(i rcut/1000) is tabulated on
// the host with the reference's erf.
static int measure_ewald_coulomb(pq_batch* b, double epsilon, double rcut, double kcut, bool trad,
                                 double kappa, double* energy) {
  if (!b || !energy || !(rcut > 0) || !(kcut > 0) || !(kappa >= 0) || !std::isfinite(kappa))
    return fail(PQ_ERR_ARG, "Ewald Coulomb estimator: bad arguments (rcut and kcut must be positive)");
  if (epsilon == 0) epsilon = 1;
  HIPCHK(hipSetDevice(b->device));
  if (rcut != b->est_rcut || kcut != b->est_kcut || (int)trad != b->est_type || kappa != b->est_kappa) {
    std::map<std::tuple<double, double, double, double>, std::tuple<EwaldFit, int64_t, int64_t>> fits;
    std::vector<double> pool, vpool;
    std::vector<EwP> ew(b->nchain);
    std::vector<int64_t> voff(b->nchain, -1);
    for (int c = 0; c < b->nchain; ++c) {
      const pq_chain_desc& d = b->descs[c];
      double kap = 0;
      if (trad) {
        kap = kappa;
        if (kap == 0) {
          int n = 0;
          for (int s2 = 0; s2 < d.nspecies; ++s2) n += d.species[s2].count;
          double lside = 1000000000000000.0;
          for (int i = 0; i < b->D; ++i) lside = std::min(lside, d.cell[i]);
          kap = std::sqrt(std::pow(3.6 * n, 1.0 / 6.0) * std::sqrt(3.1415926535897931) / lside);
        }
      }
      const auto key = std::make_tuple(d.cell[0], d.cell[1], b->D == 3 ? d.cell[2] : 0.0, kap);
      auto it = fits.find(key);
      if (it == fits.end()) {
        EwaldFit e;
        const bool ok = trad ? ewald_build_trad(e, b->D, d.cell, rcut, kcut, kap) : ewald_build(e, b->D, d.cell, rcut, kcut);
        if (!ok) return fail(PQ_ERR_ARG, "the Ewald fit for rcut %g, kcut %g is not finite", rcut, kcut);
        const int64_t off = (int64_t)pool.size();
        for (int k = 0; k < e.nk(); ++k) {
          for (int q = 0; q < 3; ++q) pool.push_back(e.kvec[3 * k + q]);
          pool.push_back(2 * e.vk[k] * e.oneOver2V);
        }
        int64_t vo = -1;
        if (trad) {
          vo = (int64_t)vpool.size();
          const double dr = rcut / 1000;
          vpool.push_back(-e.fr0());
          // 1001 grid points and one more: r just below rcut can round to
          // ig = 1000, which reads point 1001 with a weight of ~0
          for (int i = 1; i <= 1001; ++i) vpool.push_back(-e.fr(i * dr));
        }
        it = fits.emplace(key, std::make_tuple(std::move(e), off, vo)).first;
      }
      const EwaldFit& e = std::get<0>(it->second);
      std::vector<double> q;
      for (int s2 = 0; s2 < d.nspecies; ++s2)
        for (int i = 0; i < d.species[s2].count; ++i) q.push_back(d.species[s2].charge);
      EwP& p = ew[c];
      p.nk = e.nk();
      p.off = std::get<1>(it->second);
      p.self = e.self_energy(q.data(), (int)q.size());
      p.rcut = rcut;
      for (int i = 0; i < kEwaldPoly; ++i) p.coef[i] = e.coef[i];
      voff[c] = std::get<2>(it->second);
    }
    void* olds[] = {b->dew_est, b->dewk_est, b->dvoff_est, b->dvgrid_est};
    for (void* o : olds)
      if (o) (void)hipFree(o);
    b->dew_est = nullptr;
    b->dewk_est = nullptr;
    b->dvoff_est = nullptr;
    b->dvgrid_est = nullptr;
    b->est_rcut = b->est_kcut = b->est_kappa = -1;
    b->est_type = -1;
    HIPCHK(hipMalloc(&b->dew_est, sizeof(EwP) * b->nchain));
    HIPCHK(hipMalloc(&b->dewk_est, sizeof(double) * std::max<size_t>(1, pool.size())));
    HIPCHK(hipMalloc(&b->dvoff_est, sizeof(int64_t) * b->nchain));
    HIPCHK(hipMalloc(&b->dvgrid_est, sizeof(double) * std::max<size_t>(1, vpool.size())));
    HIPCHK(hipMemcpy(b->dew_est, ew.data(), sizeof(EwP) * b->nchain, hipMemcpyHostToDevice));
    if (pool.size()) HIPCHK(hipMemcpy(b->dewk_est, pool.data(), sizeof(double) * pool.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->dvoff_est, voff.data(), sizeof(int64_t) * b->nchain, hipMemcpyHostToDevice));
    if (vpool.size()) HIPCHK(hipMemcpy(b->dvgrid_est, vpool.data(), sizeof(double) * vpool.size(), hipMemcpyHostToDevice));
    b->est_rcut = rcut;
    b->est_kcut = kcut;
    b->est_kappa = kappa;
    b->est_type = trad ? 1 : 0;
  }
  void* p;
  if (int rc = scratch_buffer(b, sizeof(double) * b->nchain, &p)) return rc;
  DevCtx x = make_ctx(b);
  HIPCHK(launch_ewald_coulomb_energy(b->D, x, b->nchain, b->dew_est, b->dewk_est, b->dvoff_est, b->dvgrid_est,
                                     epsilon, (double*)p, b->stream));
  HIPCHK(hipMemcpyAsync(energy, p, sizeof(double) * b->nchain, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_measure_ewald_coulomb_energy(pq_batch* b, double epsilon, double rcut, double kcut,
                                               double* energy) {
  return measure_ewald_coulomb(b, epsilon, rcut, kcut, false, 0.0, energy);
}

extern "C" int pq_measure_trad_ewald_coulomb_energy(pq_batch* b, double epsilon, double rcut, double kcut,
                                                    double kappa, double* energy) {
  return measure_ewald_coulomb(b, epsilon, rcut, kcut, true, kappa, energy);
}

extern "C" int pq_measure_virial_energy(pq_batch* b, int nwindow, double* energy) {
  if (!b || !energy) return fail(PQ_ERR_ARG, "pq_measure_virial_energy: bad arguments");
  if (nwindow <= 0) nwindow = 1;
  HIPCHK(hipSetDevice(b->device));
  void* p;
  if (int rc = scratch_buffer(b, sizeof(double) * b->nchain, &p)) return rc;
  DevCtx x = make_ctx(b);
  HIPCHK(launch_virial_energy(b->D, x, b->nchain, nwindow, (double*)p, b->stream));
  HIPCHK(hipMemcpyAsync(energy, p, sizeof(double) * b->nchain, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

static void drop_graph(pq_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->sexec) (void)hipGraphExecDestroy(b->sexec);
  if (b->sgraph) (void)hipGraphDestroy(b->sgraph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->sexec = nullptr;
  b->sgraph = nullptr;
  b->gkey.clear();
  b->skey.clear();
}

// The Measure of block blk: thermodynamic energies into dmeas[blk] and, when
// one is set, the pair-correlation counts added to dcf.
static hipError_t launch_measure(pq_batch* b, const DevCtx& x, int blk) {
  hipError_t e = launch_energy_items(b->D, x, b->ditems_e, (int)b->items_e.size(), b->dfirst_e, b->nchain, b->slice_e,
                                     b->dpart, b->dmeas + (size_t)blk * b->nchain, b->stream);
  if (e == hipSuccess && b->cf.nsp > 0)
    e = launch_pair_cf_items(b->D, x, b->ditems_g, (int)b->items_g.size(), b->slice_g, b->cf, b->dcf, b->stream);
  return e;
}

// Captures the kernel launches that fn enqueues on the batch stream as a graph.
template <class F>
static int capture(pq_batch* b, hipGraph_t* graph, hipGraphExec_t* exec, F&& fn) {
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t e = fn();
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(PQ_ERR_HIP, "capturing the measured sweeps failed: %s", hipGetErrorString(e != hipSuccess ? e : ee));
  }
  *graph = g;
  HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return PQ_OK;
}

// nblock x (nsweep sweeps, then the Measure: ThermoEnergyEstimator on every
// chain and, when one is set, the pair-correlation estimator) on the batch
// stream: the Loop/Measure pattern of a run without a host round trip per
// block.  Untimed, the whole run is one HIP graph, captured on the first call
// and replayed while the batch's kernel arguments (schedule, seed, uniform
// source, the estimator) and (nblock, nsweep) stay the same.  Timed
// (sweep_ms given), each block launches a captured graph of its sweeps between
// two HIP events, whose elapsed time is that block's sampler time, and then
// its Measure kernels (HIP does not time events recorded inside a graph).
static int run_measured(pq_batch* b, int nblock, int nsweep, double* energies, float* sweep_ms) {
  if (!b || nblock < 1 || nsweep < 1 || !energies) return fail(PQ_ERR_ARG, "pq_run_measured: bad arguments");
  HIPCHK(hipSetDevice(b->device));
  const size_t need = (size_t)nblock * b->nchain;
  const bool timed = sweep_ms != nullptr;
  if (need > b->nmeas) {
    // the whole-run graph writes into dmeas; the sweep graph does not
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gkey.clear();
    if (b->dmeas) (void)hipFree(b->dmeas);
    b->dmeas = nullptr;
    b->nmeas = 0;
    HIPCHK(hipMalloc(&b->dmeas, sizeof(double) * need));
    b->nmeas = need;
  }
  while (timed && b->gev.size() < 2 * (size_t)nblock) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    b->gev.push_back(e);
  }
  DevCtx x = make_ctx(b);
  // the kernel arguments a graph bakes in (field by field: no padding bytes)
  const uint64_t fields[] = {
      (uint64_t)(uintptr_t)x.beads, (uint64_t)(uintptr_t)x.inj, (uint64_t)x.inj_per_chain,
      (uint64_t)x.rng_mode, x.seed, (uint64_t)x.first_stream, (uint64_t)(uintptr_t)x.inv,
      (uint64_t)(uintptr_t)x.fncache, (uint64_t)(uintptr_t)x.fnsweep, (uint64_t)nblock, (uint64_t)nsweep,
      (uint64_t)b->sched.heavy, (uint64_t)b->sched.ncsplit, (uint64_t)(uintptr_t)b->dcf,
      (uint64_t)b->cf.nsp, (uint64_t)b->cf.nbin, (uint64_t)(uintptr_t)b->dmeas, (uint64_t)(uintptr_t)b->dplan};
  std::vector<char> key(sizeof(fields) + 2 * sizeof(SweepLayoutH) + sizeof(b->cf.sp) + 2 * sizeof(double));
  char* kp = key.data();
  std::memcpy(kp, fields, sizeof(fields)); kp += sizeof(fields);
  std::memcpy(kp, &b->lay, sizeof(SweepLayoutH)); kp += sizeof(SweepLayoutH);
  std::memcpy(kp, &b->lay_heavy, sizeof(SweepLayoutH)); kp += sizeof(SweepLayoutH);
  std::memcpy(kp, b->cf.sp, sizeof(b->cf.sp)); kp += sizeof(b->cf.sp);
  std::memcpy(kp, &b->cf.rmin, sizeof(double)); kp += sizeof(double);
  std::memcpy(kp, &b->cf.rmax, sizeof(double));
  if (timed) {
    // the sweep graph depends on neither nblock nor the Measure's buffers
    std::vector<char> skey(key);
    uint64_t* sf = reinterpret_cast<uint64_t*>(skey.data());
    sf[9] = 0;                                    // nblock
    for (int k = 13; k < 17; ++k) sf[k] = 0;      // dcf, cf.nsp, cf.nbin, dmeas
    std::memset(skey.data() + sizeof(fields) + 2 * sizeof(SweepLayoutH), 0, sizeof(b->cf.sp) + 2 * sizeof(double));
    if (!b->sexec || skey != b->skey) {
      if (b->sexec) (void)hipGraphExecDestroy(b->sexec);
      if (b->sgraph) (void)hipGraphDestroy(b->sgraph);
      b->sexec = nullptr;
      b->sgraph = nullptr;
      b->skey.clear();
      if (int rc = capture(b, &b->sgraph, &b->sexec, [&] { return run_launch(b, x, nsweep); })) return rc;
      b->skey = skey;
    }
    for (int blk = 0; blk < nblock; ++blk) {
      HIPCHK(hipEventRecord(b->gev[2 * blk], b->stream));
      HIPCHK(hipGraphLaunch(b->sexec, b->stream));
      HIPCHK(hipEventRecord(b->gev[2 * blk + 1], b->stream));
      HIPCHK(launch_measure(b, x, blk));
    }
  } else {
    if (!b->gexec || key != b->gkey) {
      if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->gexec = nullptr;
      b->graph = nullptr;
      b->gkey.clear();
      if (int rc = capture(b, &b->graph, &b->gexec, [&] {
            hipError_t e = hipSuccess;
            for (int blk = 0; blk < nblock && e == hipSuccess; ++blk) {
              e = run_launch(b, x, nsweep);
              if (e == hipSuccess) e = launch_measure(b, x, blk);
            }
            return e;
          }))
        return rc;
      b->gkey = key;
    }
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
  }
  HIPCHK(hipMemcpyAsync(energies, b->dmeas, sizeof(double) * need, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (timed)
    for (int blk = 0; blk < nblock; ++blk) HIPCHK(hipEventElapsedTime(&sweep_ms[blk], b->gev[2 * blk], b->gev[2 * blk + 1]));
  return check_stream(b);
}

extern "C" int pq_run_measured(pq_batch* b, int nblock, int nsweep, double* energies) {
  return run_measured(b, nblock, nsweep, energies, nullptr);
}

extern "C" int pq_run_measured_timed(pq_batch* b, int nblock, int nsweep, double* energies, float* sweep_ms) {
  if (!sweep_ms) return fail(PQ_ERR_ARG, "pq_run_measured_timed: null sweep_ms");
  return run_measured(b, nblock, nsweep, energies, sweep_ms);
}

static int cf_args(int nspecies, const int* species, double rmin, double rmax, int nbin, PairCfArgs* g) {
  if (nspecies < 1 || nspecies > PQ_MAX_CF_SPECIES || !species)
    return fail(PQ_ERR_ARG, "pair correlation: nspecies must be in 1..%d", PQ_MAX_CF_SPECIES);
  if (nbin < 1 || nbin > 16384 || !(rmax > rmin) || !std::isfinite(rmin) || !std::isfinite(rmax))
    return fail(PQ_ERR_ARG, "pair correlation: need 1 <= nbin <= 16384 and finite rmin < rmax");
  std::memset(g, 0, sizeof *g);
  g->nsp = nspecies;
  for (int k = 0; k < nspecies; ++k) {
    if (species[k] < 0 || species[k] >= PQ_MAX_SPECIES) return fail(PQ_ERR_ARG, "pair correlation: species out of range");
    g->sp[k] = species[k];
  }
  g->nbin = nbin;
  g->rmin = rmin;
  g->rmax = rmax;
  g->deltaInv = nbin / (rmax - rmin);    // PairCFEstimator's deltaInv = nbin / (max - min)
  return PQ_OK;
}

extern "C" int pq_pair_cf_histogram(pq_batch* b, int nspecies, const int* species, double rmin, double rmax,
                                    int nbin, uint64_t* counts) {
  if (!b || !counts) return fail(PQ_ERR_ARG, "pq_pair_cf_histogram: bad arguments");
  PairCfArgs g;
  if (int rc = cf_args(nspecies, species, rmin, rmax, nbin, &g)) return rc;
  HIPCHK(hipSetDevice(b->device));
  void* p;
  const size_t bytes = sizeof(uint64_t) * (size_t)nbin * b->nchain;
  if (int rc = scratch_buffer(b, bytes, &p)) return rc;
  HIPCHK(hipMemcpyAsync(p, counts, bytes, hipMemcpyHostToDevice, b->stream));
  DevCtx x = make_ctx(b);
  HIPCHK(launch_pair_cf_items(b->D, x, b->ditems_g, (int)b->items_g.size(), b->slice_g, g, (uint64_t*)p, b->stream));
  HIPCHK(hipMemcpyAsync(counts, p, bytes, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_pair_histogram(pq_batch* b, int s1, int s2, double rmin, double rmax, int nbin,
                                 uint64_t* counts) {
  const int sp[2] = {s1, s2};
  return pq_pair_cf_histogram(b, 2, sp, rmin, rmax, nbin, counts);
}

extern "C" int pq_set_measure_pair_cf(pq_batch* b, int nspecies, const int* species, double rmin, double rmax,
                                      int nbin) {
  if (!b) return fail(PQ_ERR_ARG, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  drop_graph(b);
  if (b->dcf) (void)hipFree(b->dcf);
  b->dcf = nullptr;
  std::memset(&b->cf, 0, sizeof b->cf);
  if (nspecies == 0) return PQ_OK;
  PairCfArgs g;
  if (int rc = cf_args(nspecies, species, rmin, rmax, nbin, &g)) return rc;
  const size_t bytes = sizeof(uint64_t) * (size_t)nbin * b->nchain;
  HIPCHK(hipMalloc(&b->dcf, bytes));
  HIPCHK(hipMemset(b->dcf, 0, bytes));
  b->cf = g;
  return PQ_OK;
}

extern "C" int pq_collect_pair_cf(pq_batch* b, uint64_t* counts) {
  if (!b || !counts) return fail(PQ_ERR_ARG, "pq_collect_pair_cf: bad arguments");
  if (b->cf.nsp == 0) return fail(PQ_ERR_STATE, "no pair-correlation estimator is set");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  const size_t n = (size_t)b->cf.nbin * b->nchain;
  std::vector<uint64_t> h(n);
  HIPCHK(hipMemcpy(h.data(), b->dcf, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(b->dcf, 0, sizeof(uint64_t) * n));
  for (size_t k = 0; k < n; ++k) counts[k] += h[k];
  return PQ_OK;
}

// the queries of pq_action_difference: chains, levels and moving particles
// in range, no particle twice; the largest section and moving-bead blocks
static int check_queries(pq_batch* b, int nq, const int* chain, const int* level, const double* section,
                         const double* moving, const int* index, int nmov, double* dS, int64_t* sec_stride,
                         int64_t* mov_stride) {
  if (!b || nq < 1 || !chain || !level || !section || !moving || !index || !dS)
    return fail(PQ_ERR_ARG, "pq_action_difference: bad arguments");
  if (nmov < 1 || nmov > PQ_MAX_MOVING) return fail(PQ_ERR_ARG, "pq_action_difference: nmov must be in 1..%d", PQ_MAX_MOVING);
  *sec_stride = *mov_stride = 0;
  for (int q = 0; q < nq; ++q) {
    if (int rc = check_chain(b, chain[q])) return rc;
    const ChainP& ch = b->hc[chain[q]];
    if (level[q] < 0 || level[q] > ch.L) return fail(PQ_ERR_ARG, "query %d: level out of range", q);
    for (int m = 0; m < nmov; ++m) {
      const int i = index[(size_t)q * nmov + m];
      if (i < 0 || i >= ch.N) return fail(PQ_ERR_ARG, "query %d: index out of range", q);
      for (int m2 = 0; m2 < m; ++m2)
        if (index[(size_t)q * nmov + m2] == i) return fail(PQ_ERR_ARG, "query %d: particle %d moves twice", q, i);
    }
    *sec_stride = std::max<int64_t>(*sec_stride, (int64_t)ch.S * ch.N * ch.D);
    *mov_stride = std::max<int64_t>(*mov_stride, (int64_t)ch.S * nmov * ch.D);
  }
  return PQ_OK;
}

extern "C" int pq_action_difference(pq_batch* b, int nq, const int* chain, const int* level,
                                    const double* section, const double* moving, const int* index,
                                    int nmov, double* dS) {
  int64_t sec_stride = 0, mov_stride = 0;
  if (int rc = check_queries(b, nq, chain, level, section, moving, index, nmov, dS, &sec_stride, &mov_stride))
    return rc;
  // queries arrive packed by their own chain's sizes; restage with a fixed stride
  std::vector<double> hs((size_t)nq * sec_stride, 0.0), hm((size_t)nq * mov_stride, 0.0);
  int64_t os = 0, om = 0;
  for (int q = 0; q < nq; ++q) {
    const ChainP& ch = b->hc[chain[q]];
    const int64_t ns = (int64_t)ch.S * ch.N * ch.D, nm = (int64_t)ch.S * nmov * ch.D;
    std::memcpy(&hs[q * sec_stride], section + os, sizeof(double) * ns);
    std::memcpy(&hm[q * mov_stride], moving + om, sizeof(double) * nm);
    os += ns;
    om += nm;
  }
  HIPCHK(hipSetDevice(b->device));
  const size_t bytes = sizeof(double) * (hs.size() + hm.size() + nq) + sizeof(int) * (2 + nmov) * nq + 64;
  void* p;
  if (int rc = scratch_buffer(b, bytes, &p)) return rc;
  double* dsec = (double*)p;
  double* dmov = dsec + hs.size();
  double* dout = dmov + hm.size();
  int* dq = (int*)(dout + nq);
  HIPCHK(hipMemcpyAsync(dsec, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dmov, hm.data(), sizeof(double) * hm.size(), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq, chain, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq + nq, level, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq + 2 * nq, index, sizeof(int) * nq * nmov, hipMemcpyHostToDevice, b->stream));
  CallBufsH cb{dq, dq + nq, dsec, dmov, dq + 2 * nq, nmov, sec_stride, mov_stride, dout};
  DevCtx x = make_ctx(b);
  HIPCHK(launch_action_diff(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, nq, cb, b->stream));
  HIPCHK(hipMemcpyAsync(dS, dout, sizeof(double) * nq, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

// getBeadAction of nq beads of one chain on the paths at dbeads/dperm
// (device pointers): out[nq][3 + 2D] on the host
static int bead_action_dev(pq_batch* b, int c, const double* dbeads, const int* dperm, int nq, const int* ipart,
                           const int* islice, double* out) {
  const ChainP& ch = b->hc[c];
  for (int q = 0; q < nq; ++q) {
    if (ipart[q] < 0 || ipart[q] >= ch.N) return fail(PQ_ERR_ARG, "query %d: particle %d out of range", q, ipart[q]);
    if (islice[q] < 0 || islice[q] >= ch.M) return fail(PQ_ERR_ARG, "query %d: slice %d out of range", q, islice[q]);
  }
  const int w = 3 + 2 * ch.D;
  const size_t need = sizeof(double) * (size_t)nq * w + sizeof(int) * 2 * (size_t)nq + 64;
  void* p;
  if (int rc = scratch_buffer(b, need, &p)) return rc;
  double* dout = (double*)p;
  int* dq = (int*)(dout + (size_t)nq * w);
  HIPCHK(hipMemcpyAsync(dq, ipart, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq + nq, islice, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  DevCtx x = make_ctx(b);
  HIPCHK(launch_bead_action(b->D, x, c, dbeads, dperm, nq, dq, dq + nq, dout, b->stream));
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * (size_t)nq * w, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_bead_action(pq_batch* b, int c, const double* beads, const int* perm, int nq, const int* ipart,
                              const int* islice, double* out) {
  if (int rc = check_chain(b, c)) return rc;
  if (nq < 1 || !ipart || !islice || !out) return fail(PQ_ERR_ARG, "pq_bead_action: bad arguments");
  const ChainP& ch = b->hc[c];
  HIPCHK(hipSetDevice(b->device));
  if (!beads) {
    if (perm) return fail(PQ_ERR_ARG, "pq_bead_action: a permutation goes with caller beads (the device paths keep theirs)");
    return bead_action_dev(b, c, b->dbeads + ch.bead_off, b->dperm + ch.perm_off, nq, ipart, islice, out);
  }
  // the caller's paths: staged in a buffer of their own, with the permutation and its inverse
  const int64_t nb = (int64_t)ch.M * ch.N * ch.D;
  std::vector<int> pi(2 * (size_t)ch.N);
  for (int i = 0; i < ch.N; ++i) pi[i] = perm ? perm[i] : i;
  for (int i = 0; i < ch.N; ++i) {
    if (pi[i] < 0 || pi[i] >= ch.N) return fail(PQ_ERR_ARG, "pq_bead_action: perm is not a permutation");
    pi[ch.N + pi[i]] = i + 1;
  }
  for (int i = 0; i < ch.N; ++i) {
    if (pi[ch.N + i] == 0) return fail(PQ_ERR_ARG, "pq_bead_action: perm is not a permutation");
    pi[ch.N + i] -= 1;
  }
  const size_t need = sizeof(double) * nb + sizeof(int) * pi.size() + 64;
  if (b->bead_stage_bytes < need) {
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->dbead_stage) (void)hipFree(b->dbead_stage);
    b->dbead_stage = nullptr;
    b->bead_stage_bytes = 0;
    HIPCHK(hipMalloc(&b->dbead_stage, need));
    b->bead_stage_bytes = need;
  }
  int* dpi = reinterpret_cast<int*>(b->dbead_stage + nb);
  HIPCHK(hipMemcpyAsync(b->dbead_stage, beads, sizeof(double) * nb, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dpi, pi.data(), sizeof(int) * pi.size(), hipMemcpyHostToDevice, b->stream));
  return bead_action_dev(b, c, b->dbead_stage, dpi, nq, ipart, islice, out);
}

// ----------------------------------------- the entry point that allocates nothing
// The device memory of pq_action_difference_static: a file-scope array on
// every GPU the library runs on, held until the process ends.  In host code
// the name `scratch` is a stand-in, not a device address: host code uses the
// address hipGetSymbolAddress gives for each device, looked up once per
// device (with that device current) when a batch is created there.
#define PQ_STATIC_SCRATCH_BYTES (4 << 20)
#define PQ_MAX_DEVICES 64
static __device__ __align__(256) char scratch[PQ_STATIC_SCRATCH_BYTES];
static void* g_scratch_ptr[PQ_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int static_scratch_setup(int device) {
  if (device < 0 || device >= PQ_MAX_DEVICES) return fail(PQ_ERR_ARG, "device %d beyond %d", device, PQ_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return PQ_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(PQ_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return PQ_OK;
}

extern "C" int pq_action_difference_static(pq_batch* b, int nq, const int* chain, const int* level,
                                           const double* section, const double* moving, const int* index,
                                           int nmov, double* dS) {
  int64_t sec_stride = 0, mov_stride = 0;
  if (int rc = check_queries(b, nq, chain, level, section, moving, index, nmov, dS, &sec_stride, &mov_stride))
    return rc;
  const size_t nsec = (size_t)nq * sec_stride, nmv = (size_t)nq * mov_stride;
  const size_t bytes = sizeof(double) * (nsec + nmv + nq) + sizeof(int) * (2 + (size_t)nmov) * nq;
  if (bytes > PQ_STATIC_SCRATCH_BYTES)
    return fail(PQ_ERR_ARG, "pq_action_difference_static: %d queries need %zu bytes of device scratch, more than its "
                "%d; split them over several calls", nq, bytes, PQ_STATIC_SCRATCH_BYTES);
  void* base;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_ptr[b->device];
  }
  if (!base) return fail(PQ_ERR_STATE, "the static scratch of device %d was not set up", b->device);
  HIPCHK(hipSetDevice(b->device));
  // every byte the kernel reads is written here first: nothing needs clearing
  double* dsec = (double*)base;
  double* dmov = dsec + nsec;
  double* dout = dmov + nmv;
  int* dq = (int*)(dout + nq);
  int64_t os = 0, om = 0;
  for (int q = 0; q < nq; ++q) {
    const ChainP& ch = b->hc[chain[q]];
    const int64_t ns = (int64_t)ch.S * ch.N * ch.D, nm = (int64_t)ch.S * nmov * ch.D;
    HIPCHK(hipMemcpyAsync(dsec + q * sec_stride, section + os, sizeof(double) * ns, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(dmov + q * mov_stride, moving + om, sizeof(double) * nm, hipMemcpyHostToDevice, b->stream));
    os += ns;
    om += nm;
  }
  HIPCHK(hipMemcpyAsync(dq, chain, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq + nq, level, sizeof(int) * nq, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dq + 2 * nq, index, sizeof(int) * nq * nmov, hipMemcpyHostToDevice, b->stream));
  CallBufsH cb{dq, dq + nq, dsec, dmov, dq + 2 * nq, nmov, sec_stride, mov_stride, dout};
  DevCtx x = make_ctx(b);
  HIPCHK(launch_action_diff(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, nq, cb, b->stream));
  HIPCHK(hipMemcpyAsync(dS, dout, sizeof(double) * nq, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_bead_utau(pq_batch* b, int nq, const int* chain, const int* ipart, const int* islice,
                            double* utau) {
  if (!b || nq < 1 || !chain || !ipart || !islice || !utau) return fail(PQ_ERR_ARG, "pq_bead_utau: bad arguments");
  for (int q = 0; q < nq; ++q)
    if (int rc = check_chain(b, chain[q])) return rc;
  // one launch per run of queries on the same chain
  std::vector<double> out;
  for (int q0 = 0; q0 < nq;) {
    int q1 = q0 + 1;
    while (q1 < nq && chain[q1] == chain[q0]) ++q1;
    const int c = chain[q0], w = 3 + 2 * b->hc[c].D;
    out.assign((size_t)(q1 - q0) * w, 0.0);
    if (int rc = pq_bead_action(b, c, nullptr, nullptr, q1 - q0, ipart + q0, islice + q0, out.data())) return rc;
    for (int q = q0; q < q1; ++q) utau[q] = out[(size_t)(q - q0) * w + 1];
    q0 = q1;
  }
  return PQ_OK;
}

// fixed-node protocol on one chain; section buffers live in the scratch block
struct FnScr {
  double *sec1, *sec2, *mov, *wdet, *wscale, *out;
  int *werr, *notmine, *idx;
};

static int fn_scratch(pq_batch* b, int c, FnScr* s) {
  const ChainP& ch = b->hc[c];
  const int64_t sec = (int64_t)ch.S * ch.N * ch.D;
  const size_t bytes = sizeof(double) * (2 * sec + ch.S * PQ_MAX_MOVING * ch.D + PQ_WAVE + ch.S + 8) +
                       sizeof(int) * (PQ_WAVE + 8 + PQ_MAX_MOVING) + 256;
  if (bytes > b->fnscr_bytes) {
    // sized for the largest chain so the buffers survive between calls
    size_t big = 0;
    for (const ChainP& h : b->hc) {
      const int64_t hs = (int64_t)h.S * h.N * h.D;
      big = std::max(big, sizeof(double) * (2 * hs + h.S * PQ_MAX_MOVING * h.D + PQ_WAVE + h.S + 8) +
                              sizeof(int) * (PQ_WAVE + 8 + PQ_MAX_MOVING) + 256);
    }
    if (b->dfnscr) (void)hipFree(b->dfnscr);
    b->dfnscr = nullptr;
    b->fnscr_bytes = 0;
    HIPCHK(hipMalloc(&b->dfnscr, big));
    HIPCHK(hipMemset(b->dfnscr, 0, big));
    b->fnscr_bytes = big;
  }
  double* d = b->dfnscr;
  s->sec1 = d; d += sec;
  s->sec2 = d; d += sec;
  s->mov = d; d += ch.S * PQ_MAX_MOVING * ch.D;
  s->wdet = d; d += PQ_WAVE;
  s->wscale = d; d += ch.S;
  s->out = d; d += 8;
  s->werr = (int*)d;
  s->notmine = s->werr + PQ_WAVE;
  s->idx = s->notmine + 8;
  return PQ_OK;
}

extern "C" int pq_section_begin(pq_batch* b, int c, const double* sec1, const double* sec2) {
  if (int rc = check_chain(b, c)) return rc;
  if (!sec1 || !sec2) return fail(PQ_ERR_ARG, "null section");
  const ChainP& ch = b->hc[c];
  if (!ch.nfn) return fail(PQ_ERR_STATE, "chain %d has no FixedNodeAction", c);
  HIPCHK(hipSetDevice(b->device));
  FnScr s;
  if (int rc = fn_scratch(b, c, &s)) return rc;
  const size_t sec = sizeof(double) * ch.S * ch.N * ch.D;
  HIPCHK(hipMemcpyAsync(s.sec1, sec1, sec, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(s.sec2, sec2, sec, hipMemcpyHostToDevice, b->stream));
  FnBufsH fb{s.sec1, s.sec2, s.mov, s.wdet, s.werr, s.wscale, s.notmine, 0, 0, s.idx, 1, s.out};
  DevCtx x = make_ctx(b);
  HIPCHK(launch_fn(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, c, fb, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_double_action_difference(pq_batch* b, int c, int level, const double* moving,
                                           const int* index, int nmov, double* dS) {
  if (int rc = check_chain(b, c)) return rc;
  if (!moving || !index || !dS) return fail(PQ_ERR_ARG, "pq_double_action_difference: bad arguments");
  if (nmov < 1 || nmov > PQ_MAX_MOVING) return fail(PQ_ERR_ARG, "pq_double_action_difference: nmov must be in 1..%d", PQ_MAX_MOVING);
  const ChainP& ch = b->hc[c];
  if (!ch.nfn) { *dS = 0; return PQ_OK; }
  if (level < 0 || level > ch.L) return fail(PQ_ERR_ARG, "level out of range");
  for (int m = 0; m < nmov; ++m) {
    if (index[m] < 0 || index[m] >= ch.N) return fail(PQ_ERR_ARG, "index out of range");
    for (int m2 = 0; m2 < m; ++m2)
      if (index[m2] == index[m]) return fail(PQ_ERR_ARG, "particle %d moves twice", index[m]);
  }
  HIPCHK(hipSetDevice(b->device));
  FnScr s;
  if (int rc = fn_scratch(b, c, &s)) return rc;
  HIPCHK(hipMemcpyAsync(s.mov, moving, sizeof(double) * ch.S * nmov * ch.D, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(s.idx, index, sizeof(int) * nmov, hipMemcpyHostToDevice, b->stream));
  FnBufsH fb{s.sec1, s.sec2, s.mov, s.wdet, s.werr, s.wscale, s.notmine, 1, level, s.idx, nmov, s.out};
  DevCtx x = make_ctx(b);
  HIPCHK(launch_fn(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, c, fb, b->stream));
  HIPCHK(hipMemcpyAsync(dS, s.out, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_accept_last_move(pq_batch* b, int c) {
  if (int rc = check_chain(b, c)) return rc;
  const ChainP& ch = b->hc[c];
  if (!ch.nfn) return PQ_OK;
  HIPCHK(hipSetDevice(b->device));
  FnScr s;
  if (int rc = fn_scratch(b, c, &s)) return rc;
  FnBufsH fb{s.sec1, s.sec2, s.mov, s.wdet, s.werr, s.wscale, s.notmine, 2, 0, s.idx, 1, s.out};
  DevCtx x = make_ctx(b);
  HIPCHK(launch_fn(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, c, fb, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_node_evaluate(pq_batch* b, int c, int inode, int nmat, const double* r1, const double* r2,
                                double* det, int* err, double* d1, double* d2) {
  if (int rc = check_chain(b, c)) return rc;
  const ChainP& ch = b->hc[c];
  if (inode < 0 || inode >= ch.nfn) return fail(PQ_ERR_ARG, "node action %d out of range", inode);
  if (nmat < 1 || !r1 || !r2 || !det || !err) return fail(PQ_ERR_ARG, "pq_node_evaluate: bad arguments");
  const int ns = ch.fn[inode].ns;
  const int64_t sl = (int64_t)ch.N * ch.D;
  HIPCHK(hipSetDevice(b->device));
  const size_t nd = (size_t)nmat * sl;
  const size_t nsm = (size_t)ch.ns_max;
  const size_t bytes = sizeof(double) * (2 * nd + (size_t)nmat * nsm * nsm * (1 + ch.D) + nmat + (size_t)nmat * 2 * ns) +
                       sizeof(int) * nmat + 256;
  void* p;
  if (int rc = scratch_buffer(b, bytes, &p)) return rc;
  double* dr1 = (double*)p;
  double* dr2 = dr1 + nd;
  double* dinv = dr2 + nd;
  double* degr = dinv + (size_t)nmat * nsm * nsm;
  double* ddet = degr + (size_t)nmat * nsm * nsm * ch.D;
  double* ddist = ddet + nmat;
  int* derr = (int*)(ddist + (size_t)nmat * 2 * ns);
  HIPCHK(hipMemcpyAsync(dr1, r1, sizeof(double) * nd, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(dr2, r2, sizeof(double) * nd, hipMemcpyHostToDevice, b->stream));
  DevCtx x = make_ctx(b);
  HIPCHK(launch_node_eval(b->D, b->NC, b->precision == PQ_PREC_MIXED, x, c, inode, nmat, dr1, dr2, dinv,
                          degr, ddet, derr, ddist, b->stream));
  std::vector<double> hd((size_t)nmat * 2 * ns);
  HIPCHK(hipMemcpyAsync(det, ddet, sizeof(double) * nmat, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(err, derr, sizeof(int) * nmat, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(hd.data(), ddist, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int m = 0; m < nmat; ++m)
    for (int k = 0; k < ns; ++k) {
      if (d1) d1[(size_t)m * ns + k] = hd[((size_t)m * 2 + 0) * ns + k];
      if (d2) d2[(size_t)m * ns + k] = hd[((size_t)m * 2 + 1) * ns + k];
    }
  return PQ_OK;
}

extern "C" int pq_get_pair_table(pq_batch* b, int c, int ipair, int* meta, double* grid, double* u, double* ut) {
  if (int rc = check_chain(b, c)) return rc;
  const ChainP& ch = b->hc[c];
  if (ipair < 0 || ipair >= ch.npair) return fail(PQ_ERR_ARG, "pair %d out of range", ipair);
  const TableP& t = b->ht[ch.pair[ipair].table];
  if (meta) { meta[0] = ch.pair[ipair].s1; meta[1] = ch.pair[ipair].s2; meta[2] = t.K1 - 1; meta[3] = t.ngpts; }
  if (grid) { grid[0] = t.rgridinv; grid[1] = t.logrratioinv; }
  const size_t n = (size_t)t.ngpts * t.K1;
  HIPCHK(hipSetDevice(b->device));
  if (u) HIPCHK(hipMemcpy(u, b->du + t.off, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (ut) HIPCHK(hipMemcpy(ut, b->dut + t.off, sizeof(double) * n, hipMemcpyDeviceToHost));
  return PQ_OK;
}
