// pq_params.h -- device-side parameter blocks shared by the host setup code
// and the HIP kernels.  One ChainP per Markov chain lives in device memory;
// pair tables are deduplicated across chains and stored once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/piqmc_hip.h"

#define PQ_WAVE 64
#define PQ_MAX_PAIRS 10
#define PQ_PG_LEVELS 9      // levels 0..8: a section of up to 2^7 links and the mover's level L+1
#define PQ_GJ32_SCRATCH_DOUBLES 560   // LDS scratch per wave of the 32-wide node-matrix inversion

struct TableP {          // one tabulated pair action u_k(r), utau_k(r)
  int ngpts;
  int K1;                // norder + 1
  double rgridinv;
  double logrratioinv;
  int64_t off;           // element offset of u[ngpts][K1] (and ut) in the table pool
};

struct PairP {
  int s1, s2;
  int table;
  TableP tp;             // a copy of tables[table], read from the chain's LDS block
};

struct FnP {             // one FixedNodeAction with its node model
  int sp, model, ifirst, ns, nodal;
  int manybody;          // useManyBodyDistance
  double mass;
  double coshwt, c;      // SHONodes
  double pg_nome[3], pg_pref[3], pg_k[3];
  int pg_nmax[3];        // FreeParticleNodes periodic Gaussians, alpha = m T
  double pg_alpha[3], pg_len[3];
  int pg_img[3];         // evaluate as the sum of the three nearest images
  int updates;           // FreeParticleNodes::MatrixUpdate (useUpdates)
};

struct EwP {             // the optimized Ewald sum of a CoulombAction or estimator
  int nk;                // half-space k-vectors, 0: no Ewald sum
  int64_t off;           // element offset of [nk][4] (k_x, k_y, k_z, 2 f(k)/(2V)) in the k pool
  double self;           // EwaldSum self energy with the particles' charges
  double rcut;
  double coef[8];        // f(r) = sum_i coef[i] r^2i up to rcut, 1/r beyond
};

#define PQ_GROUPS 4      // <ChooseSection> groups per sweep (PQ_MAX_GROUPS)

struct GroupP {          // one <ChooseSection nlevel> of a sweep with its Samples
  int L, nsection, nsample;
  int sample_sp[8];
  int sample_ch[8];
  int sample_nm[8];
  int sample_rep[8];
  int sample_dr[8];
  double sample_fac[8];
  double sample_nfac[8];
};

struct ChainP {
  int N, D, M, L, S, nsp;
  double tau;
  double a[3], b[3], rcut2;
  int ifirst[4], count[4];
  double mass[4], lambda[4];
  double charge[4];
  int spring;
  int nsho;
  int sho_sp[4], sho_ndim[4];
  double sho_omega[4];
  double sho_center[4][3];
  int npair;
  PairP pair[PQ_MAX_PAIRS];
  int nfn;
  FnP fn[4];
  int mover;
  int nsample;
  int sample_sp[8];
  int sample_ch[8];      // PQ_CHOOSER_* of each Sample
  int sample_nm[8];      // moving particles of each Sample
  int sample_rep[8];     // tries of each Sample on one section (nrepeat)
  int nsection;
  double spring_c[PQ_PG_LEVELS][4];          // 1/(4 lambda tau 2^l) per level, species
  double sho_cosh[4][PQ_PG_LEVELS];          // cosh(omega tau 2^l) per SHO action
  double sho_div1[4][PQ_PG_LEVELS];          // 1/(2 sinh(omega tau 2^l))
  double div2[PQ_PG_LEVELS];                 // 1/(2 tau 2^l)
  unsigned char spg_on[PQ_PG_LEVELS][4][3];   // SpringAction periodic Gaussian used
  unsigned char mpg_on[PQ_PG_LEVELS][4][3];   // mover periodic Gaussian used
  int64_t bead_off;      // element offset of this chain's [M][N][D] beads
  int64_t inv_off;       // element offset of this chain's [S][ns_max][ns_max] inverse cache
  int64_t fn_off;        // element offset of this chain's fixed-node state (global copy)
  int64_t egr_off;       // element offset of this chain's [S][ns_max][ns_max][D] element-gradient cache
  int64_t upd_off;       // element offset of this chain's MatrixUpdate blocks (one per node action), -1 none
  int64_t fnc_off;       // element offset of this chain's slice-pair node records, -1: none
  int64_t perm_off;      // element offset of this chain's path permutation and its inverse ([2][N])
  int ns_max;            // largest fermion species with a node model
  int stat[4];           // species isStatic: no spring action, no virial kinetic terms
  int ngroup;            // section groups of a sweep; group 0 is L, nsection, nsample, sample_*
  int Lmax, Smax;        // largest group: sizes of the sampler's section buffers and caches
  GroupP grp[PQ_GROUPS];
  double ceps;           // CoulombAction epsilon
  EwP ew;                // CoulombAction's Ewald sum (useEwald)
  int sample_dr[8];      // delayed rejection at level L-1
  double sample_fac[8];  // the movers' sigma factor (defaultSigmaFactor)
  double sample_nfac[8]; // sigma factor of the delayed second proposal
};

// Doubles of one slice-pair node record of a species of ns particles:
// determinant, the scale it was evaluated at, a valid flag, and the nodal
// distances of the lower slice's and the upper slice's particles.
__host__ __device__ static inline int fn_record_doubles(int ns) { return 3 + 2 * ns; }

// Slice-pair node records of a chain: every FixedNodeAction keeps M/2 of
// them (slices t and t+M/2 form one pair); 0 when M is odd.
static inline int64_t fn_cache_doubles(const ChainP& c) {
  if (c.M & 1) return 0;
  int64_t n = 0;
  for (int a = 0; a < c.nfn; ++a) n += (int64_t)(c.M / 2) * fn_record_doubles(c.fn[a].ns);
  return n;
}

// FreeParticleNodes::MatrixUpdate block of one node action: a header of 8
// doubles (isNewMatrixUpdated first), the section slices' inverses and updated
// inverses [Smax][ns_max][ns_max], the moving particles' new matrix columns
// [Smax][ns_max][PQ_MAX_MOVING] and per-slice scratch [Smax][64].
__host__ __device__ inline int64_t upd_block_doubles(const ChainP& c) {
  const int64_t ins = c.ns_max > 0 ? c.ns_max : 1;
  return 8 + (int64_t)c.Smax * (2 * ins * ins + PQ_MAX_MOVING * ins + 64);
}
static inline bool chain_has_updates(const ChainP& c) {
  for (int a = 0; a < c.nfn; ++a)
    if (c.fn[a].updates) return true;
  return false;
}

// Sizes of the per-chain fixed-node state block (dm, newdm, dist, newdist per
// FixedNodeAction, [S] and [S][2][ns]).
static inline int64_t fn_state_doubles(const ChainP& c) {
  int64_t n = 0;
  for (int a = 0; a < c.nfn; ++a) n += 2 * c.Smax + 4 * (int64_t)c.Smax * c.fn[a].ns;
  return n;
}
