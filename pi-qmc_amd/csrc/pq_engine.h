// pq_engine.h -- a team of NT threads (one or two wavefronts) evaluates one
// chain's bisection move.
//
// Eng<D,NC,R,NT> bundles the views of one chain (section beads, moving beads,
// fixed-node state, inverse cache) with the routines of one level of
// MultiLevelSampler/DoubleMLSampler::tryMove:
//   mover()            FreeMover / FreeMoverPBC::makeMove
//   action_diff()      CompositeAction::getActionDifference (spring, SHO, pair)
//   fn_initialize()    FixedNodeAction::initialize for every node action
//   fn_diff()          CompositeDoubleAction::getActionDifference
//   fn_accept()        FixedNodeAction::acceptLastMove
// NC bounds the largest fermion species; the node-matrix Gauss-Jordan
// inversion keeps one matrix row per lane in groups of G = gj_group(NC) lanes,
// NT/G matrices per pass.
// All routines are called by every thread of the team with team-uniform
// arguments and return team-uniform results.  They are all force-inlined
// into the sampler kernel: left to itself the inliner outlines the largest
// of them, and each call then saves the callee-saved registers it uses (SGPRs
// through VGPR lanes, VGPRs through scratch) around every level of a move.
#pragma once
#include "pq_device.h"

namespace pq {

// 64-bit DPP lane move inside 16-lane rows (quad_perm / mirror patterns)
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// maximum of v over each aligned group of G lanes (G = 4, 8, 16, 32), every
// lane of the group receives it: quad swaps, half-row and row mirrors, and
// for G = 32 a swizzle across the two rows of a 32-lane half.
template <int G>
__device__ __forceinline__ uint64_t group_max_u64(uint64_t v) {
  uint64_t o = dpp_u64<0xB1>(v);            // quad_perm [1,0,3,2]
  v = o > v ? o : v;
  o = dpp_u64<0x4E>(v);                     // quad_perm [2,3,0,1]
  v = o > v ? o : v;
  if constexpr (G >= 8) {
    o = dpp_u64<0x141>(v);                  // row_half_mirror
    v = o > v ? o : v;
  }
  if constexpr (G >= 16) {
    o = dpp_u64<0x140>(v);                  // row_mirror
    v = o > v ? o : v;
  }
  if constexpr (G >= 32) {
    // v_permlane16_swap leaves each lane the values of itself and of lane ^ 16
    // (in an order that depends on the row), without a trip through LDS
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    const uint64_t a = ((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0];
    const uint64_t b = ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1];
    o = a > b ? a : b;
    v = o > v ? o : v;
  }
  return v;
}

// the value v of lane l (wave-uniform l), through scalar registers
__device__ __forceinline__ double lane_value(double v, int l) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Row-per-lane in-place Gauss-Jordan inversion with partial pivoting of one
// n x n matrix (n <= NC <= 32) per G-lane group; every lane of the wave calls
// it with the same n.  Lane r < NC holds row r of diag(A, I): rows n..NC-1
// are identity rows, which no step k < n touches (their column k is zero) and
// which make the steps k >= n exact no-ops (pivot 1, no swap), so the
// elimination exits only every fourth step and the unrolled steps keep their
// registers.  Lanes r >= NC take no part.  Rows are not moved: each lane
// tracks the position its row would have in the row-swapping elimination
// (mypos), the pivot of step k is the largest |a[k]| among positions >= k
// with the lowest position on ties (LAPACK's idamax), and every virtual swap
// flips the determinant's sign.  The search is one group maximum of a 64-bit
// key, |a[k]|'s bits with the low six replaced by 63 - mypos (non-negative
// doubles order as their bit patterns; a difference in the last six mantissa
// bits counts as a tie).  The pivot row and its reciprocal pivot reach the
// group through LDS (row: NC + 1 doubles per group).  Step k replaces column
// k by the inverse's column as it goes (pivot row: a[j]/piv, a[k] = 1/piv;
// other rows: a[j] - f a_p[j], a[k] = -f, f = a[k]/piv).  On exit the lane
// with mypos = k < n holds row k of A^-1 with its columns in pivot order:
// a[j] is column P(j), P(j) the row that was pivot at step j.  det =
// prod(pivots) * (-1)^swaps, the LU determinant with the same pivots.
template <int G, int NC>
__device__ __forceinline__ void gj_invert(double (&a)[NC], int r, int n, double* row, double& det,
                                          int& err, int& mypos) {
  det = 1.0;
  err = 0;
  mypos = r;
  const bool act = r < NC;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if ((k & 3) == 0 && k >= n) break;
    const uint64_t key = (act && mypos >= k)
        ? ((__double_as_longlong(fabs(a[k])) & ~0x3Full) | (uint64_t)(63 - mypos)) : 0ull;
    // every lane inverts its own candidate while the pivot search runs; the
    // pivot lane's reciprocal travels with its row
    const double myinv = 1.0 / a[k];
    const uint64_t kmax = group_max_u64<G>(key);
    const int ppos = 63 - (int)(kmax & 0x3F);
    const bool isp = act && key == kmax;
    if (ppos != k) det = -det;
    if (isp) {
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) row[cc] = a[cc];
      row[NC] = myinv;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const double piv = row[k];
    det *= piv;
    // a singular column leaves every row unchanged
    const bool zero = piv == 0.0;
    if (zero) err = 1;
    const double ipiv = zero ? 0.0 : row[NC];
    const double f = (act && !isp) ? a[k] * ipiv : 0.0;
    const double sp = (isp && !zero) ? ipiv : 1.0;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
      if (cc != k) a[cc] = fma(-f, row[cc], a[cc]) * sp;
    if (!zero && act) a[k] = isp ? ipiv : -f;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // the row at position k moves to the pivot's old position
    if (isp) mypos = k;
    else if (mypos == k) mypos = ppos;
  }
}

// ----------------------------------------------------------------------------
// Node matrices of up to 32 particles: one wave per matrix, blocked in-place
// Gauss-Jordan with FP64 MFMA trailing updates.
//
// The 32 x 32 matrix diag(A, I) lives in the accumulator layout of
// v_mfma_f64_16x16x4f64, four 16 x 16 tiles: lane l, tile (rt, ct), register r
// holds element (16 rt + 4 r + l / 16, 16 ct + l % 16).  The elimination runs
// in panels of four columns.  A panel is eliminated row-per-lane (lanes 0..31
// are rows) with the pivot rules of gj_invert; its four steps act on any other
// column v as v <- v - W s, W the 32 x 4 multipliers (w_j[p_j] = 1 - 1/piv_j on
// the pivot row, a_rk/piv_j elsewhere) and s the pivot rows' values run
// through the steps before them (s_j = v_pj - sum_{i<j} w_i[p_j] s_i).  That
// rank-4 update of the 28 other columns is four MFMAs; the panel's own
// columns take the row-per-lane result.  Operand layouts: A (16 x 4) lane l
// holds (l % 16, l / 16), B (4 x 16) lane l holds (l / 16, l % 16).
struct GjMfmaScratch {
  double P[32 * 4];    // panel [row][4]
  double W[32 * 4];    // multipliers [row][4]
  double RP[4 * 32];   // pivot rows before the panel [4][column]
  double S[4 * 32];    // s [4][column]
  double spare[8];
  int spare_i[4];
  int pos[32];         // position of each row on exit
  int perm[32];        // P(j): the row that was pivot at step j
};
static_assert(sizeof(GjMfmaScratch) <= PQ_GJ32_SCRATCH_DOUBLES * sizeof(double), "32-wide scratch");

typedef double pq_d4 __attribute__((ext_vector_type(4)));

// element (16 rt + 4 r + lane/16, column tile CT) of the accumulator tiles for
// a wave-uniform k = 4 rt + r, as a chain of value selects (an index into the
// tiles at run time would move them to scratch memory)
template <int CT>
__device__ __forceinline__ double acc_row(const pq_d4 (&acc)[2][2], int k) {
  double v = acc[0][CT][0];
  v = k == 1 ? acc[0][CT][1] : v;
  v = k == 2 ? acc[0][CT][2] : v;
  v = k == 3 ? acc[0][CT][3] : v;
  v = k == 4 ? acc[1][CT][0] : v;
  v = k == 5 ? acc[1][CT][1] : v;
  v = k == 6 ? acc[1][CT][2] : v;
  v = k == 7 ? acc[1][CT][3] : v;
  return v;
}

// One panel (columns 4 kb .. 4 kb + 3, in column tile CT = kb / 4) of the
// 32-wide inversion below.
template <int CT>
__device__ __forceinline__ void gj32_panel(pq_d4 (&acc)[2][2], int kb, GjMfmaScratch* sc, int lane,
                                           int& mypos, double& det, int& err) {
  const int row = lane & 31;
  const bool act = lane < 32;
  const int c0 = 4 * kb;
  const int cc0 = c0 & 15;
  const int jj0 = (lane & 15) - cc0;
  const bool holds = jj0 >= 0 && jj0 < 4;
  // (1) panel columns to LDS, [row][4]
  if (holds) {
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) sc->P[(16 * rt + 4 * r + (lane >> 4)) * 4 + jj0] = acc[rt][CT][r];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // (2) the panel's four steps, row per lane.  The pivot lane is the one
  // lane whose key is the group maximum; its row values and reciprocal
  // reach every lane through scalar registers (readlane).
  double p[4], w[4];
  int pv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) p[j] = act ? sc->P[row * 4 + j] : 0.0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = c0 + j;
    const uint64_t key = (act && mypos >= k)
        ? ((__double_as_longlong(fabs(p[j])) & ~0x3Full) | (uint64_t)(63 - mypos)) : 0ull;
    const double myinv = 1.0 / p[j];
    const uint64_t kmax = group_max_u64<32>(key);
    const int ppos = 63 - (int)(kmax & 0x3F);
    const bool isp = act && key == kmax;
    const int pl = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(isp));
    pv[j] = pl;
    if (ppos != k) det = -det;
    double prow[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) prow[q] = lane_value(p[q], pl);
    const double pinv = lane_value(myinv, pl);
    const double piv = prow[j];
    det *= piv;
    const bool zero = piv == 0.0;
    if (zero) err = 1;
    const double ipiv = zero ? 0.0 : pinv;
    const double f = (act && !isp) ? p[j] * ipiv : 0.0;
    const double sp = (isp && !zero) ? ipiv : 1.0;
    w[j] = (zero || !act) ? 0.0 : (isp ? 1.0 - ipiv : f);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q != j) p[q] = fma(-f, prow[q], p[q]) * sp;
    if (!zero && act) p[j] = isp ? ipiv : -f;
    if (isp) mypos = k;
    else if (mypos == k) mypos = ppos;
  }
  // (3) final panel and multipliers to LDS
  if (act) {
#pragma unroll
    for (int q = 0; q < 4; ++q) { sc->P[row * 4 + q] = p[q]; sc->W[row * 4 + q] = w[q]; }
  }
  // (4) the pivot rows as they were before the panel, [4][column]
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int pr = pv[j];
    if ((lane >> 4) == (pr & 3)) {
      const int k8 = (pr >> 4) * 4 + ((pr & 15) >> 2);
      sc->RP[j * 32 + (lane & 15)] = acc_row<0>(acc, k8);
      sc->RP[j * 32 + 16 + (lane & 15)] = acc_row<1>(acc, k8);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // (5) s for every column outside the panel (lanes 0..31 are columns)
  if (act) {
    const int col = lane;
    const bool inpanel = col >= c0 && col < c0 + 4;
    const int p1 = pv[1], p2 = pv[2], p3 = pv[3];
    const double s0 = sc->RP[col];
    const double s1 = sc->RP[32 + col] - sc->W[p1 * 4 + 0] * s0;
    const double s2 = sc->RP[64 + col] - sc->W[p2 * 4 + 0] * s0 - sc->W[p2 * 4 + 1] * s1;
    const double s3 = sc->RP[96 + col] - sc->W[p3 * 4 + 0] * s0 - sc->W[p3 * 4 + 1] * s1 -
                      sc->W[p3 * 4 + 2] * s2;
    sc->S[col] = inpanel ? 0.0 : s0;
    sc->S[32 + col] = inpanel ? 0.0 : s1;
    sc->S[64 + col] = inpanel ? 0.0 : s2;
    sc->S[96 + col] = inpanel ? 0.0 : s3;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // (6) rank-4 update of every tile: acc -= W s
#pragma unroll
  for (int rt = 0; rt < 2; ++rt) {
    const double a = -sc->W[(16 * rt + (lane & 15)) * 4 + (lane >> 4)];
#pragma unroll
    for (int t2 = 0; t2 < 2; ++t2) {
      const double bb = sc->S[(lane >> 4) * 32 + 16 * t2 + (lane & 15)];
      acc[rt][t2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc[rt][t2], 0, 0, 0);
    }
  }
  // (7) the panel's columns take the row-per-lane result
  if (holds) {
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[rt][CT][r] = sc->P[(16 * rt + 4 * r + (lane >> 4)) * 4 + jj0];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// In-place inversion of the n x n matrix held in acc (rows and columns n..31
// identity) by one whole wave.  On exit acc holds the inverse with columns in
// pivot order (as gj_invert), sc->pos[i] the position of row i and
// sc->perm[j] the row that was pivot at step j; det and err as gj_invert.
__device__ __forceinline__ void gj_invert_mfma32(pq_d4 (&acc)[2][2], int n, GjMfmaScratch* sc,
                                                 double& det, int& err) {
  const int lane = __lane_id();
  const int row = lane & 31;
  const bool act = lane < 32;
  int mypos = row;
  det = 1.0;
  err = 0;
  for (int kb = 0; kb < 4 && 4 * kb < n; ++kb) gj32_panel<0>(acc, kb, sc, lane, mypos, det, err);
  for (int kb = 4; kb < 8 && 4 * kb < n; ++kb) gj32_panel<1>(acc, kb, sc, lane, mypos, det, err);
  if (act) {
    sc->pos[row] = mypos;
    sc->perm[mypos] = row;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

constexpr int gj_group(int nc) { return nc <= 4 ? 4 : (nc <= 8 ? 8 : (nc <= 16 ? 16 : 32)); }

// MM bounds the moving particles of a move: the sampler's one-particle
// instances (MM = 1) fold the moving-particle loops away and never permute.
template <int D, int NC, class R, int NT, int MM = PQ_MAX_MOVING>
struct Eng {
  static constexpr int G = gj_group(NC);
  // FreeParticleNodes::MatrixUpdate is compiled into the general instances
  // only (the host selects them for any chain with use_updates), so the
  // one-particle instances keep their registers for the full-recompute path
  static constexpr bool kUpd = MM > 1;
  __device__ __forceinline__ bool upd(const FnP& f) const { return kUpd && f.updates; }
  // chooser results shared by the team through LDS
  struct ChooseBuf {
    int idx[PQ_MAX_MOVING];   // moving particles
    int cyc[PQ_MAX_MOVING];   // moving particle m ends on idx[cyc[m]]'s last bead
    unsigned long long k;     // draw counter after the chooser
    double lnTP;              // log transition probability of the permutation
    int ok;
    int pad;
    double lnT[8];            // the mover's ln(T_old/T_new) of every level of the move
    double rej[PQ_MAX_MOVING * 3];   // delayed rejection: the rejected level L-1 beads,
    double transA2B, accA2B, piBoA;  // their forward probability, acceptance and pi ratio,
    unsigned long long kC;           // and the first draw of the second proposal
  };
  const DevCtx& x;
  const ChainP& c;
  int chain;
  int tid;
  int lane;
  uint32_t moved_done = 0;   // per node action (8 bits each): moved slices evaluated this move
  int wave;
  int ins;          // row stride of the inverse cache (largest node species)
  double* red;      // [2][8] block-reduction scratch, halves used in turn
  int redsel = 0;   // the half the next bsum writes
  double* sec1;     // [S][N][D]
  double* sec2;     // [S][N][D]
  double* mov;      // [S][nmov][D]
  int nmov_ = 1;    // moving particles of the current move, nmov()
  double fac = 1.0; // the movers' sigma factor (MultiLevelSampler::factor), MM > 1 only
  const int* midx = nullptr;   // [nmov] their indices
  int mi0 = 0;      // the moving particle when MM = 1 (team-uniform register copy)
  double* fst;      // fixed-node state block
  double* inv;      // [S][ins][ins] latest inverse per section slice
  double* egr;      // [S][ins][ins][D] gradient factors of the latest FreeParticleNodes elements
  double* scale;    // [nfn]
  int* notmine;     // [nfn]
  double* wdet;     // [NT/G] scratch: determinants of a pass
  int* werr;        // [NT/G] scratch: error flags of a pass
  double* wscale;   // [S] scale used per slice in initialize
  double* gjrow;    // [NT/G][2 NC] pivot rows of the Gauss-Jordan groups
  // sampler only: chooser scratch and WalkingChooser rows (t, cump) in LDS
  ChooseBuf* cbuf = nullptr;
  double* wrows = nullptr;
  // sampler only: pair tables staged in LDS (tabofs[pair] offsets into tab)
  const R* tab = nullptr;
  const int* tabofs = nullptr;
  // sampler only: node records of every slice pair, reused across sections
  double* fcache = nullptr;
  // the records hold global particle labels: a section that wraps past slice
  // M under a non-identity path permutation sees other labels, so the
  // sampler switches them off for it
  bool rec_on = true;
  int sec_t0 = 0;          // first slice of section 1
  bool top_known = false;  // the section's last slice is never moved
  int* lasterr = nullptr;  // [nfn] error flag of the last slice at initialize

  __device__ Eng(const DevCtx& x_, const ChainP& c_, int chain_)
      : x(x_), c(c_), chain(chain_), tid(threadIdx.x), lane(threadIdx.x & 63), wave(threadIdx.x >> 6),
        ins(c_.ns_max > 0 ? c_.ns_max : 1), red(nullptr) {}

  // sum over the NT threads of the team; every thread gets the result.
  // Consecutive sums write alternate halves of red: a wave writes a half
  // again only after the next sum's barrier, which every wave reaches after
  // reading that half, so one barrier per sum suffices.
  __device__ __forceinline__ double bsum(double v) {
    v = wave_sum(v);
    if constexpr (NT == PQ_WAVE) return v;
    double* r = red + redsel;
    redsel ^= 8;
    if (lane == 0) r[wave] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int w = 0; w < NT / PQ_WAVE; ++w) s += r[w];
    return s;
  }

  __device__ __forceinline__ double U(uint64_t k) const { return uniform(x, chain, k); }
  __device__ __forceinline__ int nmov() const { return MM == 1 ? 1 : nmov_; }
  __device__ __forceinline__ int mv(int m) const { return MM == 1 ? mi0 : midx[m]; }

  __device__ __forceinline__ double* sb(double* s, int t, int i) const {
    return s + ((int64_t)t * c.N + i) * D;
  }

  // ------------------------------------------------------------ choosers
  // choosePermutation then chooseParticles of Sample smp (MultiLevelSampler.cc:
  // 68-80, DoubleMLSampler.cc:76-81).  Sets nmov, midx (in cbuf) and the
  // permutation cbuf->cyc; lnTP is the permutation chooser's log transition
  // probability; false when the permutation chooser turns the move down.
  //   SPECIES     PermutationChooser (identity, no draw) + SpeciesParticleChooser
  //               (SpeciesParticleChooser.cc:19-31)
  //   WALKING     WalkingChooser (WalkingChooser.cc:39-142), walk_choose
  //   ALL         PermutationChooser + SimpleParticleChooser over every
  //               particle, redrawing (int)(N u) == N (SimpleParticleChooser.cc:15-27)
  //   ALL_RANDOM  RandomPermutationChooser of nmov elements
  //               (RandomPermutationChooser.cc:12-24) + SimpleParticleChooser
  // The draws are sequential and cheap: one thread makes them and the team
  // reads the result back from LDS.
  __device__ __forceinline__ bool choose(int smp, uint64_t& k, double& lnTP) {
    const int kind = c.sample_ch[smp];
    const int sp = c.sample_sp[smp];
    const int nm = MM == 1 ? 1 : c.sample_nm[smp];
    ChooseBuf& cb = *cbuf;
    nmov_ = nm;
    midx = cb.idx;
    if (kind == PQ_CHOOSER_WALKING) return walk_choose(sp, nm, k, lnTP);
    if constexpr (MM == 1) {
      // one particle, no permutation: every thread makes the same draws
      lnTP = 0;
      if (kind == PQ_CHOOSER_ALL || kind == PQ_CHOOSER_ALL_RANDOM) {
        if (kind == PQ_CHOOSER_ALL_RANDOM) ++k;
        for (;;) {
          const int j = (int)(c.N * U(k++));
          if (j != c.N) { mi0 = j; return true; }
        }
      }
      int j = (int)(c.count[sp] * U(k++) + c.ifirst[sp]);
      if (j == c.count[sp] + c.ifirst[sp]) j = c.count[sp] + c.ifirst[sp] - 1;
      mi0 = j;
      return true;
    }
    if (tid == 0) {
      uint64_t kk = k;
      for (int m = 0; m < nm; ++m) cb.cyc[m] = m;
      if (kind == PQ_CHOOSER_ALL_RANDOM) {
        for (int ifrom = 0; ifrom < nm;) {
          int ito = (int)(nm * U(kk++));
          if (ito == nm) ito = nm - 1;
          for (int jf = 0; jf <= ifrom; ++jf) {
            if (jf == ifrom) { cb.cyc[ifrom++] = ito; break; }
            if (cb.cyc[jf] == ito) break;
          }
        }
      }
      const bool all = kind == PQ_CHOOSER_ALL || kind == PQ_CHOOSER_ALL_RANDOM;
      const int pool = all ? c.N : c.count[sp], first = all ? 0 : c.ifirst[sp];
      for (int im = 0; im < nm;) {
        int j;
        if (all) {
          j = (int)(pool * U(kk++));
          if (j == pool) continue;
        } else {
          j = (int)(pool * U(kk++) + first);
          if (j == pool + first) j = pool + first - 1;
        }
        for (int jm = 0; jm <= im; ++jm) {
          if (jm == im) { cb.idx[im++] = j; break; }
          if (cb.idx[jm] == j) break;
        }
      }
      cb.k = kk;
    }
    __syncthreads();
    k = cb.k;
    lnTP = 0;
    mi0 = cb.idx[0];
    // the next Sample's chooser rewrites cbuf
    __syncthreads();
    return true;
  }

  // free-particle propagator of the WalkingChooser between particle i's first
  // and particle j's last section bead, prod_d PG_{m/(2 tau 2^L)}(pbc(delta)_d)
  // (WalkingChooser.cc:89-103)
  __device__ __forceinline__ double walk_t(int sp, int i, int j) const {
    double delta[D];
#pragma unroll
    for (int d = 0; d < D; ++d) delta[d] = sb(sec1, 0, i)[d] - sb(sec1, c.S - 1, j)[d];
    pbc<D>(c, delta);
    const double alpha = c.mass[sp] / (2 * c.tau * (1 << c.L));
    double t = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double len = c.a[d];
      t *= (alpha * len * len >= 20.0) ? pg_images<double>(alpha, len, delta[d], nullptr)
                                       : pg_eval(pg_make(alpha, len), delta[d]);
    }
    return t;
  }

  // WalkingChooser::choosePermutation (WalkingChooser.cc:39-87) with init's
  // tables (:89-118) built one row per walk step: row p holds t(idx[p], j) and
  // its running sum cump, the team fills the row and one thread walks.  The
  // next particle is iSearch(prev, h u) (:120-142) with h = cump(prev, npart-1);
  // a repeated particle ends the move; the cycle m -> m+1 (mod nsize) is
  // accepted with probability revTranProb/tranProb and lnTP = log prod_m
  // t(i_m,i_m)/t(i_m,i_m+1).
  __device__ __forceinline__ bool walk_choose(int sp, int nsize, uint64_t& k, double& lnTP) {
    ChooseBuf& cb = *cbuf;
    const int npart = c.count[sp], first = c.ifirst[sp];
    double tran = 0, rev = 0;   // thread 0's running sums
    if (tid == 0) {
      uint64_t kk = k;
      int i0 = (int)(npart * U(kk++));
      if (i0 == npart) i0 = npart - 1;
      cb.idx[0] = i0;
      for (int m = 0; m < nsize; ++m) cb.cyc[m] = (m + 1) % nsize;
      cb.k = kk;
      cb.ok = 1;
    }
    __syncthreads();
    for (int ip = 0; ip < nsize; ++ip) {
      const int r = cb.idx[ip];
      double* trow = wrows + (int64_t)ip * 2 * npart;
      for (int j = tid; j < npart; j += NT) trow[j] = walk_t(sp, first + r, first + j);
      __syncthreads();
      if (tid == 0) {
        double* cump = trow + npart;
        cump[0] = trow[0];
        for (int j = 1; j < npart; ++j) cump[j] = cump[j - 1] + trow[j];
        if (ip + 1 < nsize) {
          uint64_t kk = cb.k;
          const double h = cump[npart - 1];
          const double xx = h * U(kk++);
          int nx;
          if (xx < cump[0]) nx = 0;
          else if (cump[npart - 2] < xx) nx = npart - 1;
          else {
            int ilo = 0, ihi = npart - 2, mid = 0;
            nx = -1;
            for (int q = 0; q < npart - 1; q++) {
              if (ihi == ilo + 1) { nx = ihi; break; }
              mid = (ilo + ihi) / 2;
              if (cump[mid] > xx) ihi = mid;
              else ilo = mid;
            }
            if (nx < 0) nx = mid;
          }
          cb.k = kk;
          bool dup = false;
          for (int jp = 0; jp <= ip; ++jp) dup |= cb.idx[jp] == nx;
          if (dup) {
            cb.ok = 0;
          } else {
            cb.idx[ip + 1] = nx;
            rev += h / trow[r];
            tran += h / trow[nx];
          }
        }
      }
      __syncthreads();
      if (!cb.ok) break;
    }
    if (cb.ok && tid == 0) {
      const int last = cb.idx[nsize - 1];
      const double* tl = wrows + (int64_t)(nsize - 1) * 2 * npart;
      const double h = tl[2 * npart - 1];
      rev += h / tl[last];
      tran += h / tl[cb.idx[0]];
      const double accept = rev / tran;
      double prob = 1;
      for (int m = 0; m < nsize; ++m) {
        const double* tm = wrows + (int64_t)m * 2 * npart;
        prob *= tm[cb.idx[m]] / tm[cb.idx[(m + 1) % nsize]];
      }
      cb.lnTP = log(prob);
      cb.ok = U(cb.k) < accept;
      cb.k += 1;
      for (int m = 0; m < nsize; ++m) cb.idx[m] += first;
    }
    __syncthreads();
    k = cb.k;
    lnTP = cb.lnTP;
    mi0 = cb.idx[0];
    const bool ok = cb.ok;
    __syncthreads();
    return ok;
  }

  // ------------------------------------------------------------ mover
  // FreeMover::makeMove (FreeMover.cc:59-124) / FreeMoverPBC::makeMove
  // (FreeMoverPBC.cc:59-162); lane k places the bead of moving particle
  // k % nmov at slice s(2(k / nmov) + 1).  Per slice the reference draws
  // nmov D Gaussians by Box-Muller pairs over the flattened array
  // (RandomNumGenerator.h:75-86), then, for FreeMoverPBC, D image uniforms per
  // moving particle; each lane computes its own from the draw offsets.
  // Returns ln(T_old/T_new); with fwd, FreeMover's forward probability
  // (getForwardProb: the new beads' Gaussian exponents and the log of their
  // inverse periodic-Gaussian weights) goes there.
  // one new bead of the mover: moving particle m (particle i of species sp)
  // at slice t = s(2kk+1) of level l, from its Gaussian g and the level's
  // draw offset kb; adds its ln(T_old/T_new) terms to lnT and, with fwd, its
  // forward-probability terms to fw
  __device__ __forceinline__ void place_bead(int level, int m, int i, int sp, int t, uint64_t kb, int gd,
                                             const double* g, double* fwd, double& lnT, double& fw) {
    const int s = 1 << level;
    const bool pbcmv = c.mover == PQ_MOVER_FREE_PBC;
    const double lam = c.lambda[sp];
    const double sigma = (MM > 1 ? fac : 1.0) * sqrt(lam * c.tau * s);
    const double inv2Sigma2 = 0.5 / (sigma * sigma);
    const double* mp = mov + ((t - s) * nmov() + m) * D;
    const double* mn = mov + ((t + s) * nmov() + m) * D;
    const double* op = sb(sec1, t - s, i);
    const double* on = sb(sec1, t + s, i);
    const double* oc = sb(sec1, t, i);
    double* mc = mov + (t * nmov() + m) * D;
    if (!pbcmv) {
      double mid[D], delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) mid[d] = mn[d] - mp[d];
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += mp[d]; }
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = g[d] * sigma;
      pbc<D>(c, delta);
      double nb[D];
#pragma unroll
      for (int d = 0; d < D; ++d) nb[d] = mid[d] + delta[d];
      pbc<D>(c, nb);
#pragma unroll
      for (int d = 0; d < D; ++d) mc[d] = nb[d];
      double temp = 1.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c.mpg_on[level][sp][d]) temp *= pg_eval(pg_make(c.mass[sp] / (c.tau * s), c.a[d]), delta[d]);
        else {
          const double tmp = delta[d] * delta[d] * inv2Sigma2;
          if (fwd) fw += tmp;
          lnT += tmp;
        }
      }
      temp = 1.0 / temp;
      if (fwd && temp != 1.0) fw += log(temp);
#pragma unroll
      for (int d = 0; d < D; ++d) mid[d] = on[d] - op[d];
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += op[d]; }
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = oc[d] - mid[d];
      pbc<D>(c, delta);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c.mpg_on[level][sp][d]) temp *= pg_eval(pg_make(c.mass[sp] / (c.tau * s), c.a[d]), delta[d]);
        else lnT -= delta[d] * delta[d] * inv2Sigma2;
      }
      if (temp != 1.0) lnT += log(temp);   // log 1 = 0 exactly
    } else {
      double mid[D], mid2[D];
#pragma unroll
      for (int d = 0; d < D; ++d) mid[d] = mn[d] - mp[d];
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) mid[d] *= 0.5;
#pragma unroll
      for (int d = 0; d < D; ++d) mid2[d] = mid[d] + 0.5 * c.a[d];
      pbc<D>(c, mid2);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double w1, w2;
        if (c.mpg_on[level][sp][d]) {
          PG p = pg_make(c.mass[sp] / (c.tau * s), c.a[d]);
          w1 = pg_eval(p, mid[d]);
          w2 = pg_eval(p, mid2[d]);
        } else {
          w1 = exp(-inv2Sigma2 * mid[d] * mid[d]);
          w2 = exp(-inv2Sigma2 * mid2[d] * mid2[d]);
        }
        if (U(kb + gd + m * D + d) > w1 / (w1 + w2)) mid[d] = mid2[d];
      }
      double nb[D];
#pragma unroll
      for (int d = 0; d < D; ++d) { mid[d] += mp[d]; nb[d] = mid[d] + g[d] * sigma; }
      pbc<D>(c, nb);
#pragma unroll
      for (int d = 0; d < D; ++d) mc[d] = nb[d];
      double dpn[D], dnn[D], dnew[D], dpo[D], dno[D], dold[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dpn[d] = nb[d] - mp[d];
        dnn[d] = nb[d] - mn[d];
        dnew[d] = mn[d] - mp[d];
        dpo[d] = oc[d] - op[d];
        dno[d] = oc[d] - on[d];
        dold[d] = on[d] - op[d];
      }
      pbc<D>(c, dpn); pbc<D>(c, dnn); pbc<D>(c, dnew);
      pbc<D>(c, dpo); pbc<D>(c, dno); pbc<D>(c, dold);
      double temp = 1.;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c.mpg_on[level + 1][sp][d]) {
          PG p = pg_make(c.mass[sp] / (c.tau * (s << 1)), c.a[d]);
          temp *= pg_eval(p, dpo[d]); temp *= pg_eval(p, dno[d]);
          temp /= pg_eval(p, dpn[d]); temp /= pg_eval(p, dnn[d]);
        } else {
          lnT -= dpo[d] * dpo[d] * inv2Sigma2 * 0.5;
          lnT -= dno[d] * dno[d] * inv2Sigma2 * 0.5;
          lnT += dpn[d] * dpn[d] * inv2Sigma2 * 0.5;
          lnT += dnn[d] * dnn[d] * inv2Sigma2 * 0.5;
        }
        if (level + 2 < PQ_PG_LEVELS && c.mpg_on[level + 2][sp][d]) {
          PG p = pg_make(c.mass[sp] / (c.tau * (s << 2)), c.a[d]);
          temp /= pg_eval(p, dold[d]); temp *= pg_eval(p, dnew[d]);
        } else {
          lnT += dold[d] * dold[d] * inv2Sigma2 * 0.25;
          lnT -= dnew[d] * dnew[d] * inv2Sigma2 * 0.25;
        }
      }
      if (temp != 1.0) lnT += log(temp);   // log 1 = 0 exactly
    }
  }

  // Every level of a one-particle move, proposed by the calling wave: the
  // Box-Muller Gaussians do not depend on the beads, so the lanes first draw
  // those of every level's beads at once (each bead's Gaussian is parked in
  // its own slot of mov), then the levels place their beads coarse to fine,
  // a level reading only the coarser levels' beads.  Level l's draws start
  // at k + (L - l) + pd (2^(L-1-l) - 1): the draws of the levels above and
  // one accept draw per level, as the level-by-level mover consumes them.
  // lnT[l] receives each level's ln(T_old/T_new) (0 at the top level).
  __device__ __forceinline__ void mover_all(uint64_t k, double* lnT_out) {
    const int L = c.L, nb = (1 << L) - 1;
    const bool pbcmv = c.mover == PQ_MOVER_FREE_PBC;
    const int gd = 2 * ((D + 1) / 2);
    const int pd = gd + (pbcmv ? D : 0);
    const int i = mv(0);
    const int sp = species_of(c, i);
    // one Box-Muller pair per lane: item (bead q, pair p)
    constexpr int NP = (D + 1) / 2;
    for (int it = lane; it < nb * NP; it += PQ_WAVE) {
      const int q = NP == 1 ? it : it >> 1, p = NP == 1 ? 0 : it & 1;
      const int lg = 31 - __clz(q + 1);
      const int l = L - 1 - lg, kk = q + 1 - (1 << lg);
      const int t = (1 << l) * (2 * kk + 1);
      const uint64_t kb = k + (L - l) + (uint64_t)pd * ((1u << lg) - 1) + (uint64_t)kk * pd;
      double t1 = 1 - 0.9999999999 * U(kb + 2 * p);
      double t2 = U(kb + 2 * p + 1);
      double rr = sqrt(-2.0 * log(t1));
      double sn, cs;
      sincos(6.283185306 * t2, &sn, &cs);
      mov[t * D + 2 * p] = rr * cs;
      if (2 * p + 1 < D) mov[t * D + 2 * p + 1] = rr * sn;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane == 0) lnT_out[L] = 0.0;
    for (int l = L - 1; l >= 0; --l) {
      const int lg = L - 1 - l;
      double lnT = 0, fw = 0;
      for (int kk = lane; kk < (1 << lg); kk += PQ_WAVE) {
        const int t = (1 << l) * (2 * kk + 1);
        const uint64_t kb = k + (L - l) + (uint64_t)pd * ((1u << lg) - 1) + (uint64_t)kk * pd;
        double g[D];
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = mov[t * D + d];
        place_bead(l, 0, i, sp, t, kb, gd, g, nullptr, lnT, fw);
      }
      lnT = wave_sum(lnT);
      if (lane == 0) lnT_out[l] = lnT;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }

  // uniforms the mover draws at a level (FreeMover / FreeMoverPBC)
  __device__ __forceinline__ int mover_draws(int level) const {
    const int nmoved = (level < c.L) ? (1 << (c.L - 1 - level)) : 0;
    const int ng = nmov() * D;
    return nmoved * (2 * ((ng + 1) / 2) + (c.mover == PQ_MOVER_FREE_PBC ? ng : 0));
  }

  __device__ __forceinline__ double mover(int level, uint64_t k0, int* ndraws, double* fwd = nullptr) {
    const int s = 1 << level;
    const int nmoved = (level < c.L) ? (1 << (c.L - 1 - level)) : 0;
    const bool pbcmv = c.mover == PQ_MOVER_FREE_PBC;
    const int ng = nmov() * D;
    const int gd = 2 * ((ng + 1) / 2);
    const int pd = gd + (pbcmv ? ng : 0);
    *ndraws = nmoved * pd;
    double lnT = 0, fw = 0;
    for (int it = tid; it < nmoved * nmov(); it += NT) {
      int kk = it, m = 0;
      if (nmov() > 1) { kk = it / nmov(); m = it - kk * nmov(); }
      const int i = mv(m);
      const int sp = species_of(c, i);
      const int t = s * (2 * kk + 1);
      const uint64_t kb = k0 + (uint64_t)kk * pd;
      double g[D];
      if (nmov() == 1) {
        double gg[4];
#pragma unroll
        for (int p = 0; p < (D + 1) / 2; ++p) {
          double t1 = 1 - 0.9999999999 * U(kb + 2 * p);
          double t2 = U(kb + 2 * p + 1);
          double rr = sqrt(-2.0 * log(t1));
          double sn, cs;
          sincos(6.283185306 * t2, &sn, &cs);
          gg[2 * p] = rr * cs;
          gg[2 * p + 1] = rr * sn;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = gg[d];
      } else {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const int e = m * D + d, p = e >> 1;
          double t1 = 1 - 0.9999999999 * U(kb + 2 * p);
          double t2 = U(kb + 2 * p + 1);
          double rr = sqrt(-2.0 * log(t1));
          double sn, cs;
          sincos(6.283185306 * t2, &sn, &cs);
          g[d] = (e & 1) ? rr * sn : rr * cs;
        }
      }
      place_bead(level, m, i, sp, t, kb, gd, g, fwd, lnT, fw);
    }
    const double r = bsum(lnT);
    if (fwd) {
      const double f = bsum(fw);
      if (tid == 0) *fwd = f;
    }
    return r;
  }

  // FreeMover::makeDelayedMove (FreeMover.cc:126-184) at level L-1, whose one
  // moved slice t = 2^(L-1) lies between the section's end slices: ln(T_old/
  // T_new) of proposing the rejected beads B (rej, [nmov][D]) from the second
  // proposal C (in mov) against C from B, both about the same midpoint, and
  // its forward probability (minus B's Gaussian exponents plus the log of
  // B's periodic-Gaussian weights) in fwd.
  __device__ __forceinline__ double delayed_mover(int level, const double* rej, double* fwd) {
    const int s = 1 << level;
    const int t = s;
    double lnT = 0, fw = 0;
    for (int m = tid; m < nmov(); m += NT) {
      const int i = mv(m);
      const int sp = species_of(c, i);
      const double sigma = fac * sqrt(c.lambda[sp] * c.tau * s);
      const double inv2Sigma2 = 0.5 / (sigma * sigma);
      const double* mp = mov + ((t - s) * nmov() + m) * D;
      const double* mn = mov + ((t + s) * nmov() + m) * D;
      const double* mc = mov + (t * nmov() + m) * D;
      const double* rb = rej + m * D;
      double mid[D], delta[D];
#pragma unroll
      for (int d = 0; d < D; ++d) mid[d] = mn[d] - mp[d];
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += mp[d]; }
      pbc<D>(c, mid);
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = rb[d] - mid[d];
      pbc<D>(c, delta);
      double temp = 1.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c.mpg_on[level][sp][d]) temp *= pg_eval(pg_make(c.mass[sp] / (c.tau * s), c.a[d]), delta[d]);
        else {
          const double tmp = delta[d] * delta[d] * inv2Sigma2;
          fw -= tmp;
          lnT += tmp;
        }
      }
      fw += log(temp);
      temp = 1.0 / temp;
      // the end slices are the same in B and C, and so is the midpoint
#pragma unroll
      for (int d = 0; d < D; ++d) delta[d] = mc[d] - mid[d];
      pbc<D>(c, delta);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c.mpg_on[level][sp][d]) temp *= pg_eval(pg_make(c.mass[sp] / (c.tau * s), c.a[d]), delta[d]);
        else lnT -= delta[d] * delta[d] * inv2Sigma2;
      }
      lnT += log(temp);
    }
    const double r = bsum(lnT);
    *fwd = bsum(fw);
    return r;
  }

  // ---------------------------------------------------- action layer
  // Spring (SpringAction.cc:56-98), SHO (SHOAction.cc:19-65) and Coulomb pair
  // (PairAction.cc:219-360, CoulombAction.cc:100-105) action differences of
  // the moving particles at a level.  Lanes split the (link, moving particle)
  // and (moving particle, partner, link) items; a pair of two moving particles
  // counts once, from the higher index, with both new paths.
  // the partner species of a pair action for a moving particle of species
  // sp, or -1 when the pair action does not involve sp
  __device__ __forceinline__ static int pair_partner(const PairP& P, int sp) {
    return sp == P.s1 ? P.s2 : (sp == P.s2 ? P.s1 : -1);
  }
  __device__ __forceinline__ double action_diff(int level) {
    const int s = 1 << level;
    const int nl = (c.S - 1) >> level;
    const int nA = nl * nmov();   // spring and SHO links, lanes 0..nA-1
    Acc<R> acc;
    for (int it = tid; it < nA; it += NT) {
      int k = it, m = 0;
      if (nmov() > 1) { k = it / nmov(); m = it - k * nmov(); }
      const int i = mv(m);
      const int sp = species_of(c, i);
      const int t = s * (k + 1);
      const double* m1 = mov + (t * nmov() + m) * D;
      const double* m0 = mov + ((t - s) * nmov() + m) * D;
      const double* o1 = sb(sec1, t, i);
      const double* o0 = sb(sec1, t - s, i);
      if (c.spring && !c.stat[sp]) {
        const R inv2Sigma2 = (R)c.spring_c[level][sp];
        double dn[D], dd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) { dn[d] = m1[d] - m0[d]; dd[d] = o1[d] - o0[d]; }
        pbc<D>(c, dn);
        pbc<D>(c, dd);
        double temp = 1.0;
        R part = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if (c.spg_on[level][sp][d]) {
            PG p = pg_make(c.mass[sp] / (2 * c.tau * s), c.a[d]);
            temp *= pg_eval(p, dd[d]) / pg_eval(p, dn[d]);
          } else {
            part += ((R)dn[d] * (R)dn[d] - (R)dd[d] * (R)dd[d]) * inv2Sigma2;
          }
        }
        acc.add(part);
        if (temp != 1.0) acc.add((R)log(temp));
      }
      for (int a = 0; a < c.nsho; ++a) {
        const int ssp = c.sho_sp[a];
        if (i < c.ifirst[ssp] || i >= c.ifirst[ssp] + c.count[ssp]) continue;
        const R m = (R)c.mass[ssp], omega = (R)c.sho_omega[a];
        const R coshwt = (R)c.sho_cosh[a][level], div1 = (R)c.sho_div1[a][level];
        const R div2 = (R)c.div2[level];
        const int d0 = D - (c.sho_ndim[a] ? c.sho_ndim[a] : D);
        R v[2];
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          const double* p1 = pass ? o1 : m1;
          const double* p0 = pass ? o0 : m0;
          R r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            if (d < d0) continue;
            R r1 = (R)(p1[d] - c.sho_center[a][d]), r0 = (R)(p0[d] - c.sho_center[a][d]);
            R dl = r1 - r0;
            r0r0 += r0 * r0; r0r1 += r0 * r1; r1r1 += r1 * r1; delta2 += dl * dl;
          }
          v[pass] = m * omega * ((r1r1 + r0r0) * coshwt - (R)2.0 * r0r1) * div1 - m * delta2 * div2;
        }
        acc.add(v[0]);
        acc.add(-v[1]);
      }
    }
    // pair actions: one index space over the (pair action, moving particle)
    // blocks of partner x link items, spread over the whole team after the
    // lanes of the links above
    int npi = 0;
    for (int pp = 0; pp < c.npair; ++pp)
      for (int m = 0; m < nmov(); ++m) {
        const int js = pair_partner(c.pair[pp], species_of(c, mv(m)));
        if (js >= 0) npi += c.count[js] * nl;
      }
    for (int it = tid; it < nA + npi; it += NT) {
      if (it < nA) continue;
      // the block holding item it - nA
      int r = it - nA, pp = 0, m = 0, js = -1;
      for (int q = 0; q < c.npair * nmov(); ++q) {
        pp = q / nmov();
        m = q - pp * nmov();
        js = pair_partner(c.pair[pp], species_of(c, mv(m)));
        if (js < 0) continue;
        const int cnt = c.count[js] * nl;
        if (r < cnt) break;
        r -= cnt;
      }
      const int i = mv(m);
      const TableP& T = c.pair[pp].tp;
      const R* u = tab ? tab + tabofs[pp] : table_ptr<R>(x, T);
      const bool offdiag = level < 3 && T.K1 > 1;
      {
        const int j = c.ifirst[js] + r / nl;
        if (j == i) continue;
        int jm = -1;
        if constexpr (MM > 1)
          for (int q = 0; q < nmov(); ++q)
            if (midx[q] == j) jm = q;
        if (jm >= 0 && i <= j) continue;
        const int t = s * (r % nl + 1);
        const double* pj1 = sb(sec1, t, j);
        const double* pj0 = sb(sec1, t - s, j);
        R vals[2];
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          const double* pi1 = pass ? sb(sec1, t, i) : mov + (t * nmov() + m) * D;
          const double* pi0 = pass ? sb(sec1, t - s, i) : mov + ((t - s) * nmov() + m) * D;
          if (pass == 0 && jm >= 0) {
            pj1 = mov + (t * nmov() + jm) * D;
            pj0 = mov + ((t - s) * nmov() + jm) * D;
          } else if (pass == 1) {
            pj1 = sb(sec1, t, j);
            pj0 = sb(sec1, t - s, j);
          }
          double d1[D], d0[D];
#pragma unroll
          for (int d = 0; d < D; ++d) { d1[d] = pi1[d] - pj1[d]; d0[d] = pi0[d] - pj0[d]; }
          pbc<D>(c, d1);
          pbc<D>(c, d0);
          R r1 = 0, r0 = 0, sv = 0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            R a1 = (R)d1[d], a0 = (R)d0[d], e = a1 - a0;
            r1 += a1 * a1; r0 += a0 * a0; sv += e * e;
          }
          r1 = r_sqrt(r1);
          r0 = r_sqrt(r0);
          R q = (R)0.5 * (r1 + r0);
          R s2 = offdiag ? sv / (q * q) : (R)0;
          vals[pass] = pair_u<R>(T, u, q, s2, offdiag) * (R)s;
        }
        acc.add(vals[0]);
        acc.add(-vals[1]);
      }
    }
    const double links = bsum(acc.value());
    if (!(c.ew.nk && level == 0)) return links;
    return links + ewald_diff();
  }

  // CoulombAction's long-range Ewald action at level 0 (CoulombAction.cc:
  // 107-119): tau/eps sum over slices 1..2^L-1 of the k-space sum's change
  // when the moving particles take their new beads.  Lanes over (slice,
  // k-vector) items; per item rho_k of the section slice and the moving
  // particles' change d rho_k give w_k (|rho + d rho|^2 - |rho|^2) =
  // w_k (2 Re(rho* d rho) + |d rho|^2), in FP64 in both precision modes.
  __device__ __forceinline__ double ewald_diff() {
    const double* kp = x.ewk + c.ew.off;
    const int nk = c.ew.nk, items = (c.S - 2) * nk;
    double e = 0;
    for (int it = tid; it < items; it += NT) {
      const int t = 1 + it / nk;
      const double* kv = kp + 4 * (it - (t - 1) * nk);
      double rr = 0, ri = 0, dr = 0, di = 0;
      for (int j = 0; j < c.N; ++j) {
        double co, si;
        ewald_phase<D>(kv, sb(sec1, t, j), co, si);
        const double q = c.charge[species_of(c, j)];
        rr += q * co;
        ri += q * si;
      }
      for (int m = 0; m < nmov(); ++m) {
        const int i = mv(m);
        const double q = c.charge[species_of(c, i)];
        double cn, sn, co, so;
        ewald_phase<D>(kv, mov + (t * nmov() + m) * D, cn, sn);
        ewald_phase<D>(kv, sb(sec1, t, i), co, so);
        dr += q * (cn - co);
        di += q * (sn - so);
      }
      e += kv[3] * (2 * (rr * dr + ri * di) + dr * dr + di * di);
    }
    return bsum(e) * c.tau / c.ceps;
  }

  // -------------------------------------------------------- node model
  __device__ __forceinline__ double* fn_dm(int a) const { return fst + fn_base(a); }
  __device__ __forceinline__ double* fn_newdm(int a) const { return fst + fn_base(a) + c.S; }
  __device__ __forceinline__ double* fn_dist(int a) const { return fst + fn_base(a) + 2 * c.S; }
  __device__ __forceinline__ double* fn_newdist(int a) const {
    return fst + fn_base(a) + 2 * c.S + 2 * c.S * c.fn[a].ns;
  }
  __device__ __forceinline__ int64_t fn_base(int a) const {
    int64_t b = 0;
    for (int k = 0; k < a; ++k) b += 2 * c.S + 4 * (int64_t)c.S * c.fn[k].ns;
    return b;
  }

  // position of particle p in section 1 at slice t, with the moving particles
  // replaced by their moving beads when use_mov
  __device__ __forceinline__ const double* r1pos(int t, int p, bool use_mov) const {
    if (use_mov)
      for (int m = 0; m < nmov(); ++m)
        if (mv(m) == p) return mov + (t * nmov() + m) * D;
    return sb(sec1, t, p);
  }

  // Slice of position p in the coarse-first order of a move's changed slices:
  // the last slice first when a cycle moved it (the top level tests it), then
  // level L-1's slice, level L-2's two, ..., level 0's 2^(L-1), each level in
  // increasing slice order.  Level l holds slices 2^l (2i + 1).
  __device__ __forceinline__ int moved_slice(int p) const {
    if (!top_known) {
      if (p == 0) return c.S - 1;
      --p;
    }
    const int q = p + 1;
    const int lg = 31 - __clz(q);
    return (1 << (c.L - 1 - lg)) * (2 * (q - (1 << lg)) + 1);
  }
  // positions of that order a test at the given level needs evaluated
  __device__ __forceinline__ int moved_needed(int level) const {
    const int top = top_known ? 0 : 1;
    return level >= c.L ? top : top + (1 << (c.L - level)) - 1;
  }

  // NodeModel::evaluate (SHONodes.cc:51-128, FreeParticleNodes.cc:72-157) on
  // slices t0, t0+step, ... (nt of them), or with step 0 on the moved slices
  // at positions t0..t0+nt-1 of moved_slice's order, with the given scale;
  // inverses go to the cache, determinants and error flags to wdet/werr[0..nt).
  __device__ __forceinline__ int matrix_slice(int t0, int step, int m) const {
    return step ? t0 + m * step : moved_slice(t0 + m);
  }
  __device__ __forceinline__ void node_matrices(int fa, int t0, int step, int nt, bool use_mov, double scl) {
    const FnP& f = c.fn[fa];
    const int n = f.ns;
    // fill: every thread takes matrix elements; they are staged in the inverse
    // slot of their slice, FreeParticleNodes' gradient factors go to their own
    // cache (SHONodes' distances recompute theirs from the positions)
    const int per = n * n;
    const bool sho = f.model == PQ_NODES_SHO;
    for (int q = tid; q < nt * per; q += NT) {
      const int m = q / per, rem = q - m * per;
      const int ii = rem / n, jj = rem - ii * n;
      const int t = matrix_slice(t0, step, m);
      const double* ri = sb(sec2, t, f.ifirst + ii);
      const double* rj = r1pos(t, f.ifirst + jj, use_mov);
      double gf[D];
      const int64_t e = ((int64_t)t * ins + ii) * ins + jj;
      inv[e] = elem(f, rj, ri, scl, gf);
      if (!sho) {
#pragma unroll
        for (int d = 0; d < D; ++d) egr[e * D + d] = gf[d];
      }
    }
    __syncthreads();
    if constexpr (NC == 32) {
      // one wave per matrix, MFMA trailing updates
      constexpr int NW = NT / PQ_WAVE;
      GjMfmaScratch* sc = reinterpret_cast<GjMfmaScratch*>(gjrow + (int64_t)wave * PQ_GJ32_SCRATCH_DOUBLES);
      for (int base = 0; base < nt; base += NW) {
        const int m = base + wave;
        if (m < nt) {
          const int t = matrix_slice(t0, step, m);
          double* slot = inv + (int64_t)t * ins * ins;
          pq_d4 acc[2][2];
#pragma unroll
          for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
              for (int r4 = 0; r4 < 4; ++r4) {
                const int i = 16 * rt + 4 * r4 + (lane >> 4), j = 16 * ct + (lane & 15);
                acc[rt][ct][r4] = (i < n && j < n) ? slot[i * ins + j] : (i == j ? 1.0 : 0.0);
              }
          double det;
          int err;
          gj_invert_mfma32(acc, n, sc, det, err);
#pragma unroll
          for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
              for (int r4 = 0; r4 < 4; ++r4) {
                const int i = 16 * rt + 4 * r4 + (lane >> 4), j = 16 * ct + (lane & 15);
                if (i < n && j < n) slot[sc->pos[i] * ins + sc->perm[j]] = acc[rt][ct][r4];
              }
          if (lane == 0) { wdet[m] = det; werr[m] = err; }
          // the next matrix of this wave reuses the scratch
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      __syncthreads();
      return;
    }
    const int g = tid / G, r = tid % G;
    constexpr int NG = NT / G;
    constexpr int GPW = PQ_WAVE / G;   // groups per wave
    double* row = gjrow + (int64_t)g * 2 * NC;
    for (int base = 0; base < nt; base += NG) {
      // a wave whose groups all lie past the last matrix has nothing to do
      if (base + wave * GPW < nt) {
        const int m = base + g;
        const bool live = m < nt;
        const int t = live ? matrix_slice(t0, step, m) : t0;
        double a[NC];
        const double* src = inv + ((int64_t)t * ins + r) * ins;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
          a[cc] = (live && r < n && cc < n) ? src[cc] : (cc == r ? 1.0 : 0.0);
        double det;
        int err, pos;
        gj_invert<G, NC>(a, r, n, row, det, err, pos);
        // column j of the result is inverse column P(j): P(pos) = r
        int* perm = reinterpret_cast<int*>(row);
        if (live && r < n) perm[pos] = r;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (live) {
          if (r < n) {
            double* dst = inv + ((int64_t)t * ins + pos) * ins;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
              if (cc < n) dst[perm[cc]] = a[cc];
          }
          if (r == 0) { wdet[m] = det; werr[m] = err; }
        }
        // the next pass rewrites the row buffer that perm aliases
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    __syncthreads();
  }

  // density-matrix element rho(r2_i, r1_j) times the scale, and the factors the
  // nodal distance needs: SHONodes the unscaled element (gf[0]);
  // FreeParticleNodes the gradient (PG'(x_d)/(PG(x_d)+1e-300)) prod_k PG(x_k)
  // in the order of FreeParticleNodes.cc:196-207.
  // Always FP64: at low temperature the SHO matrix is close to rank one and
  // its determinant cancels, so element rounding decides the sign test.
  // SHONodes' unscaled element exp(-c[(r_i^2 + r_j^2) cosh - 2 r_i.r_j])
  __device__ __forceinline__ double sho_e(const FnP& f, const double* rj, const double* ri) const {
    double rj2 = 0, ri2 = 0, rirj = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double aj = rj[d], ai = ri[d];
      rj2 += aj * aj; ri2 += ai * ai; rirj += ai * aj;
    }
    return exp(-f.c * ((ri2 + rj2) * f.coshwt - 2.0 * rirj));
  }
  __device__ __forceinline__ double elem(const FnP& f, const double* rj, const double* ri, double scl,
                                         double* gf) const {
    if (f.model == PQ_NODES_SHO) {
      const double e = sho_e(f, rj, ri);
      gf[0] = e;
#pragma unroll
      for (int d = 1; d < D; ++d) gf[d] = 0.0;
      return scl * e;
    }
    double delta[D], value[D];
#pragma unroll
    for (int d = 0; d < D; ++d) delta[d] = rj[d] - ri[d];
    pbc<D>(c, delta);
    double ear2 = scl;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (f.pg_img[d]) {
        double g2;
        const double v2 = pg_images<double>(f.pg_alpha[d], f.pg_len[d], delta[d], &g2);
        value[d] = v2;
        gf[d] = ratio<double>(g2, v2);
      } else {
        double gr;
        value[d] = pg_eval(f.pg_nome[d], f.pg_pref[d], f.pg_k[d], f.pg_nmax[d], delta[d], &gr);
        gf[d] = gr / (value[d] + 1e-300);
      }
      ear2 *= value[d];
    }
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int d = 0; d < D; ++d) gf[d] *= value[k];
    return ear2;
  }

  // evaluateDistance (SHONodes.cc:130-170, FreeParticleNodes.cc:184-234) for
  // slices t0..t0+nt-1 into dist[t][2][ns]; slice t uses scale scl[t] when
  // scl is given, else scl0.
  __device__ __forceinline__ void node_distances(int fa, int t0, int nt, bool use_mov, const double* scl,
                                 double scl0, double* dist) {
    const FnP& f = c.fn[fa];
    const int n = f.ns;
    const int items = nt * 2 * n;
    for (int it = tid; it < items; it += NT) {
      const int t = t0 + it / (2 * n);
      const int pass = (it / n) & 1;
      const int aa = it % n;
      const double sc = scl ? scl[t] : scl0;
      const double* iv = inv + (int64_t)t * ins * ins;
      double lg[D];
#pragma unroll
      for (int d = 0; d < D; ++d) lg[d] = 0;
      // the inverse's row or column eight elements at a time, their loads
      // in flight together before the elements are used
      for (int b0 = 0; b0 < n; b0 += 8) {
        double wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int bb = b0 + u;
          wv[u] = bb < n ? iv[pass == 0 ? aa * ins + bb : bb * ins + aa] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int bb = b0 + u;
          if (bb >= n) break;
          const int jp = pass == 0 ? aa : bb, ip = pass == 0 ? bb : aa;
          const double w = wv[u];
          if (f.model == PQ_NODES_SHO) {
            const double* rj = r1pos(t, f.ifirst + jp, use_mov);
            const double* ri = sb(sec2, t, f.ifirst + ip);
            const double e = sho_e(f, rj, ri);
#pragma unroll
            for (int d = 0; d < D; ++d) {
              double gg = pass == 0 ? 2 * f.c * (ri[d] - f.coshwt * rj[d]) * e
                                    : 2 * f.c * (rj[d] - f.coshwt * ri[d]) * e;
              lg[d] += w * sc * gg;
            }
          } else {
            const double* gf = egr + (((int64_t)t * ins + ip) * ins + jp) * D;
            // the reversed separation negates the odd gradient factor exactly
#pragma unroll
            for (int d = 0; d < D; ++d) lg[d] += w * (pass == 0 ? gf[d] : -gf[d]) * sc;
          }
        }
      }
      double g2 = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) g2 += lg[d] * lg[d];
      dist[((int64_t)t * 2 + pass) * n + aa] = sqrt(2 * f.mass / ((g2 + 1e-15) * c.tau));
    }
    __syncthreads();
  }

  // ------------------------------------- FreeParticleNodes::MatrixUpdate
  // Node actions with useUpdates keep, through a section, the inverse of
  // every slice's matrix from initialize (FixedNodeAction::initialize
  // evaluates them all) in their own block, and a move evaluates the
  // determinant ratio of the moving particles' new columns instead of the
  // new matrix (FreeParticleNodes.cc:894-1058).
  __device__ __forceinline__ double* upd_block(int fa) const {
    return x.upd + c.upd_off + (int64_t)fa * upd_block_doubles(c);
  }
  __device__ __forceinline__ double* upd_inv(int fa) const { return upd_block(fa) + 8; }
  __device__ __forceinline__ double* upd_new(int fa) const {
    return upd_inv(fa) + (int64_t)c.Smax * ins * ins;
  }
  __device__ __forceinline__ double* upd_phi(int fa) const {
    return upd_new(fa) + (int64_t)c.Smax * ins * ins;
  }
  __device__ __forceinline__ double* upd_scr(int fa) const {
    return upd_phi(fa) + (int64_t)c.Smax * ins * PQ_MAX_MOVING;
  }
  // the moving particles of the node action's species, in the order of the
  // moving particles: how many, and the position of the jm-th among them
  __device__ __forceinline__ int upd_count(const FnP& f) const {
    int k = 0;
    for (int m = 0; m < nmov(); ++m) k += mv(m) >= f.ifirst && mv(m) < f.ifirst + f.ns;
    return k;
  }
  __device__ __forceinline__ int upd_moving(const FnP& f, int jm) const {
    for (int m = 0; m < nmov(); ++m)
      if (mv(m) >= f.ifirst && mv(m) < f.ifirst + f.ns) {
        if (jm == 0) return m;
        --jm;
      }
    return 0;
  }

  // MatrixUpdate::evaluateChange (FreeParticleNodes.cc:894-948) on slices t0,
  // t0+step, ... (nt of them): the moving particles' new columns
  // phi(k, jm) = scale prod_d PG(x_jm - x2_k), the block
  // b(im, jm) = sum_k phi(k, jm) inv(index1(im), k) and its determinant by
  // LU with partial pivoting (dgetf2 order), the ratio det(new)/det(old);
  // newDMValue = ratio * dmValue (FixedNodeAction.cc:61-63).
  __device__ __forceinline__ void upd_change(int fa, int t0, int step, int nt) {
    const FnP& f = c.fn[fa];
    const int n = f.ns, nm = upd_count(f);
    const double scl = scale[fa];
    double* phi = upd_phi(fa);
    const double* bi = upd_inv(fa);
    double* scr = upd_scr(fa);
    if (tid == 0) upd_block(fa)[0] = 0.0;   // isNewMatrixUpdated = false
    for (int q = tid; q < nt * n * nm; q += NT) {
      const int sl = q / (n * nm), rem = q - sl * n * nm;
      const int k = rem / nm, jm = rem - k * nm;
      const int t = t0 + sl * step;
      double gf[D];
      phi[((int64_t)t * ins + k) * PQ_MAX_MOVING + jm] =
          elem(f, mov + (t * nmov() + upd_moving(f, jm)) * D, sb(sec2, t, f.ifirst + k), scl, gf);
    }
    __syncthreads();
    for (int q = tid; q < nt * nm * nm; q += NT) {
      const int sl = q / (nm * nm), rem = q - sl * nm * nm;
      const int im = rem / nm, jm = rem - im * nm;
      const int t = t0 + sl * step;
      const double* row = bi + ((int64_t)t * ins + (mv(upd_moving(f, im)) - f.ifirst)) * ins;
      const double* ph = phi + (int64_t)t * ins * PQ_MAX_MOVING + jm;
      double v = 0;
      for (int k = 0; k < n; ++k) v += ph[k * PQ_MAX_MOVING] * row[k];
      scr[(int64_t)t * 64 + im * nm + jm] = v;
    }
    __syncthreads();
    double* ndm = fn_newdm(fa);
    const double* dm = fn_dm(fa);
    for (int sl = tid; sl < nt; sl += NT) {
      const int t = t0 + sl * step;
      double* a = scr + (int64_t)t * 64;
      double det = 1;
      for (int cc = 0; cc < nm; ++cc) {
        int p = cc;
        double amax = fabs(a[cc * nm + cc]);
        for (int r = cc + 1; r < nm; ++r)
          if (fabs(a[r * nm + cc]) > amax) { amax = fabs(a[r * nm + cc]); p = r; }
        if (a[p * nm + cc] != 0.0) {
          if (p != cc)
            for (int q = 0; q < nm; ++q) {
              const double tmp = a[cc * nm + q];
              a[cc * nm + q] = a[p * nm + q];
              a[p * nm + q] = tmp;
            }
          for (int r = cc + 1; r < nm; ++r) a[r * nm + cc] /= a[cc * nm + cc];
        }
        for (int r = cc + 1; r < nm; ++r)
          for (int q = cc + 1; q < nm; ++q) a[r * nm + q] -= a[r * nm + cc] * a[cc * nm + q];
        det *= a[cc * nm + cc];
        det *= (p == cc) ? 1 : -1;
      }
      ndm[t] = det * dm[t];
    }
    __syncthreads();
  }

  // The Sherman-Morrison row updates of evaluateNewInverse and acceptLastMove
  // (FreeParticleNodes.cc:950-974, 1026-1057) on the matrices m of slices
  // t0..t0+nt-1, one moving particle jm after the other: with
  // b_i = sum_k phi(k, jm) m(i, k), row jp (the particle) is scaled by 1/b_jp
  // and every other row takes m(i, k) -= m(jp, k) b_i.
  __device__ __forceinline__ void upd_rows(int fa, double* m, int t0, int nt) {
    const FnP& f = c.fn[fa];
    const int n = f.ns, nm = upd_count(f);
    const double* phi = upd_phi(fa);
    double* scr = upd_scr(fa);
    for (int jm = 0; jm < nm; ++jm) {
      const int jp = mv(upd_moving(f, jm)) - f.ifirst;
      for (int q = tid; q < nt * n; q += NT) {
        const int t = t0 + q / n, i = q % n;
        const double* row = m + ((int64_t)t * ins + i) * ins;
        const double* ph = phi + (int64_t)t * ins * PQ_MAX_MOVING + jm;
        double v = 0;
        for (int k = 0; k < n; ++k) v += ph[k * PQ_MAX_MOVING] * row[k];
        scr[(int64_t)t * 64 + i] = v;
      }
      __syncthreads();
      for (int q = tid; q < nt * n * n; q += NT) {
        const int t = t0 + q / (n * n), rem = q % (n * n);
        const int i = rem / n, k = rem % n;
        if (i == jp) continue;
        double* mt = m + (int64_t)t * ins * ins;
        const double sj = mt[jp * ins + k] * (1. / scr[(int64_t)t * 64 + jp]);
        mt[i * ins + k] -= sj * scr[(int64_t)t * 64 + i];
      }
      __syncthreads();
      for (int q = tid; q < nt * n; q += NT) {
        const int t = t0 + q / n, k = q % n;
        m[((int64_t)t * ins + jp) * ins + k] *= 1. / scr[(int64_t)t * 64 + jp];
      }
      __syncthreads();
    }
  }

  // MatrixUpdate::evaluateNewInverse (FreeParticleNodes.cc:950-974) of slices
  // 1..S-1 from the section's inverses and the move's new columns
  __device__ __forceinline__ void upd_new_inverse(int fa) {
    const int n = c.fn[fa].ns, nt = c.S - 1;
    const double* bi = upd_inv(fa);
    double* nw = upd_new(fa);
    for (int q = tid; q < nt * n * n; q += NT) {
      const int t = 1 + q / (n * n), rem = q % (n * n);
      const int64_t e = ((int64_t)t * ins + rem / n) * ins + rem % n;
      nw[e] = bi[e];
    }
    __syncthreads();
    upd_rows(fa, nw, 1, nt);
    if (tid == 0) upd_block(fa)[0] = 1.0;   // isNewMatrixUpdated
    __syncthreads();
  }

  // MatrixUpdate::evaluateNewDistance (FreeParticleNodes.cc:976-1024) for
  // slices 1..S-1 into dist[t][2][ns] from the updated inverses.  Unlike
  // evaluateDistance it leaves the scale out of the log gradient, and in its
  // first loop (d1) each gradient component takes its own dimension's value
  // only (grad[i] *= value[i]); the second loop (d2) multiplies the whole
  // vector by every value.
  __device__ __forceinline__ void upd_distances(int fa, double* dist) {
    const FnP& f = c.fn[fa];
    const int n = f.ns, nt = c.S - 1;
    const double* mat = upd_new(fa);
    for (int it = tid; it < nt * 2 * n; it += NT) {
      const int t = 1 + it / (2 * n);
      const int pass = (it / n) & 1;
      const int aa = it % n;
      const double* mt = mat + (int64_t)t * ins * ins;
      double lg[D];
#pragma unroll
      for (int d = 0; d < D; ++d) lg[d] = 0;
      for (int bb = 0; bb < n; ++bb) {
        const int jp = pass == 0 ? aa : bb, ip = pass == 0 ? bb : aa;
        const double* rj = r1pos(t, f.ifirst + jp, true);
        const double* ri = sb(sec2, t, f.ifirst + ip);
        double delta[D], value[D], grad[D];
#pragma unroll
        for (int d = 0; d < D; ++d) delta[d] = pass == 0 ? rj[d] - ri[d] : ri[d] - rj[d];
        pbc<D>(c, delta);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double gr;
          if (f.pg_img[d]) value[d] = pg_images<double>(f.pg_alpha[d], f.pg_len[d], delta[d], &gr);
          else value[d] = pg_eval(f.pg_nome[d], f.pg_pref[d], f.pg_k[d], f.pg_nmax[d], delta[d], &gr);
          grad[d] = ratio<double>(gr, value[d]);
        }
        if (pass == 0) {
#pragma unroll
          for (int d = 0; d < D; ++d) grad[d] *= value[d];
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k)
#pragma unroll
            for (int d = 0; d < D; ++d) grad[d] *= value[k];
        }
        const double w = mt[jp * ins + ip];
#pragma unroll
        for (int d = 0; d < D; ++d) lg[d] += w * grad[d];
      }
      double g2 = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) g2 += lg[d] * lg[d];
      dist[((int64_t)t * 2 + pass) * n + aa] = sqrt(2 * f.mass / ((g2 + 1e-15) * c.tau));
    }
    __syncthreads();
  }

  // MatrixUpdate::acceptLastMove (FreeParticleNodes.cc:1026-1057): slices
  // 1..S-2 take the updated inverses, or are updated in place when no
  // level-0 nodal action made them; the last slice keeps its inverse
  __device__ __forceinline__ void upd_accept(int fa) {
    const int n = c.fn[fa].ns, nt = c.S - 2;
    if (nt <= 0) return;
    double* bi = upd_inv(fa);
    if (upd_block(fa)[0] != 0.0) {
      const double* nw = upd_new(fa);
      for (int q = tid; q < nt * n * n; q += NT) {
        const int t = 1 + q / (n * n), rem = q % (n * n);
        const int64_t e = ((int64_t)t * ins + rem / n) * ins + rem % n;
        bi[e] = nw[e];
      }
      __syncthreads();
    } else {
      upd_rows(fa, bi, 1, nt);
    }
  }

  // Node record of section slice j for node action fa.  Section 1 slice j is
  // absolute slice a and section 2 slice j is a + M/2 (mod M), so slices
  // t and t + M/2 share one record, kept with the lower slice as r1.  Both
  // models' density matrices are symmetric, rho(r, r') = rho(r', r): seen
  // from the upper slice the matrix is the transpose, with the same
  // determinant and the two distance sides exchanged.  lower tells whether
  // section 1 holds the lower slice.
  __device__ __forceinline__ double* slot_rec(int fa, int j, bool& lower) const {
    int a = sec_t0 + j;
    a = a < 0 ? a + c.M : (a >= c.M ? a - c.M : a);
    const int half = c.M >> 1;
    lower = a < half;
    int64_t off = 0;
    for (int k = 0; k < fa; ++k) off += (int64_t)half * fn_record_doubles(c.fn[k].ns);
    return fcache + off + (int64_t)(lower ? a : a - half) * fn_record_doubles(c.fn[fa].ns);
  }

  // initialize from the records when every slice of the section has one:
  // the rescale loop runs on the recorded determinants, det(s) = det(s_r)
  // (s/s_r)^n, in slice order as the evaluating loop does.  Returns false
  // (nothing changed) when a record is missing.
  __device__ __forceinline__ bool fn_init_cached(int fa, int& nerr) {
    const FnP& f = c.fn[fa];
    const int n = f.ns, S = c.S;
    double* dm = fn_dm(fa);
    double* flag = fn_newdm(fa);
    double* dd = fn_dist(fa);
    double* nd0 = fn_newdist(fa);
    // every record of the section in one pass: determinant, scale, flag and
    // the distances (the latter are overwritten when a record is missing),
    // as one index space of single loads, four in flight per thread before
    // their stores.  The threads storing the flags note a missing record,
    // and slice 0's distances also go to newDist (newDist(0) = dist(0)).
    int missing = 0;
    const int nrec = 3 * S;
    const int tot = nrec + (f.nodal ? S * 2 * n : 0);
    for (int q0 = tid; q0 < tot; q0 += 4 * NT) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = q0 + u * NT;
        v[u] = 0.0;
        if (q < nrec) {
          const int j = q / 3;
          bool lower;
          v[u] = slot_rec(fa, j, lower)[q - 3 * j];
        } else if (q < tot) {
          const int qd = q - nrec;
          const int j = qd / (2 * n), pass = (qd / n) & 1, k = qd % n;
          bool lower;
          const double* r = slot_rec(fa, j, lower);
          v[u] = r[3 + (lower ? pass : 1 - pass) * n + k];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = q0 + u * NT;
        if (q < nrec) {
          const int j = q / 3, fld = q - 3 * j;
          (fld == 0 ? dm : fld == 1 ? wscale : flag)[j] = v[u];
          if (fld == 2) missing |= v[u] != 1.0;
        } else if (q < tot) {
          const int qd = q - nrec;
          dd[qd] = v[u];
          if (qd < 2 * n) nd0[qd] = v[u];
        }
      }
    }
    // every slice has a record
    if (__syncthreads_or(missing)) return false;
    // the rescale loop of initialize on the recorded determinants,
    // det(s) = det(s_r) (s/s_r)^n.  The first wave evaluates every slice at
    // the current scale at once; when none leaves [1e-50, 1e50] the loop
    // makes no pass and those are its values.  Otherwise one thread runs it.
    if (wave == 0) {
      const double scl0 = scale[fa];
      bool in = true;
      double dt = 0;
      if (S <= PQ_WAVE && lane < S) {
        const double lr = (double)n * log(scl0 / wscale[lane]);
        dt = dm[lane] * exp(lr);
        const double ad = fabs(dt);
        in = !(ad < 1e-50 || ad > 1e50);
      }
      const bool fast = S <= PQ_WAVE && __builtin_amdgcn_ballot_w64(!in) == 0;
      if (fast) {
        const double dm0 = lane_value(dt, 0);
        const bool e = lane < S && dt * dm0 <= 1e-200;
        const int errs = __builtin_popcountll(__builtin_amdgcn_ballot_w64(e));
        if (lane < S) { dm[lane] = dt; wscale[lane] = scl0; }
        if (lane == 0) { werr[0] = errs; lasterr[fa] = 0; flag[0] = dm0; }
      } else if (lane == 0) {
        double scl = scl0, dm0 = 0;
        int errs = 0, passes = 0;
        for (int j = 0; j < S;) {
          const double lr = (double)n * log(scl / wscale[j]);
          const double dt = dm[j] * exp(lr);
          const double ad = fabs(dt);
          if (ad < 1e-50 || ad > 1e50) {
            ++passes;
            const double ld = log(fabs(dm[j])) + lr;   // log|det| at the current scale
            if (ld < -575.6463) scl *= 2;               // |det| < 1e-250
            else scl *= exp(-ld / n);
            if (passes < 200) continue;
          }
          dm[j] = dt;
          wscale[j] = scl;
          if (j == 0) dm0 = dt;
          if (dt * dm0 <= 1e-200) errs++;
          passes = 0;
          ++j;
        }
        scale[fa] = scl;
        werr[0] = errs;
        lasterr[fa] = 0;
        flag[0] = dm0;   // newDMValue(0) = dmValue(0)
      }
    }
    __syncthreads();
    nerr += werr[0];
    return true;
  }
  // write the records of section slices j0..j1-1 from dm/scale/dist; a slice
  // whose evaluation erred (flag) gets no record
  __device__ __forceinline__ void fn_store_records(int fa, int j0, int j1, const double* det, const double* scl,
                                   double scl0, const double* errflag, const double* dist) {
    const int n = c.fn[fa].ns, rl = fn_record_doubles(n);
    for (int q = tid; q < (j1 - j0) * rl; q += NT) {
      const int j = j0 + q / rl, e = q % rl;
      bool lower;
      double* r = slot_rec(fa, j, lower);
      double v;
      if (e == 0) v = det[j];
      else if (e == 1) v = scl ? scl[j] : scl0;
      else if (e == 2) v = (errflag && errflag[j] != 0.0) ? 0.0 : 1.0;
      else {
        const int side = (e - 3) / n, k = (e - 3) % n;
        v = c.fn[fa].nodal ? dist[((int64_t)j * 2 + (lower ? side : 1 - side)) * n + k] : 0.0;
      }
      r[e] = v;
    }
  }

  // FixedNodeAction::initialize (FixedNodeAction.cc:331-357) with the
  // evaluate(...,scaleMagnitude=true) rescale loop (SHONodes.cc:113-126):
  // slices are evaluated in passes with the current scale; on the first slice
  // whose |det| leaves [1e-50,1e50] the scale is updated and evaluation
  // resumes from that slice, exactly as the sequential loop would.  In the
  // sampler the slice-pair records replace the evaluation when all are
  // present.  Returns the number of crossed or singular slices.
  __device__ __forceinline__ int fn_initialize() {
    int nerr = 0;
    constexpr int NG = NT / G;
    double* const inv_shared = inv;
    for (int fa = 0; fa < c.nfn; ++fa) {
      const FnP& f = c.fn[fa];
      double* dm = fn_dm(fa);
      double* eflag = fn_newdm(fa);   // per-slice error flags until the end
      // with useUpdates the inverses stay in the node action's own block for
      // the section's moves, and every slice is evaluated
      inv = upd(f) ? upd_inv(fa) : inv_shared;
      const bool recs = fcache && rec_on && !upd(f);
      const bool cached = recs && fn_init_cached(fa, nerr);
      if (!cached) {
        double scl = scale[fa];
        int t = 0, passes = 0;
        int errs = 0;
        double dm0 = 0;
        while (t < c.S) {
          const int nt = min(NG, c.S - t);
          node_matrices(fa, t, 1, nt, false, scl);
          int bad = -1;
          for (int m = 0; m < nt; ++m) {
            double dt = wdet[m];
            if (werr[m] || fabs(dt) < 1e-50 || fabs(dt) > 1e50) { bad = m; break; }
          }
          const int ncommit = (bad < 0) ? nt : bad;
          for (int m = 0; m < ncommit; ++m) {
            if (tid == 0) { dm[t + m] = wdet[m]; wscale[t + m] = scl; eflag[t + m] = werr[m] ? 1.0 : 0.0; }
            if (t + m == 0) dm0 = wdet[m];
            if (werr[m] || wdet[m] * dm0 <= 1e-200) errs++;
          }
          if (ncommit) passes = 0;
          t += ncommit;
          if (bad >= 0) {
            ++passes;
            const double dt = wdet[bad];
            const int er = werr[bad];
            if (!er) {
              if (fabs(dt) < 1e-250) scl *= 2;
              else scl *= pow(fabs(dt), -1. / f.ns);
            }
            if (er || passes >= 200) {
              // the reference repeats an erring evaluation unchanged; the
              // oracle stops after 200 passes: commit the slice as it stands
              if (tid == 0) { dm[t] = dt; wscale[t] = scl; eflag[t] = er ? 1.0 : 0.0; }
              if (t == 0) dm0 = dt;
              if (er || dt * dm0 <= 1e-200) errs++;
              ++t;
              passes = 0;
            }
          }
          __syncthreads();
        }
        nerr += errs;
        if (tid == 0) {
          scale[fa] = scl;
          if (lasterr) lasterr[fa] = eflag[c.S - 1] != 0.0;
        }
        __syncthreads();
        if (f.nodal) node_distances(fa, 0, c.S, false, wscale, 0.0, fn_dist(fa));
        if (recs) fn_store_records(fa, 0, c.S, dm, wscale, 0.0, eflag, fn_dist(fa));
        __syncthreads();
      }
      inv = inv_shared;
      if (!cached) {
        // newDMValue(0)=dmValue(0); newDist(0)=dist(0) (the cached path set them)
        const int n = f.ns;
        double* nd = fn_newdist(fa);
        const double* od = fn_dist(fa);
        for (int k = tid; k < 2 * n; k += NT) nd[k] = od[k];
        if (tid == 0) fn_newdm(fa)[0] = dm[0];
        __syncthreads();
      }
    }
    return nerr;
  }

  // CompositeDoubleAction::getActionDifference over FixedNodeAction::
  // getActionDifference (FixedNodeAction.cc:39-131) for the moving particles.
  __device__ __forceinline__ double fn_diff(int level) {
    double total = 0;
    for (int fa = 0; fa < c.nfn; ++fa) {
      const FnP& f = c.fn[fa];
      bool mine = false;
      for (int m = 0; m < nmov(); ++m) mine |= mv(m) >= f.ifirst && mv(m) < f.ifirst + f.ns;
      if (tid == 0) notmine[fa] = mine ? 0 : 1;
      if (!mine) continue;
      const int nStride = 1 << (level + 1);
      const int t0 = nStride / 2;
      const int nt = (c.S - 1 - t0) / nStride + 1;
      const double dm0 = fn_dm(fa)[0];
      int crossed = 0;
      double* ndm = fn_newdm(fa);
      if (upd(f)) {
        upd_change(fa, t0, nStride, nt);
        for (int m = 0; m < nt; ++m) crossed |= ndm[t0 + m * nStride] * dm0 <= 1e-200;
      } else if (top_known && level == c.L) {
        // the top level evaluates only the last slice, which no level moves:
        // its determinant is the one initialize committed at the current scale
        const double dl = fn_dm(fa)[c.S - 1];
        crossed = lasterr[fa] || dl * dm0 <= 1e-200;
        if (tid == 0) ndm[c.S - 1] = dl;
      } else {
        node_matrices(fa, t0, nStride, nt, true, scale[fa]);
        for (int m = 0; m < nt; ++m) {
          if (werr[m] || wdet[m] * dm0 <= 1e-200) { crossed = 1; break; }
          if (tid == 0) ndm[t0 + m * nStride] = wdet[m];
        }
      }
      __syncthreads();
      if (crossed) { total += 2e100; continue; }
      double dA = 0;
      if (f.nodal && level == 0 && upd(f)) {
        double* nd = fn_newdist(fa);
        upd_new_inverse(fa);
        upd_distances(fa, nd);
        dA = nodal_action(f, fn_dist(fa), nd);
      } else if (f.nodal && level == 0) {
        double* nd = fn_newdist(fa);
        const int n = f.ns;
        const double* od = fn_dist(fa);
        if (top_known) {
          // the last slice keeps its distances
          for (int k = tid; k < 2 * n; k += NT) nd[(int64_t)(c.S - 1) * 2 * n + k] = od[(int64_t)(c.S - 1) * 2 * n + k];
          node_distances(fa, 1, c.S - 2, true, nullptr, scale[fa], nd);
        } else {
          node_distances(fa, 1, c.S - 1, true, nullptr, scale[fa], nd);
        }
        dA = nodal_action(f, od, nd);
      }
      total += dA;
    }
    return total;
  }

  // FixedNodeAction::getActionDifference evaluates, at each level, the
  // determinants of the slices that level moves (FixedNodeAction.cc:59-74).
  // The sampler proposes every level of a move before trying any of them, so
  // the slices can be evaluated ahead of their level: fn_diff_moved evaluates
  // them coarse level first, a pass of the team's matrices at a time (the
  // pass a level needs is filled with the next levels' slices), and a move
  // rejected at a coarse level never evaluates the finer ones.  newdm(t)
  // holds what the level that moves slice t evaluates; an evaluation that
  // errs is stored as 0, which the sign test rejects, as the reference
  // rejects the error.  The top slice comes from initialize when no cycle
  // changed it.  fn_eval_moved starts a move (or the delayed second
  // proposal) with nothing evaluated.
  __device__ __forceinline__ void fn_eval_moved() {
    moved_done = 0;
    for (int fa = 0; fa < c.nfn; ++fa) {
      const FnP& f = c.fn[fa];
      bool mine = false;
      for (int m = 0; m < nmov(); ++m) mine |= mv(m) >= f.ifirst && mv(m) < f.ifirst + f.ns;
      if (tid == 0) notmine[fa] = mine ? 0 : 1;
      if (mine && top_known && !upd(f) && tid == 0) fn_newdm(fa)[c.S - 1] = lasterr[fa] ? 0.0 : fn_dm(fa)[c.S - 1];
    }
    __syncthreads();
  }

  // evaluate node action fa's moved slices up to those the given level tests
  __device__ __forceinline__ void fn_eval_level(int fa, int level) {
    if (upd(c.fn[fa])) {
      // MatrixUpdate::evaluateChange on the level's slices, the last slice
      // at the top level included (FixedNodeAction.cc:59-63)
      const int nStride = 2 << level;
      const int t0 = nStride / 2;
      upd_change(fa, t0, nStride, (c.S - 1 - t0) / nStride + 1);
      return;
    }
    const int done = (moved_done >> (8 * fa)) & 0xFF;
    const int need = moved_needed(level);
    if (done >= need) return;
    constexpr int PASS = NC == 32 ? NT / PQ_WAVE : NT / G;
    const int total = moved_needed(0);
    const int upto = min(total, done + (need - done + PASS - 1) / PASS * PASS);
    node_matrices(fa, done, 0, upto - done, true, scale[fa]);
    double* ndm = fn_newdm(fa);
    for (int m = tid; m < upto - done; m += NT) ndm[moved_slice(done + m)] = werr[m] ? 0.0 : wdet[m];
    __syncthreads();
    moved_done = (moved_done & ~(0xFFu << (8 * fa))) | ((uint32_t)upto << (8 * fa));
  }

  // FixedNodeAction::getActionDifference (FixedNodeAction.cc:39-131) at one
  // level from the determinants fn_eval_moved left in newdm; level 0 adds the
  // nodal action from the distances of every slice.
  // the nodal action of FixedNodeAction::getActionDifference at level 0
  // (FixedNodeAction.cc:75-126) over links t-1..t, t = 1..S-1, of both
  // sections, old distances od against new nd ([S][2][ns]): per particle
  // log[(1 - e^{-d(t) d(t-1)}) / (1 - e^{-d'(t) d'(t-1)})], or with
  // useManyBodyDistance one term per link and section with the distance
  // 1/sqrt(sum_i 1/d_i^2), where the other species' particles enter with the
  // node models' 200 (SHONodes.cc:132, FreeParticleNodes.cc:187)
  __device__ __forceinline__ double nodal_action(const FnP& f, const double* od, const double* nd) {
    const int n = f.ns;
    double part = 0;
    if (f.manybody) {
      const double others = (double)(c.N - n) / (200.0 * 200.0);
      for (int it = tid; it < (c.S - 1) * 2; it += NT) {
        const int t = 1 + (it >> 1), pass = it & 1;
        const double* o1 = od + ((int64_t)t * 2 + pass) * n;
        const double* o0 = od + ((int64_t)(t - 1) * 2 + pass) * n;
        const double* n1 = nd + ((int64_t)t * 2 + pass) * n;
        const double* n0 = nd + ((int64_t)(t - 1) * 2 + pass) * n;
        double so1 = 0, so0 = 0, sn1 = 0, sn0 = 0;
        for (int k = 0; k < n; ++k) {
          so1 += 1. / (o1[k] * o1[k]);
          so0 += 1. / (o0[k] * o0[k]);
          sn1 += 1. / (n1[k] * n1[k]);
          sn0 += 1. / (n0[k] * n0[k]);
        }
        so1 += others; so0 += others; sn1 += others; sn0 += others;
        part += log((1 - exp(-1. / sqrt(so1 * so0))) / (1 - exp(-1. / sqrt(sn1 * sn0))));
      }
    } else {
      for (int it = tid; it < (c.S - 1) * n; it += NT) {
        const int t = 1 + it / n, k = it % n;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          const int64_t a1 = ((int64_t)t * 2 + pass) * n + k;
          const int64_t a0 = ((int64_t)(t - 1) * 2 + pass) * n + k;
          part += log((1 - exp(-od[a1] * od[a0])) / (1 - exp(-nd[a1] * nd[a0])));
        }
      }
    }
    return bsum(part);
  }

  __device__ __forceinline__ double fn_diff_moved(int level) {
    double total = 0;
    for (int fa = 0; fa < c.nfn; ++fa) {
      if (notmine[fa]) continue;
      const FnP& f = c.fn[fa];
      fn_eval_level(fa, level);
      const int nStride = 1 << (level + 1);
      const double dm0 = fn_dm(fa)[0];
      const double* ndm = fn_newdm(fa);
      bool crossed = false;
      for (int t = nStride / 2; t < c.S; t += nStride) crossed |= ndm[t] * dm0 <= 1e-200;
      if (crossed) { total += 2e100; continue; }
      double dA = 0;
      if (f.nodal && level == 0) {
        double* nd = fn_newdist(fa);
        const int n = f.ns;
        const double* od = fn_dist(fa);
        if (upd(f)) {
          // evaluateNewInverse + evaluateNewDistance (FixedNodeAction.cc:79-91)
          upd_new_inverse(fa);
          upd_distances(fa, nd);
        } else if (top_known) {
          for (int k = tid; k < 2 * n; k += NT) nd[(int64_t)(c.S - 1) * 2 * n + k] = od[(int64_t)(c.S - 1) * 2 * n + k];
          node_distances(fa, 1, c.S - 2, true, nullptr, scale[fa], nd);
        } else {
          node_distances(fa, 1, c.S - 1, true, nullptr, scale[fa], nd);
        }
        dA = nodal_action(f, od, nd);
      }
      total += dA;
    }
    return total;
  }

  // FixedNodeAction::acceptLastMove (FixedNodeAction.cc:359-367); in the
  // sampler the moved slices' records take the accepted values.  A move that
  // permuted the last slice evaluated it without error, so the error flag
  // initialize gave that slice no longer applies.
  __device__ __forceinline__ void fn_accept(bool permuted = false) {
    for (int fa = 0; fa < c.nfn; ++fa) {
      if (notmine[fa]) continue;
      if (permuted && lasterr && tid == 0) lasterr[fa] = 0;
      const int n = c.fn[fa].ns;
      double* dm = fn_dm(fa);
      const double* ndm = fn_newdm(fa);
      double* dd = fn_dist(fa);
      const double* nd = fn_newdist(fa);
      for (int k = tid; k < c.S; k += NT) dm[k] = ndm[k];
      for (int k = tid; k < 2 * c.S * n; k += NT) dd[k] = nd[k];
      if (upd(c.fn[fa])) upd_accept(fa);
      else if (fcache && rec_on) fn_store_records(fa, 1, c.S - 1, ndm, nullptr, scale[fa], nullptr, nd);
    }
    __syncthreads();
  }
};

}  // namespace pq
