/*
 * piqmc_hip.h -- C ABI of the MI355X (gfx950) bisection-move action library.
 *
 * The library evaluates the multilevel-bisection Monte Carlo hot path of
 * pi-qmc on one GPU: spring, harmonic-oscillator and tabulated Coulomb pair
 * action differences, the fixed-node action with SHO or free-particle
 * density-matrix determinants, the free-particle midpoint mover, the
 * Metropolis decision per level, and the thermodynamic-energy and g(r)
 * estimators.  Every entry point takes plain pointers and sizes; host arrays
 * are borrowed for the duration of the call and the library owns all device
 * memory.  Every call returns PQ_OK (0) or a negative status; the message of
 * the last failure on the calling thread is returned by pq_last_error().
 *
 * Reference interfaces replaced (paths relative to the pi-qmc source tree):
 *   Action::getActionDifference            src/action/Action.h:44
 *   Action::getBeadAction                  src/action/Action.h:53-54
 *   DoubleAction / FixedNodeAction         src/action/DoubleAction.h:32-58,
 *                                          src/fixednode/FixedNodeAction.cc:39-131,331-367
 *   NodeModel::evaluate / evaluateDistance src/fixednode/NodeModel.h:93-100
 *   Mover::makeMove                        src/advancer/mover/Mover.h:7-17
 *   MultiLevelSampler::run / tryMove       src/advancer/MultiLevelSampler.cc:68-177
 *   DoubleMLSampler::run / tryMove         src/advancer/DoubleMLSampler.cc:66-245
 *   SectionChooser::run                    src/advancer/SectionChooser.cc:30-48
 *   DoubleSectionChooser::run              src/advancer/DoubleSectionChooser.cc:33-65
 *   ThermoEnergyEstimator::handleLink      src/estimator/ThermoEnergyEstimator.cc:32-54
 *   PairCFEstimator<1>::handleLink         src/estimator/PairCFEstimator.h:72-93
 *   VirialEnergyEstimator, CoulombEnergyEstimator, EwaldCoulombEstimator
 *                                          src/estimator/{Virial,Coulomb,EwaldCoulomb}*.cc
 */
#ifndef PIQMC_HIP_H
#define PIQMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQ_OK 0
#define PQ_ERR_ARG (-1)         /* invalid argument or unsupported configuration */
#define PQ_ERR_HIP (-2)         /* HIP runtime failure */
#define PQ_ERR_NOMEM (-3)       /* device or host allocation failed */
#define PQ_ERR_STATE (-4)       /* call made in the wrong order */

#define PQ_MAX_SPECIES 4
#define PQ_MAX_SHO 4
#define PQ_MAX_NODES 4
#define PQ_MAX_SAMPLES 8
#define PQ_MAX_MOVING 8
#define PQ_MAX_NORDER 4
#define PQ_MAX_NODE_PART 32     /* largest fermion species handled by a node model */
#define PQ_MAX_GROUPS 4         /* <ChooseSection> groups per sweep */

#define PQ_MOVER_FREE 0         /* FreeMover      (src/advancer/mover/FreeMover.cc:59-124) */
#define PQ_MOVER_FREE_PBC 1     /* FreeMoverPBC   (src/advancer/mover/FreeMoverPBC.cc:59-162) */

#define PQ_NODES_FREE 0         /* FreeParticleNodes (src/fixednode/FreeParticleNodes.cc:72-234) */
#define PQ_NODES_SHO 1          /* SHONodes          (src/fixednode/SHONodes.cc:51-170) */

/* Particle and permutation choosers of a <Sample npart=nmoving>
   (src/parser/PIMCParser.cc:364-453). */
#define PQ_CHOOSER_SPECIES 0    /* species given, noPermutation="true": PermutationChooser +
                                   SpeciesParticleChooser (SpeciesParticleChooser.cc:19-31) */
#define PQ_CHOOSER_WALKING 1    /* species given, noPermutation absent or false: WalkingChooser
                                   (WalkingChooser.cc:39-142), an nmoving-cycle */
#define PQ_CHOOSER_ALL 2        /* species "" or "all", noPermutation="true": PermutationChooser +
                                   SimpleParticleChooser (SimpleParticleChooser.cc:15-27) */
#define PQ_CHOOSER_ALL_RANDOM 3 /* species "" or "all", noPermutation false: RandomPermutationChooser
                                   (RandomPermutationChooser.cc:12-24) + SimpleParticleChooser */

#define PQ_PREC_FP64 0          /* every operation in FP64 */
#define PQ_PREC_MIXED 1         /* FP32 spring/SHO/pair link arithmetic with compensated
                                   sums; FP64 mover, node matrices, sign test, exponent */

#define PQ_RNG_PHILOX 0         /* Philox4x32-10 counter stream per chain */
#define PQ_RNG_INJECTED 1       /* uniforms supplied by the caller, consumed in order */

/* One particle species (<Species> element; src/base/Species.h). */
typedef struct pq_species_desc {
  int count;
  double mass;
  double charge;
  int is_static;       /* isStatic="true": SpringAction and the virial estimator skip it
                          (src/action/SpringAction.cc:69,115, VirialEnergyEstimator.cc:36,51,60,66) */
} pq_species_desc;

/* <SHOAction omega=".." species=".." x y z ndim> (src/parser/ActionParser.cc:307-324). */
typedef struct pq_sho_desc {
  int species;
  double omega;
  double center[3];
  int ndim;            /* 0 means the full dimension */
} pq_sho_desc;

/* <CoulombAction epsilon norder rmin rmax ngridPoints useEwald ewaldRcut ewaldKcut>
   without exchange (src/parser/ActionParser.cc:118-153, src/action/CoulombAction.cc:19-91).
   useEwald takes the default ewaldType (optEwald) with ewaldNDim = ndim: the pair
   tables lose tau q1 q2 f(r) from u_0 and level 0 adds the k-space sum's change
   times tau/epsilon on slices 1..2^L-1 (CoulombAction.cc:28-44,107-119,240-242). */
typedef struct pq_coulomb_desc {
  int enabled;
  double epsilon;      /* 0 means 1 */
  int norder;
  double rmin;         /* 0 means 0.005/(2 mu |q1 q2|) of the first interacting pair */
  double rmax;         /* 0 means the length of the cell diagonal */
  int ngpts;           /* 0 means 500 */
  int use_ewald;       /* long-range Ewald part: OptEwaldSum (src/util/OptEwaldSum.cc:27-173)
                          or, with ewald_type 1, TradEwaldSum (src/util/TradEwaldSum.cc:13-36) */
  double ewald_rcut;   /* 0 means a[0]/2 */
  double ewald_kcut;   /* k-vectors with 0 < |k| < ewald_kcut */
  int ewald_type;      /* 0 optEwald (the default), 1 tradEwald with ewaldImages = 1 */
  double ewald_kappa;  /* tradEwald kappa; 0 means the parser's default
                          sqrt((3.6 N)^(1/6) sqrt(pi) / min a) (ActionParser.cc:137-146) */
} pq_coulomb_desc;

/* <FixedNodeAction species model omega temperature noNodalAction>
   (src/parser/ActionParser.cc:433-452,759-815). */
typedef struct pq_node_desc {
  int species;
  int model;           /* PQ_NODES_FREE or PQ_NODES_SHO */
  double omega;        /* SHONodes only */
  double temperature;  /* 0 means the simulation temperature */
  int nodal_action;    /* 1 unless noNodalAction="true" */
  int many_body_distance;  /* useManyBodyDistance="true": one distance per slice and
                              section, 1/sqrt(sum_i 1/d_i^2) over all particles
                              (FixedNodeAction.cc:99-116) */
  int use_updates;     /* FreeParticleNodes useUpdates="true": determinant ratios and
                          inverses of the moved slices by FreeParticleNodes::MatrixUpdate
                          (FreeParticleNodes.cc:894-1058, FixedNodeAction.cc:61-92,366)
                          instead of a new evaluation; PQ_NODES_FREE only */
} pq_node_desc;

/* <Sample npart species noPermutation nrepeat delayedRejection
   defaultSigmaFactor newSigmaFactor> inside <ChooseSection>
   (PIMCParser.cc:312-322). */
typedef struct pq_sample_desc {
  int species;         /* ignored by the PQ_CHOOSER_ALL* choosers */
  int nmoving;         /* npart: 1..PQ_MAX_MOVING, at most the particles it chooses from */
  int chooser;         /* PQ_CHOOSER_* */
  int nrepeat;         /* <Sample nrepeat>: tries on the same section, 0 means 1
                          (MultiLevelSampler.cc:68-98, DoubleMLSampler.cc:66-120) */
  int delayed_rejection;   /* a second proposal at level nlevel-1 after a rejection
                              (MultiLevelSampler.cc:114-146, DoubleMLSampler.cc:150-209);
                              needs PQ_MOVER_FREE (FreeMoverPBC leaves its forward
                              probability unset, FreeMoverPBC.h:18) */
  double sigma_factor;     /* defaultSigmaFactor: the movers' sigma scale, 0 means 1 */
  double new_sigma_factor; /* newSigmaFactor of the second proposal, 0 means 0.75 */
} pq_sample_desc;

/* A further <ChooseSection nlevel> of the sweep with its <Sample> elements,
   repeated nsection times (a <Loop nrepeat> around it), e.g. the hydrogen-atom
   input's nlevel 7, 3 x nlevel 6 and 5 x nlevel 5 sections between Measures
   (test/system/hatom/pimc.xml). */
typedef struct pq_section_group {
  int nlevel;
  int nsample;
  pq_sample_desc sample[PQ_MAX_SAMPLES];
  int nsection;
} pq_section_group;

/* One Markov chain (one clone of the reference).  Particle i of species s has
   global index species_first(s)+i, species in the order given.  A sweep runs
   nsection sections of (nlevel, sample[0..nsample)) and then, in order, the
   ngroup further groups. */
typedef struct pq_chain_desc {
  int ndim;            /* 2 or 3 */
  int nslice;          /* M */
  int nlevel;          /* L of <ChooseSection nlevel>; a section holds 2^L+1 slices */
  int nspecies;
  double tau;
  double cell[3];      /* SuperCell a */
  pq_species_desc species[PQ_MAX_SPECIES];
  int spring;          /* <SpringAction> present */
  int nsho;
  pq_sho_desc sho[PQ_MAX_SHO];
  pq_coulomb_desc coulomb;
  int nnodes;
  pq_node_desc nodes[PQ_MAX_NODES];
  int mover;           /* PQ_MOVER_* */
  int nsample;         /* <Sample> elements per <ChooseSection>, run in order */
  pq_sample_desc sample[PQ_MAX_SAMPLES];
  int nsection;        /* <ChooseSection> repetitions per sweep */
  int ngroup;          /* further section groups, 0..PQ_MAX_GROUPS-1 */
  pq_section_group group[PQ_MAX_GROUPS - 1];
} pq_chain_desc;

/* Per-chain counters kept on the device by the sampler. */
typedef struct pq_chain_stats {
  uint64_t draws;                /* uniforms consumed */
  uint64_t bead_updates;         /* beads proposed and evaluated */
  uint64_t accepted_bead_updates;
  uint64_t tries[8];             /* per level: moves that reached this level */
  uint64_t accepts[8];           /* per level: moves accepted at this level */
  uint64_t node_errors;          /* crossings or singular matrices met in initialize */
  uint64_t sweeps;               /* sweeps run */
  uint64_t t_start, t_end;       /* device wall clock (pq_clock_hz ticks per second) when the
                                    chain's workgroup of the last sampler launch started and ended */
} pq_chain_stats;

/* A pi-qmc input file read for the device path (pq_read_xml).  The sweep is
   the body of the first <Loop> that holds a <Measure>: its <ChooseSection>
   elements, and <Loop>s of one <ChooseSection> each, become the section
   groups of one sweep, so a batch of that descriptor runs one Measure per
   sweep; nmeasure is that loop's nrepeat and nblock the nrepeat of an
   enclosing <Loop> with a <Collect>.  The first <Loop> before it without a
   Measure is the thermalisation, a descriptor of its own repeated
   nthermalize times. */
#define PQ_XML_MAX_CF_SPECIES 8
typedef struct pq_xml_input {
  pq_chain_desc sweep;
  int nmeasure;
  int nblock;
  int has_thermalize;
  pq_chain_desc thermalize;
  int nthermalize;
  int has_seed;            /* <RandomGenerator>; iseed 0 when the attribute is absent (srand(0)) */
  int iseed;
  char read_paths[256];    /* <ReadPaths file>, empty when absent */
  int read_bfactor;        /* 1 when absent */
  int thermal_energy;      /* <ThermalEnergyEstimator> */
  int virial_nwindow;      /* <VirialEnergyEstimator nwindow> (absent: 1), 0: no estimator */
  int coulomb_energy;      /* <CoulombEnergyEstimator epsilon useEwald rcut kcut ewaldType kappa> */
  double coulomb_epsilon;
  int coulomb_ewald;
  double coulomb_rcut, coulomb_kcut;
  int coulomb_ewald_type;  /* 0 optEwald, 1 tradEwald */
  double coulomb_kappa;
  int pair_cf_nspecies;    /* <PairCFEstimator nspecies species1..> with <Radial nbin min max>; 0: none */
  int pair_cf_species[PQ_XML_MAX_CF_SPECIES];
  double pair_cf_rmin, pair_cf_rmax;
  int pair_cf_nbin;
  int nignored;            /* top-level PIMC elements with no device meaning (path
                              initialisers, writers), left to the caller */
} pq_xml_input;

typedef struct pq_batch pq_batch;

const char* pq_last_error(void);

/* Reads a pi-qmc input file into descriptors for a reference binary built
   with NDIM = ndim, following SimInfoParser (SimInfoParser.cc:30-105),
   ActionParser::parseActions / parseNodeModel (ActionParser.cc:102-456,
   759-815), EstimatorParser (EstimatorParser.cc:150-215,477-488,643-725) and
   PIMCParser::parseAlgorithm (PIMCParser.cc:125-490,560-640): attribute
   values as the reference reads them, units scaled to Hartree atomic units
   with the reference's table.  Elements and attributes that the device path
   does not model (other actions, node models or movers, exchange, image
   Ewald sums, ...) are refused with PQ_ERR_ARG and a message naming them,
   never skipped.  pq_parse_xml reads the same from a string.  Touches no
   device. */
int pq_read_xml(const char* filename, int ndim, pq_xml_input* out);
int pq_parse_xml(const char* text, int ndim, pq_xml_input* out);

/* Builds pair tables, uploads parameters and allocates per-chain device state
   (beads slice-major [M][N][D] in FP64).  precision is PQ_PREC_*. */
int pq_batch_create(const pq_chain_desc* chains, int nchain, int precision,
                    int device, pq_batch** out);
void pq_batch_destroy(pq_batch* batch);
int pq_batch_info(const pq_batch* batch, int chain, int* npart, int* nslice,
                  int* ndim, int64_t* bead_offset);

/* Path coordinates of one chain, slice-major [M][N][D] FP64
   (Beads<NDIM> layout, src/base/Beads.h:215,225; identity permutation). */
int pq_set_paths(pq_batch* batch, int chain, const double* beads);
int pq_get_paths(pq_batch* batch, int chain, double* beads);
/* All chains at once, concatenated in chain order. */
int pq_set_all_paths(pq_batch* batch, const double* beads, int64_t count);
/* Path permutation of a chain: the bead after (i, M-1) is (perm[i], 0)
   (SerialPaths permutation / inversePermutation, src/base/SerialPaths.cc:68-117;
   ReadPaths sets it, src/algorithm/PathReader.cc:38-50,113-117).  Setting
   paths resets it to the identity; permutation sampling changes it. */
int pq_set_permutation(pq_batch* batch, int chain, const int* perm);
int pq_get_permutation(pq_batch* batch, int chain, int* perm);
int pq_get_all_paths(pq_batch* batch, double* beads, int64_t count);

/* Path files of <WritePaths file> / <ReadPaths file bfactor>
   (src/algorithm/WritePaths.cc:36-112, src/algorithm/PathReader.cc:17-132):
   a "#Permutations" line, a "#Path coordinates: N particles in D-d" line,
   then one line per slice with the N*D coordinates at 15 significant digits.
   Reading accepts the reference's files with or without the permutation
   line (without it, a trailing wrap slice) and fills copy k of the file's
   slices, k < bfactor, relabelled by the k-th power of the file's
   permutation p, as ReadPaths' putBeads of each copy's last slice leaves
   them; perm receives p^bfactor (null: a non-identity p is refused).  Writing
   takes the permutation (null: the identity).  beads are
   [nslice][npart][ndim].  These two touch no device. */
int pq_write_paths_file(const char* filename, const double* beads, const int* perm,
                        int nslice, int npart, int ndim);
int pq_read_paths_file(const char* filename, double* beads, int* perm, int nslice,
                       int npart, int ndim, int bfactor);
/* pimc.dat of an ASCII report (src/stats/ascii/AsciiReportBuilder.cc:28-43,
   AsciiScalarReportWriter.cpp:12-35): a header line "#" + "\"name (unit)\" "
   per scalar estimator, then one line per block of (value + shift) * scale
   at printf %g, each followed by a space.  values are [nblock][nscalar], the
   block means already averaged over clones; units, shift and scale may be
   null (no unit, 0, 1).  Touches no device. */
int pq_write_ascii_report(const char* filename, int nscalar, const char* const* names,
                          const char* const* units, const double* shift, const double* scale,
                          int nblock, const double* values);
/* Every chain of a batch, paths and permutation; chain c uses filename + c
   when the batch has more than one chain, as the reference appends the clone
   ID when nclone > 1. */
int pq_write_all_paths(pq_batch* batch, const char* filename);
int pq_read_all_paths(pq_batch* batch, const char* filename, int bfactor);

/* Uniform stream of every chain: Philox4x32-10 keyed by seed, stream id =
   first_stream + chain, or caller-supplied uniforms (per_chain values per
   chain, chain-major).  Replaces RandomNumGenerator (src/util/RandomNumGenerator.h:26-86).
   A chain that draws past the end of its supplied stream makes pq_sync,
   pq_get_stats and pq_get_all_stats return PQ_ERR_STATE. */
int pq_seed(pq_batch* batch, uint64_t seed, uint32_t first_stream);
int pq_set_uniforms(pq_batch* batch, const double* u, int64_t per_chain);

/* Device-resident sampler: nsweep sweeps of every chain, each sweep nsection
   x (choose section, initialize node caches, run each Sample, put beads back).
   One fused kernel launch per call; asynchronous on the batch stream. */
int pq_run_sweeps(pq_batch* batch, int nsweep);
/* Same launch bracketed by HIP events on the batch stream; waits for it and
   returns the kernel's elapsed time in milliseconds. */
int pq_run_sweeps_timed(pq_batch* batch, int nsweep, float* ms);
int pq_sync(pq_batch* batch);
/* Schedule of the sampler kernel: team = threads per chain (64 = one wave,
   128 = two waves, 256 or 512 = four or eight waves for batches of few
   chains, -1 automatic), inv_lds = node inverse cache in LDS (1), global memory (0) or
   automatic (-1).  Results agree to rounding whatever the schedule. */
int pq_set_schedule(pq_batch* batch, int team, int inv_lds);
int pq_get_schedule(const pq_batch* batch, int* team, int* inv_lds, int* lds_bytes);
/* Placement of the sampler's pair tables (u, the arithmetic type's copy) and
   fixed-node state: LDS (1), global memory (0) or automatic (-1).  Automatic
   placement fills a 40 KB LDS budget per chain (four chains per CU) with the
   tables first, then the node state, then the inverse cache. */
int pq_set_staging(pq_batch* batch, int tables_lds, int state_lds);
int pq_get_staging(const pq_batch* batch, int* tables_lds, int* state_lds);
/* The heaviest heavy_chains chains (the batch launches chains in decreasing
   order of estimated cost) run in a launch of their own with heavy_team
   threads per chain (128, 256 or 512), concurrently with the rest; 0 turns
   the split off.  The sweep's length is the heaviest chain's, so giving it
   more waves shortens every launch of an uneven batch.  heavy_chains -1
   splits at the node-matrix width instead, when chains with 17-32-particle
   node species sit beside narrower ones (the batch then launches the wide
   ones first): those run with heavy_team threads on the 32-wide instance,
   the rest on the narrower Gauss-Jordan instance with an LDS layout sized to
   them; pq_get_split reports -1 for it. */
int pq_set_split(pq_batch* batch, int heavy_chains, int heavy_team);
int pq_get_split(const pq_batch* batch, int* heavy_chains, int* heavy_team);
/* Times every distinct schedule (team 64/128/256/512 x tables, node state and
   inverse cache in LDS or global memory, then splits of the heaviest chains
   on the eight best of those) over nsweep sampler sweeps of the batch, after
   one untimed launch of nsweep sweeps; the two fastest single launches and
   the three fastest splits are timed three more times each, in turn, and the
   lowest mean wins.  best_ms receives that
   mean per sweep and ntimed the number of launches run, the untimed one
   included (ntimed * nsweep sweeps were run).
   The tuning sweeps advance the chains like pq_run_sweeps: schedules change
   only where data lives and how many threads share a chain, so results agree
   to rounding.  Replaces nothing in the reference (it has no device
   schedule). */
int pq_autotune(pq_batch* batch, int nsweep, float* best_ms, int* ntimed);
/* Sweeps per launch of each chain: chain c runs nsweep * sweeps[c] sweeps
   in every sampler launch (pq_run_sweeps, pq_run_measured and the rest),
   1 <= sweeps[c] <= 4096; a null array gives every chain one.  Chains are
   independent Markov chains, so a chain's decisions and paths after n sweeps
   do not depend on the plan; a plan that gives light chains more sweeps
   keeps the device busy while the heaviest chains run.  No counterpart in
   the reference (its clones run at their own pace in separate processes). */
int pq_set_sweep_plan(pq_batch* batch, const int* sweeps);
int pq_get_sweep_plan(const pq_batch* batch, int* sweeps);
/* Ticks per second of the device clock of pq_chain_stats.t_start/t_end. */
int pq_clock_hz(const pq_batch* batch, double* hz);
int pq_get_stats(pq_batch* batch, int chain, pq_chain_stats* out);
int pq_get_all_stats(pq_batch* batch, pq_chain_stats* out);
int pq_reset_stats(pq_batch* batch);
/* Node-model magnitude scale of each FixedNodeAction of a chain
   (SHONodes.cc:113-122 / FreeParticleNodes.cc:142-151 state). */
int pq_get_node_scale(pq_batch* batch, int chain, double* scale);

/* nblock x (nsweep sweeps of every chain, then the thermodynamic energy of
   every chain) captured once into a HIP graph and replayed on later calls
   with the same arguments and batch settings: energies[block][chain].  The
   same results as nblock x (pq_run_sweeps(nsweep), pq_measure_energy), without a
   host round trip per block (the <Loop>/<Measure> pattern). */
int pq_run_measured(pq_batch* batch, int nblock, int nsweep, double* energies);
/* pq_run_measured with sweep_ms[block] = the duration of that block's
   sweeps: each block launches a captured graph of its sampler launch between
   two HIP events on the batch stream, then its Measure kernels. */
int pq_run_measured_timed(pq_batch* batch, int nblock, int nsweep, double* energies, float* sweep_ms);
/* Adds a pair-correlation estimator (<PairCFEstimator nspecies species1..n>
   with <Radial nbin min max>, src/parser/EstimatorParser.cc:477-488,643-725)
   to every Measure of pq_run_measured: its integer counts, as
   pq_pair_cf_histogram gives them, accumulate on the device over the
   Measures; nspecies 0 removes it.  pq_collect_pair_cf adds the accumulated
   counts[chain][nbin] to the caller's array and restarts the accumulation
   (a <Collect>: the reference divides each Measure's counts by nslice and
   averages a block's Measures, which the caller does with nslice and the
   number of Measures). */
int pq_set_measure_pair_cf(pq_batch* batch, int nspecies, const int* species, double rmin,
                           double rmax, int nbin);
int pq_collect_pair_cf(pq_batch* batch, uint64_t* counts);

/* Thermodynamic energy estimate of every chain's current paths:
   E = (1/M) sum_{t,i} utau(i,t) (ThermoEnergyEstimator + CompositeAction::getBeadAction),
   with CoulombAction's Ewald k-space sum / epsilon of every slice when useEwald is set. */
int pq_measure_energy(pq_batch* batch, double* energy);
/* Coulomb potential energy of every chain's current paths
   (<CoulombEnergyEstimator epsilon>, src/estimator/CoulombEnergyEstimator.cc:30-40,
   without Ewald): (1/M) sum_t sum_{j<i} q_i q_j / (epsilon |pbc(r_i - r_j)|);
   epsilon 0 means 1, as the parser's default. */
int pq_measure_coulomb_energy(pq_batch* batch, double epsilon, double* energy);
/* Coulomb energy with the optimized Ewald sum (<CoulombEnergyEstimator useEwald="true"
   rcut kcut epsilon>, ewaldType optEwald: EwaldCoulombEstimator,
   src/estimator/EwaldCoulombEstimator.cc:63-140, src/parser/EstimatorParser.cc:169-190):
   per slice, sum_{j<i, r<rcut} q_i q_j (1/(eps r) - f(r)/eps) with f read off a
   1001-point grid the way the reference reads it, plus the k-space sum / eps;
   (1/M) over slices.  The fit is built once per (rcut, kcut) and cell. */
int pq_measure_ewald_coulomb_energy(pq_batch* batch, double epsilon, double rcut, double kcut,
                                    double* energy);
/* The same estimator with ewaldType="tradEwald" and ewaldImages 1 (the
   default): a TradEwaldSum, f(r) = erf(kappa r)/r with the reference's erf,
   kappa 0 meaning the parser's default sqrt((3.6 N)^(1/6) sqrt(pi) / Lside),
   Lside the shortest cell side (EwaldCoulombEstimator.cc:17-61,
   EstimatorParser.cc:191-209).  ewaldImages > 1 is not modelled. */
int pq_measure_trad_ewald_coulomb_energy(pq_batch* batch, double epsilon, double rcut, double kcut,
                                         double kappa, double* energy);
/* Virial energy of every chain's current paths (<VirialEnergyEstimator nwindow>,
   src/estimator/VirialEnergyEstimator.cc:27-78): utau plus the centroid-virial
   force terms of getBeadAction over windows of nwindow links (0 means 1, the
   parser's default; nwindow = 1 is the thermodynamic estimator).  The
   FixedNodeActions contribute zero force and utau, as in the reference. */
int pq_measure_virial_energy(pq_batch* batch, int nwindow, double* energy);
/* Radial pair histogram of every chain (PairCFEstimator<1>::handleLink with
   PairRadial, src/estimator/PairCFEstimator.h:72-93): integer counts over all
   slices, particle i of species[0] against every particle j != i of each
   partner species species[1..nspecies-1] in list order (a partner listed
   twice counts twice), bin int((r - rmin) nbin/(rmax - rmin)), dropped when r
   < rmin or the bin exceeds nbin - 1; counts[chain][nbin], added to what the
   array holds.  A chain that lacks one of the species counts nothing. */
int pq_pair_cf_histogram(pq_batch* batch, int nspecies, const int* species, double rmin,
                         double rmax, int nbin, uint64_t* counts);
/* The same with one partner species: species {species1, species2}. */
int pq_pair_histogram(pq_batch* batch, int species1, int species2, double rmin,
                      double rmax, int nbin, uint64_t* counts);

/* Reference call-level entry points.  A query is one section (S=2^L+1 slices,
   [S][N][D]) with the moving beads ([S][nmov][D]) and their particle indices.
   dS[q] = CompositeAction::getActionDifference (spring + SHO + Coulomb pair, and
   at level 0 the Ewald k-space term) of query q at level[q] for chain chain[q]. */
int pq_action_difference(pq_batch* batch, int nquery, const int* chain,
                         const int* level, const double* section,
                         const double* moving, const int* index, int nmov,
                         double* dS);
/* Action::getBeadAction (src/action/Action.h:53-54) of the chain's
   CompositeAction (CompositeAction.cc:57-67: SpringAction, SHOAction,
   CoulombAction's pair actions and its long-range Ewald part) for nquery
   beads (ipart[q], islice[q]), 0 <= islice < M, of caller-supplied paths:
   beads [M][N][D] with perm their path permutation (null: the identity), or,
   with beads and perm both null, the chain's device-resident paths.
   out[q][3 + 2D] = u, utau, ulambda (0 for these actions), fm[D] (the force
   from the link into slice islice), fp[D] (from the link out of it). */
int pq_bead_action(pq_batch* batch, int chain, const double* beads, const int* perm,
                   int nquery, const int* ipart, const int* islice, double* out);
/* pq_action_difference for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc); it has no counterpart in the reference.  The same queries and
   results; its staging buffers live in a 4 MiB device array per GPU,
   reserved when the library is loaded and looked up at pq_batch_create, so
   the call makes no hipMalloc or hipFree.  Queries needing more than 4 MiB
   ((S N D + S nmov D + 1) doubles + (2 + nmov) ints each, S N D the largest
   among the queried chains) are refused with PQ_ERR_ARG.  Calls on one GPU share that
   array: calls that overlap on one GPU (other host threads) race. */
int pq_action_difference_static(pq_batch* batch, int nquery, const int* chain,
                                const int* level, const double* section,
                                const double* moving, const int* index, int nmov,
                                double* dS);
/* Per-link utau of CompositeAction::getBeadAction for a set of (chain, ipart,
   islice) on the current device paths. */
int pq_bead_utau(pq_batch* batch, int nquery, const int* chain,
                 const int* ipart, const int* islice, double* utau);
/* FixedNodeAction protocol on one chain: section_begin = initialize with both
   sections ([S][N][D] each, sec2 at +M/2); double_action_difference =
   CompositeDoubleAction::getActionDifference at one level (2e100 on a crossing);
   accept_last_move = acceptLastMove. */
int pq_section_begin(pq_batch* batch, int chain, const double* sec1,
                     const double* sec2);
int pq_double_action_difference(pq_batch* batch, int chain, int level,
                                const double* moving, const int* index,
                                int nmov, double* dS);
int pq_accept_last_move(pq_batch* batch, int chain);
/* NodeModel::evaluate + evaluateDistance on nmat slice pairs (r1,r2 [nmat][N][D])
   for FixedNodeAction inode of chain, without changing the chain's scale. */
int pq_node_evaluate(pq_batch* batch, int chain, int inode, int nmat,
                     const double* r1, const double* r2, double* det,
                     int* err, double* d1, double* d2);

/* Tabulated pair action ipair of a chain (CoulombAction -> PairAction grid,
   src/action/CoulombAction.cc:45-90, src/action/PairAction.cc:161-185):
   meta = {species1, species2, norder, ngpts}, grid = {1/rmin, (ngpts-1)/ln(rmax/rmin)},
   u and ut [ngpts][norder+1]. */
int pq_get_pair_table(pq_batch* batch, int chain, int ipair, int* meta,
                      double* grid, double* u, double* ut);

#ifdef __cplusplus
}
#endif
#endif
