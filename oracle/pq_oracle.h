/*
 * pq_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of pi-qmc's bisection-move hot path, used as the
 * checker of the HIP library (tests/, __graft_entry__.smoke(), and the
 * cpu_baseline leg of bench.py).  Product code never links or calls it.
 * Layouts: section/path beads slice-major [slice][particle][dim] FP64.
 */
#ifndef PQ_ORACLE_H
#define PQ_ORACLE_H
#include <stdint.h>
#include "../include/piqmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oq_chain oq_chain;

/* uniform streams */
void oq_glibc_sequence(unsigned int seed, int n, int32_t* out);
void oq_philox_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t oq_philox_u32(uint64_t seed, uint32_t stream, uint64_t k);
double oq_philox_uniform(uint64_t seed, uint32_t stream, uint64_t k);

/* closed-form pieces with reference known-answer tests */
double oq_coulomb1d(double stau, int which, double reff);   /* which: -1 origin, 0..4 */
double oq_coulomb3d(double stau, int which, double reff);   /* which: 0..4 */
int oq_pg_nterms(double alpha, double length);
void oq_pg_eval(double alpha, double length, double x, double out[3]);
void oq_pbc(const double cell[3], int ndim, double* v);

/* chain lifecycle */
oq_chain* oq_create(const pq_chain_desc* desc);
void oq_destroy(oq_chain* ch);
void oq_set_beads(oq_chain* ch, const double* beads);
void oq_get_beads(const oq_chain* ch, double* beads);
/* path permutation: bead (i, t + M) is bead (perm[i], t); set returns -1
   unless perm is a permutation of 0..N-1 */
int oq_set_permutation(oq_chain* ch, const int* perm);
/* Permutation::prepend (append = 0) or append (append = 1) of q onto p */
void oq_perm_compose(int* p, const int* q, int n, int append);
void oq_get_permutation(const oq_chain* ch, int* perm);
int oq_npair(const oq_chain* ch);
int oq_pair_table(const oq_chain* ch, int ipair, int* meta /*s1,s2,norder,ngpts*/,
                  double* grid /*rgridinv,logrratioinv*/, double* u, double* ut);

/* RNG modes: 0 glibc rand() clone, 1 philox(seed, stream), 2 injected */
void oq_rng_glibc(oq_chain* ch, unsigned int seed);
void oq_rng_philox(oq_chain* ch, uint64_t seed, uint32_t stream);
void oq_rng_inject(oq_chain* ch, const double* u, int64_t n);
double oq_rng_next(oq_chain* ch);
uint64_t oq_rng_count(const oq_chain* ch);
/* 1 when an injected stream was read past its end (those draws were 0.5) */
int oq_rng_exhausted(const oq_chain* ch);

/* action layer on explicit sections.  mask bit0 spring, bit1 SHO, bit2
   CoulombAction (pair tables and the Ewald term), bit3 the Ewald term alone */
double oq_action_difference(oq_chain* ch, const double* sec, const double* mov,
                            const int* idx, int nmov, int level, int mask);
double oq_bead_utau(const oq_chain* ch, int ipart, int islice, int mask);

/* node models and the fixed-node action */
int oq_node_evaluate(oq_chain* ch, int inode, const double* r1, const double* r2,
                     int islice, int scale_magnitude, double* det);
void oq_node_distance(oq_chain* ch, int inode, const double* r1, const double* r2,
                      int islice, double* d1, double* d2);
double oq_node_scale(const oq_chain* ch, int inode);
void oq_fn_initialize(oq_chain* ch, const double* sec1, const double* sec2);
double oq_fn_action_difference(oq_chain* ch, const double* mov, const int* idx,
                               int nmov, int level);
void oq_fn_accept(oq_chain* ch);
/* node action inode's determinants: dmValue (which 0) or newDMValue (1) of
   the S section slices, and its distances dist / newDist [S][2][ns] */
void oq_fn_state(const oq_chain* ch, int inode, int which, double* det, double* dist);

/* mover on explicit section/moving buffers (mov modified in place) */
double oq_mover(oq_chain* ch, const double* sec, double* mov, const int* idx,
                int nmov, int level);

/* full sampler */
void oq_run_sweeps(oq_chain* ch, int nsweep);
void oq_get_stats(const oq_chain* ch, pq_chain_stats* st);
double oq_thermo_energy(const oq_chain* ch);
double oq_coulomb_energy(const oq_chain* ch, double epsilon);
/* CoulombEnergyEstimator useEwald="true" (optEwald) with its own rcut, kcut */
double oq_ewald_coulomb_energy(const oq_chain* ch, double epsilon, double rcut, double kcut);
double oq_ewald_coulomb_energy_type(const oq_chain* ch, double epsilon, double rcut, double kcut,
                                    int trad, double kappa);
double oq_trad_estimator_kappa(const oq_chain* ch);
/* the chain's CoulombAction Ewald sum: fit, k-vector ik, self energy */
int oq_ewald_info(const oq_chain* ch, double* coef, int* nk, double* self, int ik, int* kv, double* vk);
double oq_ewald_long_range(const oq_chain* ch, const double* r);
double oq_ewald_fr(const oq_chain* ch, double r);
/* the reference's erf (src/util/erf.cpp) */
double oq_ref_erf(double x);
/* getBeadAction of the composite action: u, utau, forces fm/fp [D] */
void oq_bead_action(const oq_chain* ch, int ipart, int islice, double* u, double* utau,
                    double* fm, double* fp);
double oq_virial_energy(const oq_chain* ch, int nwindow);
void oq_energy_blocks(oq_chain* ch, int nblock, int nmeasure, double* block);
void oq_pair_histogram(const oq_chain* ch, int s1, int s2, double rmin,
                       double rmax, int nbin, uint64_t* counts);
/* species[0] against the partner species species[1..nspec-1] */
void oq_pair_cf_histogram(const oq_chain* ch, int nspec, const int* species, double rmin,
                          double rmax, int nbin, uint64_t* counts);

/* many chains on nthreads host threads (OpenMP), Philox streams
   first_stream+c; beads concatenated per chain; returns bead updates */
uint64_t oq_run_batch(const pq_chain_desc* descs, int nchain, double* beads,
                      int nsweep, uint64_t seed, uint32_t first_stream,
                      int nthreads, double* energy);

/* the same with a Measure after every sweep: thermal energy summed per chain
   and pair-correlation counts [nchain][nbin] of species[0] against
   species[1..nspec-1] */
uint64_t oq_run_batch_measured(const pq_chain_desc* descs, int nchain, double* beads, int nsweep,
                               uint64_t seed, uint32_t first_stream, int nthreads, int nspec,
                               const int* species, double rmin, double rmax, int nbin,
                               double* energy_sum, uint64_t* counts);

#ifdef __cplusplus
}
#endif
#endif
