// pq_ewald.h -- host side of the optimized Ewald sum of CoulombAction
// (useEwald, ewaldType optEwald): the fitted short-range function f(r) and
// the half-space k-vectors with their weights, which the device sums over.
#pragma once
#include <vector>

namespace pq {

constexpr int kEwaldPoly = 8;   // OptEwaldSum npoly (CoulombAction.cc:36)

struct EwaldFit {
  int D = 3;
  bool trad = false;              // TradEwaldSum: f(r) = erf(kappa r)/r
  double kappa = 0;
  double rcut = 0, kcut = 0;
  double coef[kEwaldPoly] = {};   // f(r) = sum_i coef[i] r^2i for r <= rcut, 1/r beyond
  double oneOver2V = 0;
  std::vector<double> kvec;       // [nk][3] k-vector components of the half space
  std::vector<double> vk;         // [nk] f(|k|)
  int nk() const { return (int)vk.size(); }
  double fr(double r) const;      // evalFR
  double fr0() const;             // evalFR0
  double fk0() const;             // evalFK0
  // EwaldSum::evalSelfEnergy with charges q[0..n)
  double self_energy(const double* q, int n) const;
};

// OptEwaldSum(cell, npart, rcut, kcut, 4 kcut, 8) for a cell of sides a[0..D)
// (src/util/OptEwaldSum.cc:27-173, src/util/EwaldSum.cc:16-97).  Returns
// false when the fit is not finite (the 2D series of OptEwaldSum.cc:92-118
// loses every digit once k rcut in the halo nears 30).
bool ewald_build(EwaldFit& e, int D, const double* a, double rcut, double kcut);

// TradEwaldSum(cell, npart, rcut, kcut, kappa) (src/util/TradEwaldSum.cc:13-36):
// f(r) = erf(kappa r)/r with the reference's own erf, f(k) = 4 PI
// exp(-k^2/4 kappa^2)/k^2, on the same k-vectors.
bool ewald_build_trad(EwaldFit& e, int D, const double* a, double rcut, double kcut, double kappa);

// erf by the incomplete gamma function P(1/2, x^2) (H. Okumura, C-gengo ni
// yoru saishin algorithm jiten, 1991, p. 227), the erf the reference links
// (src/util/erf.cpp)
double okumura_erf(double x);

}  // namespace pq
