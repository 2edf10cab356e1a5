// Test infrastructure only: a C entry point over the reference's own
// Coulomb1DLinkAction, Coulomb3DLinkAction and PeriodicGaussian classes and
// its erf, compiled from the pi-qmc sources into oracle/_ref to cross-check
// the oracle.
#include <cmath>
#include "action/coulomb/Coulomb1DLinkAction.h"
#include "action/coulomb/Coulomb3DLinkAction.h"
#include "util/PeriodicGaussian.h"
#include "util/erf.h"

extern "C" double ref_coulomb1d(double stau, int which, double reff) {
  Coulomb1DLinkAction c(stau);
  switch (which) {
    case -1: return c.calculateValueAtOrigin();
    case 0: return c.calculateU0(c.calculateValueAtOrigin(), reff);
    case 1: return c.calculateU1(reff);
    case 2: return c.calculateU2(reff);
    case 3: return c.calculateU3(reff);
    default: return c.calculateU4(reff);
  }
}

extern "C" double ref_coulomb3d(double stau, int which, double reff) {
  Coulomb1DLinkAction c(stau);
  Coulomb3DLinkAction c3(c);
  switch (which) {
    case 0: return c3.calculateU0(reff);
    case 1: return c3.calculateU1(reff);
    case 2: return c3.calculateU2(reff);
    case 3: return c3.calculateU3(reff);
    default: return c3.calculateU4(reff);
  }
}

extern "C" void ref_pg_eval(double alpha, double length, double x, double* out) {
  PeriodicGaussian pg(alpha, length);
  out[0] = pg.evaluate(x);
  out[1] = pg.getGradient();
  out[2] = pg.getSecondDerivative();
}

extern "C" int ref_pg_nterms(double alpha, double length) {
  return PeriodicGaussian::numberOfTerms(alpha, length);
}

// the reference's incomplete-gamma erf (src/util/erf.cpp), bound inside this
// library
extern "C" double ref_erf(double x) { return erf(x); }
