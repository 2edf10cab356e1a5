// pq_ewald.cpp -- the Ewald sums' f(r), f(k) and k-vectors on the host
// (src/util/OptEwaldSum.cc:27-280, src/util/TradEwaldSum.cc:13-36,
// src/util/EwaldSum.cc:16-112).  The pair
// tables subtract tau q1 q2 f(r) from u_0 with this f, and the device sums
// w_k |rho_k|^2 over these k-vectors, so both follow the reference's
// arithmetic term by term; the file is built without floating-point
// contraction so that the fit comes out the same on every host.
#include "pq_ewald.h"

#include <cmath>
#include <cstring>

namespace pq {
namespace {

constexpr double kPi = 3.14159265358979;   // EwaldSum::PI (EwaldSum.cc:187)
constexpr int NP = kEwaldPoly;

// 3D: I_n = int_0^rcut r^n e^{ikr} dr by parts from I_-1 = (e^{ik rcut}-1)/(ik);
// fm1 = 4 PI Im(I_-1)/k and ft[j] = 4 PI Im(I_2j)/k, as the fit weighs them
// (OptEwaldSum.cc:76-89)
void series3d(double k, double rcut, double& fm1, double* ft) {
  const double c = std::cos(k * rcut), s = std::sin(k * rcut);
  double re = s / k, im = -(c - 1.) / k;
  fm1 = 4. * kPi * im / k;
  double rn = rcut;
  for (int n = 0; n < 2 * NP - 1; ++n) {
    const double nr = rn * c - (n + 1.) * re, ni = rn * s - (n + 1.) * im;
    re = ni / k;
    im = -nr / k;
    rn *= rcut;
    if (n % 2 == 0) ft[n / 2] = 4. * kPi * im / k;
  }
}

// 2D: the series of the Bessel-function integrals (OptEwaldSum.cc:92-118)
void series2d(double k, double rcut, double& fm1, double* ft) {
  const double factor = -0.25 * k * k * rcut * rcut;
  double term = 2. * kPi;
  double fn = term * rcut;
  for (int m = 1; std::fabs(term) > 1e-100; ++m) {
    term *= factor / (m * m);
    fn += term * rcut / (2 * m + 1);
  }
  fm1 = fn;
  double rnplus2 = rcut * rcut;
  for (int n = 0; n < 2 * NP; n += 2) {
    double t = 2. * kPi;
    double f = t * rnplus2 / (n + 2);
    for (int m = 1; std::fabs(t) > 1e-100; ++m) {
      t *= factor / (m * m);
      f += t * rnplus2 / (2 * m + n + 2);
    }
    ft[n / 2] = f;
    rnplus2 *= rcut * rcut;
  }
}

double frcut(double rcut, const double* t, int nd) {
  double v = 0;
  for (int i = 0; i < NP; ++i) {
    double x = std::pow(rcut, 2 * i - nd);
    for (int j = 0; j < nd; ++j) x *= 2 * i - j;
    v += t[i] * x;
  }
  return v;
}

// unblocked LU with partial pivoting, row-major, and its solve (the
// dgetrf/dgetrs pair of the fit)
void lu(double* a, int n, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
    if (a[k * n + k] == 0) continue;
    const double r = 1.0 / a[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = a[i * n + k] * r;
      a[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
  }
}

void lu_solve(const double* a, int n, const int* piv, double* b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= a[i * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= a[i * n + j] * b[j];
    b[i] /= a[i * n + i];
  }
}

// the incomplete gamma functions P(a, x) and Q(a, x) = 1 - P(a, x): the
// power series below x = 1 + a and the Laguerre continued fraction above,
// each summed until the sum stops changing (at most 1000 terms)
double q_gamma(double a, double x, double loggamma_a);

double p_gamma(double a, double x, double loggamma_a) {
  if (x >= 1 + a) return 1 - q_gamma(a, x, loggamma_a);
  if (x == 0) return 0;
  double term = std::exp(a * std::log(x) - x - loggamma_a) / a;
  double sum = term;
  for (int k = 1; k < 1000; ++k) {
    term *= x / (a + k);
    const double prev = sum;
    sum += term;
    if (sum == prev) break;
  }
  return sum;
}

double q_gamma(double a, double x, double loggamma_a) {
  if (x < 1 + a) return 1 - p_gamma(a, x, loggamma_a);
  double la = 1, lb = 1 + x - a;
  double w = std::exp(a * std::log(x) - x - loggamma_a);
  double sum = w / lb;
  for (int k = 2; k < 1000; ++k) {
    const double t = ((k - 1 - a) * (lb - la) + (k + x) * lb) / k;
    la = lb;
    lb = t;
    w *= (k - 1 - a) / k;
    const double prev = sum;
    sum += w / (la * lb);
    if (sum == prev) break;
  }
  return sum;
}

}  // namespace

double okumura_erf(double x) {
  const double log_gamma_half = 0.572364942924700087071713675675;   // ln(pi)/2
  if (!std::isfinite(x)) return x > 0 ? 1.0 : -1.0;
  return x >= 0 ? p_gamma(0.5, x * x, log_gamma_half) : -p_gamma(0.5, x * x, log_gamma_half);
}

double EwaldFit::fr0() const { return trad ? 2 * kappa / std::sqrt(kPi) : coef[0]; }

double EwaldFit::fr(double r) const {
  if (trad) return okumura_erf(kappa * r) / r;
  if (r > rcut) return 1. / r;
  double v = 0., r2 = r * r;
  for (int i = NP - 1; i >= 0; --i) {
    v *= r2;
    v += coef[i];
  }
  return v;
}

double EwaldFit::fk0() const {
  if (trad) return -kPi / (kappa * kappa);
  if (D == 3) {
    double fn = 2. * kPi * rcut * rcut, v = -fn;
    fn *= rcut / 3.;
    for (int n = 0; n < 2 * NP; n += 2) {
      v += coef[n / 2] * fn;
      fn *= (rcut * rcut) * (n + 3.) / (n + 5.);
    }
    return v;
  }
  double fn = 2. * kPi * rcut, v = -fn;
  fn *= rcut / 2.;
  for (int n = 0; n < 2 * NP; n += 2) {
    v += coef[n / 2] * fn;
    fn *= (rcut * rcut) * (2 + n) / (4 + n);
  }
  return v;
}

double EwaldFit::self_energy(const double* q, int n) const {
  double Q = 0, Q2 = 0;
  for (int i = 0; i < n; ++i) { Q += q[i]; Q2 += q[i] * q[i]; }
  return -0.5 * Q2 * fr0() + oneOver2V * fk0() * Q * Q;
}

// EwaldSum's half-space k-vectors with 0 < |k| < kcut (EwaldSum.cc:16-88);
// vk holds |k|^2 until the caller replaces it by f(|k|)
static void ewald_kvectors(EwaldFit& e, int D, const double* a, double kcut) {
  int ikmax[3] = {0, 0, 0};
  double dk[3] = {0, 0, 0}, vol = 1;
  for (int i = 0; i < D; ++i) {
    ikmax[i] = (int)(a[i] * kcut / (2 * kPi));
    dk[i] = 2 * kPi * (1. / a[i]);
    vol *= a[i];
  }
  e.oneOver2V = 0.5 / vol;
  e.kvec.clear();
  e.vk.clear();
  for (int kx = 0; kx <= ikmax[0]; ++kx) {
    const double kx2 = kx * kx * dk[0] * dk[0];
    for (int ky = (kx == 0) ? 0 : -ikmax[1]; ky <= ikmax[1]; ++ky) {
      const double ky2 = ky * ky * dk[1] * dk[1];
      const int kzlo = (D == 3) ? ((kx == 0 && ky == 0) ? 0 : -ikmax[2]) : 0;
      const int kzhi = (D == 3) ? ikmax[2] : 0;
      for (int kz = kzlo; kz <= kzhi; ++kz) {
        double k2 = kx2 + ky2;
        if (D == 3) k2 += kz * kz * dk[2] * dk[2];
        if (k2 < kcut * kcut && k2 != 0) {
          e.kvec.push_back(kx * dk[0]);
          e.kvec.push_back(ky * dk[1]);
          e.kvec.push_back(kz * dk[2]);
          e.vk.push_back(k2);
        }
      }
    }
  }
}

bool ewald_build_trad(EwaldFit& e, int D, const double* a, double rcut, double kcut, double kappa) {
  e.D = D;
  e.trad = true;
  e.kappa = kappa;
  e.rcut = rcut;
  e.kcut = kcut;
  ewald_kvectors(e, D, a, kcut);
  // TradEwaldSum::evalFK, the 3D transform in 2D too, as the reference has it
  for (double& w : e.vk) {
    const double k = std::sqrt(w);
    w = 4 * kPi * std::exp(-(k * k) / (4 * kappa * kappa)) / (k * k);
    if (!std::isfinite(w)) return false;
  }
  return std::isfinite(kappa) && kappa > 0;
}

bool ewald_build(EwaldFit& e, int D, const double* a, double rcut, double kcut) {
  e.D = D;
  e.trad = false;
  e.rcut = rcut;
  e.kcut = kcut;
  // --- the constrained fit over the halo kcut < k < 4 kcut (OptEwaldSum.cc:33-173)
  const double khalo = 4 * kcut, kcut2 = kcut * kcut, khalo2 = khalo * khalo;
  double deltak[3] = {0, 0, 0}, dk2[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) {
    deltak[i] = 2 * kPi / a[i];
    dk2[i] = deltak[i] * deltak[i];
  }
  double A[NP * NP] = {}, b[NP] = {};
  auto accumulate = [&](double k2, int degen) {
    if (!(k2 > kcut2)) return;
    const double k = std::sqrt(k2);
    double v, fm1, ft[NP];
    if (D == 3) {
      series3d(k, rcut, fm1, ft);
      v = -4. * kPi / (k * k) + fm1;
    } else {
      v = -2. * kPi / k;
      if (k * rcut > 30.) {
        v = 0;
        for (int p = 0; p < NP; ++p) ft[p] = 0.;
      } else {
        series2d(k, rcut, fm1, ft);
        v += fm1;
      }
    }
    for (int p = 0; p < NP; ++p) {
      b[p] += degen * v * ft[p];
      for (int q = 0; q < NP; ++q) A[p * NP + q] += degen * ft[p] * ft[q];
    }
  };
  for (int i = 0; i < khalo / deltak[0]; ++i) {
    const double kx2 = i * i * dk2[0];
    const int gx = (i == 0) ? 1 : 2;
    for (int j = 0; j < std::sqrt(khalo2 - kx2) / deltak[1]; ++j) {
      const double kxy2 = kx2 + j * j * dk2[1];
      const int gxy = gx * ((j == 0) ? 1 : 2);
      if (D == 2) {
        accumulate(kxy2, gxy);
        continue;
      }
      for (int k = 0; k < std::sqrt(khalo2 - kxy2) / deltak[2]; ++k)
        accumulate(kxy2 + k * k * dk2[2], gxy * ((k == 0) ? 1 : 2));
    }
  }
  // LAPACK reads the C-ordered normal matrix as its transpose
  double At[NP * NP];
  for (int p = 0; p < NP; ++p)
    for (int q = 0; q < NP; ++q) At[p * NP + q] = A[q * NP + p];
  int piv[NP];
  lu(At, NP, piv);
  double c[NP], d[NP], t0[NP], tc[NP], td[NP], t[NP];
  for (int i = 0; i < NP; ++i) {
    c[i] = std::pow(rcut, 2 * i);
    d[i] = (2 * i) * std::pow(rcut, 2 * i - 1);
  }
  std::memcpy(t0, b, sizeof t0);
  lu_solve(At, NP, piv, t0);
  const double f0 = frcut(rcut, t0, 0), df0 = frcut(rcut, t0, 1);
  // Lagrange multipliers of f(rcut) = 1/rcut and f'(rcut) = -1/rcut^2 by
  // linear extrapolation from two guesses
  double lamc = 10 * rcut, lamd = -rcut * rcut;
  for (int i = 0; i < NP; ++i) {
    tc[i] = b[i] + lamc * c[i];
    td[i] = b[i] + lamd * d[i];
  }
  lu_solve(At, NP, piv, tc);
  lu_solve(At, NP, piv, td);
  const double m00 = (frcut(rcut, tc, 0) - f0) / lamc, m01 = (frcut(rcut, td, 0) - f0) / lamd;
  const double m10 = (frcut(rcut, tc, 1) - df0) / lamc, m11 = (frcut(rcut, td, 1) - df0) / lamd;
  double rhs[2] = {1. / rcut - f0, -1. / (rcut * rcut) - df0};
  double mt[4] = {m00, m10, m01, m11};   // transposed, as LAPACK reads it
  int piv2[2];
  lu(mt, 2, piv2);
  lu_solve(mt, 2, piv2, rhs);
  lamc = rhs[0];
  lamd = rhs[1];
  for (int i = 0; i < NP; ++i) t[i] = b[i] + lamc * c[i] + lamd * d[i];
  lu_solve(At, NP, piv, t);
  for (int i = 0; i < NP; ++i) {
    if (!std::isfinite(t[i])) return false;
    e.coef[i] = t[i];
  }
  ewald_kvectors(e, D, a, kcut);
  // setLongRangeArray: vk = evalFK(|k|) (OptEwaldSum.cc:192-250)
  for (double& w : e.vk) {
    const double k = std::sqrt(w);
    double v;
    if (D == 3) {
      v = 4. * kPi / (k * k);
      const double cr = std::cos(k * rcut), sr = std::sin(k * rcut);
      double re = sr / k, im = -(cr - 1.) / k;
      const double fpioverk = 4. * kPi / k;
      v -= fpioverk * im;
      double rn = rcut;
      for (int n = 0; n < 2 * NP - 1; ++n) {
        const double nr = rn * cr - (n + 1.) * re, ni = rn * sr - (n + 1.) * im;
        re = ni / k;
        im = -nr / k;
        rn *= rcut;
        if (n % 2 == 0) v += e.coef[n / 2] * im * fpioverk;
      }
    } else {
      v = 2. * kPi / k;
      if (k * rcut < 30.) {
        double fm1, ft[NP];
        series2d(k, rcut, fm1, ft);
        v -= fm1;
        for (int i = 0; i < NP; ++i) v += e.coef[i] * ft[i];
      }
    }
    if (!std::isfinite(v)) return false;
    w = v;
  }
  return true;
}

}  // namespace pq
