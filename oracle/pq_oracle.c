/*
 * pq_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C FP64 restatement of pi-qmc's multilevel-bisection hot path.  It
 * follows the reference loop order, uniform-draw order and level-dependent
 * approximations so that, given the same uniform stream, it makes the same
 * accept/reject decisions.  Each function names the reference lines it
 * restates (paths relative to the pi-qmc source tree).
 *
 * Deliberate differences from the reference, none of which changes a result:
 *  - SHONodes/FreeParticleNodes run the Hungarian assignment only to fill
 *    kindex, whose sole reader (fgrad) is discarded when useHungarian is off
 *    (SHONodes.cc:77-81,145-147; FreeParticleNodes.cc:105-112,203).  It is
 *    skipped here.
 *  - Node distances are stored for the node model's own species only; the
 *    reference stores 200 for the other particles (SHONodes.cc:132), whose
 *    nodal-action term is log((1-e^-40000)/(1-e^-40000)) = 0 exactly.
 *  - LAPACK dgetrf/dgetri are replaced by a textbook unblocked LU with
 *    partial pivoting (first maximal |pivot|, as idamax) and an inverse from
 *    the LU factors; the determinant is the product of U's diagonal with the
 *    pivot sign, as SHONodes.cc:97-101.
 *  - The rescale loop of evaluate(...,scaleMagnitude=true) gives up after 200
 *    passes and reports an error instead of spinning (SHONodes.cc:55-126).
 *  - The CompositeAction sums spring, then SHO actions, then Coulomb.
 *  - OptEwaldSum's dgetrf/dgetrs/dgesv solves of its 8 x 8 and 2 x 2 fit
 *    systems use the same textbook LU (on the transposed matrices LAPACK
 *    reads); the fit meets its two constraints to a few parts in 1e9 and
 *    reproduces the reference's Madelung test, but LAPACK's own operation
 *    order, and so the last digits of the coefficients, is not pinned.
 */
#include "pq_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 256

/* ------------------------------------------------------------------ RNG */

typedef struct {
  int mode;               /* 0 glibc, 1 philox, 2 injected */
  int32_t st[31];
  int f, r;
  uint64_t seed;
  uint32_t stream;
  const double* inj;
  int64_t ninj;
  uint64_t count;
} oq_rng;

/* glibc srandom_r/random_r TYPE_3 (degree 31, separation 3), the generator
   behind rand() in RandomNumGenerator::getRand (src/util/RandomNumGenerator.h:42). */
static void glibc_seed(oq_rng* g, unsigned int seed) {
  if (seed == 0) seed = 1;
  g->st[0] = (int32_t)seed;
  int32_t word = (int32_t)seed;
  for (int i = 1; i < 31; ++i) {
    long hi = word / 127773;
    long lo = word % 127773;
    word = (int32_t)(16807 * lo - 2836 * hi);
    if (word < 0) word += 2147483647;
    g->st[i] = word;
  }
  g->f = 3;
  g->r = 0;
  for (int k = 0; k < 310; ++k) {
    uint32_t v = (uint32_t)g->st[g->f] + (uint32_t)g->st[g->r];
    g->st[g->f] = (int32_t)v;
    if (++g->f >= 31) { g->f = 0; ++g->r; }
    else if (++g->r >= 31) g->r = 0;
  }
}

static int32_t glibc_next(oq_rng* g) {
  uint32_t v = (uint32_t)g->st[g->f] + (uint32_t)g->st[g->r];
  g->st[g->f] = (int32_t)v;
  int32_t res = (int32_t)(v >> 1);
  if (++g->f >= 31) { g->f = 0; ++g->r; }
  else if (++g->r >= 31) g->r = 0;
  return res;
}

void oq_glibc_sequence(unsigned int seed, int n, int32_t* out) {
  oq_rng g;
  glibc_seed(&g, seed);
  for (int i = 0; i < n; ++i) out[i] = glibc_next(&g);
}

/* Philox4x32-10 (Salmon et al., SC'11).  Draw k of stream s under seed:
   counter (k>>2, k>>34, s, 0), key (seed lo, seed hi), word k&3, u = w*2^-32. */
static void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int rnd = 0; rnd < 10; ++rnd) {
    if (rnd > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

void oq_philox_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = ctr[i];
  philox4x32_10(out, key[0], key[1]);
}

uint32_t oq_philox_u32(uint64_t seed, uint32_t stream, uint64_t k) {
  uint32_t c[4] = {(uint32_t)(k >> 2), (uint32_t)(k >> 34), stream, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return c[k & 3];
}

double oq_philox_uniform(uint64_t seed, uint32_t stream, uint64_t k) {
  return (double)oq_philox_u32(seed, stream, k) * 2.3283064365386963e-10;
}

static double rng_next(oq_rng* g) {
  double u;
  if (g->mode == 0) {
    /* rand()*SCALE_RAND, SCALE_RAND = 1./RAND_MAX (RandomNumGenerator.cc:3) */
    const double scale = 1. / 2147483647;
    u = glibc_next(g) * scale;
  } else if (g->mode == 1) {
    u = oq_philox_uniform(g->seed, g->stream, g->count);
  } else {
    /* past the end of an injected stream: a placeholder draw, and the
       stream reports itself exhausted (oq_rng_exhausted) */
    u = (g->count < (uint64_t)g->ninj) ? g->inj[g->count] : 0.5;
  }
  g->count++;
  return u;
}

/* RandomNumGenerator::makeGaussRand (src/util/RandomNumGenerator.h:75-86) */
static void gauss_rand(oq_rng* g, double* a, int n) {
  for (int i = 0; i + 1 < n; i += 2) {
    double t1 = 1 - 0.9999999999 * rng_next(g);
    double t2 = rng_next(g);
    a[i] = sqrt(-2.0 * log(t1)) * cos(6.283185306 * t2);
    a[i + 1] = sqrt(-2.0 * log(t1)) * sin(6.283185306 * t2);
  }
  if (n % 2 == 1) {
    double t1 = 1 - 0.9999999999 * rng_next(g);
    double t2 = rng_next(g);
    a[n - 1] = sqrt(-2.0 * log(t1)) * cos(6.283185306 * t2);
  }
}

/* -------------------------------------------------- SuperCell, PG */

typedef struct { double a[3], b[3], rcut2; int D; } cell_t;

/* SuperCell::computeRecipricalVectors / pbc (src/util/SuperCell.cc:14-31) */
static void cell_init(cell_t* c, const double* a, int D) {
  c->D = D;
  for (int d = 0; d < 3; ++d) { c->a[d] = a[d]; c->b[d] = (d < D) ? 1. / a[d] : 0; }
  double l2 = a[0] * a[0];
  c->rcut2 = l2;
  for (int i = 1; i < D; ++i) { l2 = a[i] * a[i]; c->rcut2 = (c->rcut2 > l2) ? l2 : c->rcut2; }
  c->rcut2 *= 0.25;
}

static void pbc(const cell_t* c, double* v) {
  double r2 = 0;
  for (int d = 0; d < c->D; ++d) r2 += v[d] * v[d];
  if (r2 > c->rcut2) {
    for (int d = 0; d < c->D; ++d) {
      double i = floor(c->b[d] * v[d] + 1000.5) - 1000;
      v[d] -= c->a[d] * i;
    }
  }
}

void oq_pbc(const double cell[3], int ndim, double* v) {
  cell_t c;
  cell_init(&c, cell, ndim);
  pbc(&c, v);
}

/* PeriodicGaussian (src/util/PeriodicGaussian.cc:7-40) */
typedef struct { double nome, prefactor, k; int nmax; } pg_t;
static const double PG_PI = 3.141592653589793;

int oq_pg_nterms(double alpha, double length) {
  return 1 + (int)(2.0 * sqrt(alpha * length * length));
}

static void pg_init(pg_t* p, double alpha, double length) {
  p->nome = exp(-PG_PI * PG_PI / (alpha * length * length));
  p->nmax = oq_pg_nterms(alpha, length);
  p->prefactor = sqrt(PG_PI / (alpha * length * length));
  p->k = 2 * PG_PI / length;
}

/* value, gradient, second derivative; complex arithmetic written out */
static double pg_eval3(const pg_t* p, double x, double* grad, double* d2) {
  double wr = cos(p->k * x), wi = sin(p->k * x);
  double s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0;
  double tr = 1.0, ti = 0.0;
  double k2 = -p->k * p->k;
  double nomeN = 1.0;
  for (int n = 1; n < p->nmax; ++n) {
    double f = nomeN * nomeN * p->nome;
    /* term *= f*w  (std::complex: (f*w) computed first, then term*(f*w)) */
    double fwr = f * wr, fwi = f * wi;
    double nr = tr * fwr - ti * fwi, ni = tr * fwi + ti * fwr;
    tr = nr; ti = ni;
    s0r += tr; s0i += ti;
    /* term * ik * n : (term*ik) then *n */
    double ar = -ti * p->k, ai = tr * p->k;
    s1r += ar * (double)n; s1i += ai * (double)n;
    s2r += tr * k2 * (double)(n * n); s2i += ti * k2 * (double)(n * n);
    nomeN *= p->nome;
  }
  (void)s0i; (void)s1i; (void)s2i;
  if (grad) *grad = p->prefactor * 2 * s1r;
  if (d2) *d2 = p->prefactor * 2 * s2r;
  return p->prefactor * (1 + 2 * s0r);
}

static double pg_eval(const pg_t* p, double x) { return pg_eval3(p, x, 0, 0); }

void oq_pg_eval(double alpha, double length, double x, double out[3]) {
  pg_t p;
  pg_init(&p, alpha, length);
  out[0] = pg_eval3(&p, x, &out[1], &out[2]);
}

/* ------------------------------------------------ Coulomb link action */

/* Coulomb1DLinkAction (src/action/coulomb/Coulomb1DLinkAction.cpp:4-133) */
typedef struct { double stau, m1, m2, m3, m4, m5; } c1d_t;

static void c1d_init(c1d_t* c, double stau) {
  c->stau = stau;
  c->m1 = 1.0 / stau;
  c->m2 = c->m1 * c->m1;
  c->m3 = c->m2 * c->m1;
  c->m4 = c->m3 * c->m1;
  c->m5 = c->m4 * c->m1;
}

static double c1d_origin(const c1d_t* c) {
  double s = c->stau;
  return s * (1.772453851 + s * (-0.074137740 + s * (0.005834805 + s * (-0.000382686
         + s * (0.000008738 + s * 0.000002138)))));
}

static double c1d_u0(const c1d_t* c, double u0, double reff) {
  double a = 0.25300593 * c->m1 + 0.01432126;
  double b = 0.07936898 * c->m2 - 0.01634421 / c->stau;
  double cc = 0.07263383 * c->m3;
  double d = 0.00940013 * c->m4;
  double e = 0.03160181 * c->m5;
  double f = 0.18976335 * c->m1;
  double g = 0.00343053 * c->m2;
  double v = u0 + reff * ((u0 * a - 1.) + reff * (f + reff * (g + reff * e * c->stau * c->stau)));
  v /= 1. + reff * (a + reff * (b + reff * (cc + reff * (d + reff * e))));
  return v;
}

static double c1d_u1(const c1d_t* c, double reff) {
  double a = 0.07776906 * c->m1, b = -0.04695226 * c->m2, cc = 0.01291994 * c->m3;
  double d = 0.60942687 * c->m1, e = -1.09588064 * c->m2, f = 1.15591565 * c->m3;
  double g = -0.57485278 * c->m4, h = 0.15212765 * c->m5;
  double v = reff * reff * (a + reff * (b + reff * cc));
  v /= (1 + reff * (d + reff * (e + reff * (f + reff * (g + reff * h)))));
  return v;
}

static double c1d_u2(const c1d_t* c, double reff) {
  double a = 9.52130652e-04 * c->m3, b = -6.39946071e-04 * c->m4, cc = 1.14359558e-04 * c->m5;
  double d = -7.59575215e-01 * c->m3, e = 5.53798980e-01 * c->m4, f = -7.15626226e-02 * c->m5;
  double g = -3.49081618e-02 * c->m5 * c->m1, h = 8.40084881e-03 * c->m5 * c->m2;
  return pow(reff, 4) * (a + reff * (b + reff * cc))
         / (1 + pow(reff, 3) * (d + reff * (e + reff * (f + reff * (g + reff * h)))));
}

static double c1d_u3(const c1d_t* c, double reff) {
  double a = -2.06984256e-1 * c->m5, b = -9.51423947e-3 * c->m5 * c->m1;
  double cc = 8.97532561e-2 * c->m5 * c->m2, d = 6.10876797e+4 * c->m3;
  double e = -1.19016292e+5 * c->m4, f = 9.29836599e+4 * c->m5;
  double g = -3.49919130e+4 * c->m5 * c->m1, h = 6.42490539e+3 * c->m5 * c->m2;
  double i = -6.02729064e+2 * c->m5 * c->m3, j = 6.01285550e+1 * c->m5 * c->m4;
  return pow(reff, 6) * (a + reff * (b + reff * cc))
         / (1 + pow(reff, 3) * (d + reff * (e + reff * (f + reff * (g + reff * (h + reff * (i + reff * j)))))));
}

static double c1d_u4(const c1d_t* c, double reff) {
  double a = -2.84102296e-6 * c->m5, b = 6.26961672e-7 * c->m5 * c->m2;
  double cc = -2.20166473e-1 * c->m3, d = 1.07981903e-1 * c->m5;
  double e = -1.94671111e-2 * c->m5 * c->m2, f = 1.56217930e-3 * c->m5 * c->m4;
  return pow(reff, 6) * (a + b * reff * reff)
         / (1 + reff * reff * reff * (cc + reff * reff * (d + reff * reff * (e + reff * reff * f))));
}

double oq_coulomb1d(double stau, int which, double reff) {
  c1d_t c;
  c1d_init(&c, stau);
  switch (which) {
    case -1: return c1d_origin(&c);
    case 0: return c1d_u0(&c, c1d_origin(&c), reff);
    case 1: return c1d_u1(&c, reff);
    case 2: return c1d_u2(&c, reff);
    case 3: return c1d_u3(&c, reff);
    default: return c1d_u4(&c, reff);
  }
}

/* Coulomb3DLinkAction (src/action/coulomb/Coulomb3DLinkAction.cpp:13-90) */
static double c3d(const c1d_t* c, int which, double reff) {
  double stau = c->stau;
  if (which == 0) {
    double u0 = c1d_origin(c);
    double u1_0 = c1d_u0(c, u0, reff);
    double u1_1 = c1d_u1(c, reff);
    double a = 0.5 * reff * reff / (stau * stau);
    double b = exp(-2. * a);
    double cc = 1.0 / (1.0 + 2 * (1.0 - b) * u1_1 / (a + 1e-200));
    return u1_0 + log(cc);
  } else if (which == 1) {
    double u1_1 = c1d_u1(c, reff), u1_2 = c1d_u2(c, reff);
    double a = 0.5 * reff * reff / (stau * stau);
    double b = exp(-2. * a);
    double cc = 1.0 / (1.0 + 2 * (1.0 - b) * u1_1 / (a + 1e-200));
    return u1_1 + cc * (b * u1_1 - 4 * (1 - b) * u1_2 / (a + 1e-200));
  } else if (which == 2) {
    double u1_1 = c1d_u1(c, reff), u1_2 = c1d_u2(c, reff), u1_3 = c1d_u3(c, reff);
    double a = 2. * stau * stau / (reff * reff + 1e-50);
    double b = exp(-2. / (a + 1e-200));
    double cc = 1. / (1 + 2 * a * (1 - b) * u1_1);
    return u1_2 + 0.25 * cc / (a + 1e-200)
           * (cc * (b * u1_1 * (1 + 2 * a * u1_1) + 8 * a * b * u1_2
                    + 32 * a * a * a * (1 - b) * (1 - b) * u1_2 * u1_2)
              - 24 * a * a * (1 - b) * u1_3);
  } else if (which == 3) {
    double u1_1 = c1d_u1(c, reff), u1_2 = c1d_u2(c, reff);
    double u1_3 = c1d_u3(c, reff), u1_4 = c1d_u4(c, reff);
    double a = 2. * stau * stau / (reff * reff + 1e-50);
    double b = exp(-2. / (a + 1e-200));
    double cc = 1. / (1 + 2 * a * (1 - b) * u1_1);
    return u1_3 + cc * cc * cc / (24. * a * a + 1e-200)
      * (128 * a * a * a * a * a * b * b * b * (4 * u1_2 * u1_2 * u1_2
      - 9 * u1_1 * u1_2 * u1_3 + 6 * u1_1 * u1_1 * u1_4)
      + 64 * a * a * a * (a * u1_2 * (
      -8 * a * u1_2 * u1_2 + 9 * (1 + 2 * a * u1_1) * u1_3)
      - 3 * (1 + 2 * a * u1_1) * (1 + 2 * a * u1_1)
      * u1_4) + 2 * a * b * b
      * (2 * a * u1_1 * u1_1 * u1_1
      + 96 * a * a * u1_2 * (u1_2 - 8 * a * a * u1_2 * u1_2
      + 3 * a * u1_3) + 12 * a * u1_1 * (u1_2 - 6 * a * u1_3
      + 144 * a * a * a * u1_2 * u1_3
      - 32 * a * a * u1_4) + u1_1 * u1_1
      * (1 - 1152 * a * a * a * a * u1_4))
      + b * (4 * a * a * u1_1 * u1_1 * u1_1
      + 12 * a * (u1_2 - 16 * a * a * u1_2 * u1_2
      + 128 * a * a * a * a * u1_2 * u1_2 * u1_2 + 6 * a * u1_3 - 96 * a * a * a
      * u1_2 * u1_3 + 16 * a * a * u1_4) + 4 * u1_1 * u1_1
      * (a + 576 * a * a * a * a * a * u1_4)
      + u1_1 * (1 + 24 * a * a * (u1_2 + 6 * a * u1_3
              - 144 * a * a * a * u1_2 * u1_3 + 64 * a * a * u1_4))));
  }
  return c1d_u4(c, reff);
}

double oq_coulomb3d(double stau, int which, double reff) {
  c1d_t c;
  c1d_init(&c, stau);
  return c3d(&c, which, reff);
}

/* ------------------------------------------------------------ Ewald sum */

/* OptEwaldSum (src/util/OptEwaldSum.cc:27-280) on the k-vectors of
   EwaldSum (src/util/EwaldSum.cc:16-112): f(r) = 1/r beyond rcut and an even
   polynomial of npoly = 8 terms inside, fitted over the k-space halo
   kcut < k < 4 kcut with value and slope pinned at rcut. */
#define EW_NPOLY 8
static const double EW_PI = 3.14159265358979;   /* EwaldSum::PI (EwaldSum.cc:187) */

typedef struct {
  int D, N, nk;
  int trad;              /* TradEwaldSum (TradEwaldSum.cc:13-36) */
  double kappa;
  double rcut, kcut;
  double coef[EW_NPOLY];
  int ikmax[3];
  double dk[3];          /* 2 PI b_d */
  int* kv;               /* [nk][3] half-space integer k-vectors */
  double* vk;            /* [nk] evalFK(|k|) */
  double oneOver2V, self;
  double q[MAXN];
} ewald_t;

/* the erf the reference links (src/util/erf.cpp): the incomplete gamma
   function P(1/2, x^2) by its series below x^2 = 3/2 and by the Laguerre
   continued fraction of Q = 1 - P above (H. Okumura, 1991, p. 227) */
static double gamma_q(double a, double x, double lga);
static double gamma_p(double a, double x, double lga) {
  if (x >= 1 + a) return 1 - gamma_q(a, x, lga);
  if (x == 0) return 0;
  double term = exp(a * log(x) - x - lga) / a, sum = term;
  for (int k = 1; k < 1000; k++) {
    term *= x / (a + k);
    double prev = sum;
    sum += term;
    if (sum == prev) break;
  }
  return sum;
}
static double gamma_q(double a, double x, double lga) {
  if (x < 1 + a) return 1 - gamma_p(a, x, lga);
  double la = 1, lb = 1 + x - a, w = exp(a * log(x) - x - lga), sum = w / lb;
  for (int k = 2; k < 1000; k++) {
    double t = ((k - 1 - a) * (lb - la) + (k + x) * lb) / k;
    la = lb; lb = t;
    w *= (k - 1 - a) / k;
    double prev = sum;
    sum += w / (la * lb);
    if (sum == prev) break;
  }
  return sum;
}
static double ref_erf(double x) {
  const double lgh = 0.572364942924700087071713675675;
  if (!isfinite(x)) return x > 0 ? 1.0 : -1.0;
  return x >= 0 ? gamma_p(0.5, x * x, lgh) : -gamma_p(0.5, x * x, lgh);
}
double oq_ref_erf(double x) { return ref_erf(x); }

/* evalFR (OptEwaldSum.cc:178-190; TradEwaldSum.cc:24-26) */
static double ew_fr(const ewald_t* e, double r) {
  double v = 0.;
  if (e->trad) return ref_erf(e->kappa * r) / r;
  if (r > e->rcut) return 1. / r;
  double r2 = r * r;
  for (int i = EW_NPOLY - 1; i >= 0; --i) { v *= r2; v += e->coef[i]; }
  return v;
}

/* the 2D even-power transforms of OptEwaldSum.cc:92-118 / 206-232: the
   Bessel series of the n = -1 term (*fm1) and of r^n for even n < 2 npoly */
static void ew_series2d(double k, double rcut, double* fm1, double* fn_even) {
  double factor = -0.25 * k * k * rcut * rcut;
  double term = 2. * EW_PI;
  double fn = term * rcut;
  int m = 1;
  while (fabs(term) > 1e-100) { term *= factor / (m * m); fn += term * rcut / (2 * m + 1); ++m; }
  *fm1 = fn;
  double rnplus2 = rcut * rcut;
  for (int n = 0; n < 2 * EW_NPOLY; n += 2) {
    double t = 2. * EW_PI;
    double f = t * rnplus2 / (n + 2);
    int mm = 1;
    while (fabs(t) > 1e-100) { t *= factor / (mm * mm); f += t * rnplus2 / (2 * mm + n + 2); ++mm; }
    fn_even[n / 2] = f;
    rnplus2 *= rcut * rcut;
  }
}

/* the 3D recursion of OptEwaldSum.cc:76-89 / 193-205: temp_n = int_0^rcut
   r^n e^{ikr} dr by parts; ft[n/2] = 4 PI Im(temp_n)/k for even n, and the
   n = -1 term 4 PI Im((e^{ik rcut} - 1)/(ik))/k in *fm1 */
static void ew_series3d(double k, double rcut, double* fm1, double* ft) {
  double er = cos(k * rcut), ei = sin(k * rcut);
  /* temp = (eikr - 1)/(ik) = (er - 1 + i ei)(-i/k) */
  double tr = ei / k, ti = -(er - 1.) / k;
  *fm1 = 4. * EW_PI * ti / k;
  double rn = rcut;
  for (int n = 0; n < 2 * EW_NPOLY - 1; ++n) {
    /* temp = (rn eikr - (n+1) temp)/(ik) */
    double nr = rn * er - (n + 1.) * tr, ni = rn * ei - (n + 1.) * ti;
    tr = ni / k; ti = -nr / k;
    rn *= rcut;
    if (n % 2 == 0) ft[n / 2] = 4. * EW_PI * ti / k;
  }
}

/* evalFK (OptEwaldSum.cc:192-250), in its own operation order: (4 PI/k) Im */
static double ew_fk(const ewald_t* e, double k) {
  double fm1, f[EW_NPOLY];
  /* TradEwaldSum::evalFK, the 3D transform in 2D too (TradEwaldSum.cc:32-34) */
  if (e->trad) return 4 * EW_PI * exp(-(k * k) / (4 * e->kappa * e->kappa)) / (k * k);
  if (e->D == 3) {
    const double rcut = e->rcut;
    double v = 4. * EW_PI / (k * k);
    double er = cos(k * rcut), ei = sin(k * rcut);
    double tr = ei / k, ti = -(er - 1.) / k;
    const double fpioverk = 4. * EW_PI / k;
    v -= fpioverk * ti;
    double rn = rcut;
    for (int n = 0; n < 2 * EW_NPOLY - 1; ++n) {
      double nr = rn * er - (n + 1.) * tr, ni = rn * ei - (n + 1.) * ti;
      tr = ni / k; ti = -nr / k;
      rn *= rcut;
      if (n % 2 == 0) v += e->coef[n / 2] * ti * fpioverk;
    }
    return v;
  }
  double v = 2. * EW_PI / k;
  if (k * e->rcut < 30.) {
    ew_series2d(k, e->rcut, &fm1, f);
    v -= fm1;
    for (int i = 0; i < EW_NPOLY; ++i) v += e->coef[i] * f[i];
  }
  return v;
}

/* evalFK0 (OptEwaldSum.cc:252-270) */
static double ew_fk0(const ewald_t* e) {
  const double rc = e->rcut;
  if (e->trad) return -EW_PI / (e->kappa * e->kappa);
  if (e->D == 3) {
    double fn = 2. * EW_PI * rc * rc, v = -fn;
    fn *= rc / 3.;
    for (int n = 0; n < 2 * EW_NPOLY; n += 2) { v += e->coef[n / 2] * fn; fn *= (rc * rc) * (n + 3.) / (n + 5.); }
    return v;
  }
  double fn = 2. * EW_PI * rc, v = -fn;
  fn *= rc / 2.;
  for (int n = 0; n < 2 * EW_NPOLY; n += 2) { v += e->coef[n / 2] * fn; fn *= (rc * rc) * (2 + n) / (4 + n); }
  return v;
}

/* evalFRcut (OptEwaldSum.cc:272-280) */
static double ew_frcut(const ewald_t* e, const double* t, int nd) {
  double v = 0;
  for (int i = 0; i < EW_NPOLY; ++i) {
    double x = pow(e->rcut, 2 * i - nd);
    for (int j = 0; j < nd; ++j) x *= 2 * i - j;
    v += t[i] * x;
  }
  return v;
}

/* LU with partial pivoting of a row-major n x n matrix in place (dgetrf) and
   the solve of A x = b with it (dgetrs 'N') */
static void ew_lu(double* a, int n, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i) if (fabs(a[i * n + k]) > fabs(a[p * n + k])) p = i;
    piv[k] = p;
    if (p != k) for (int j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
    if (a[k * n + k] == 0) continue;
    double r = 1.0 / a[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = a[i * n + k] * r;
      a[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
  }
}
static void ew_lu_solve(const double* a, int n, const int* piv, double* b) {
  for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
  for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) b[i] -= a[i * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= a[i * n + j] * b[j];
    b[i] /= a[i * n + i];
  }
}

/* OptEwaldSum constructor: the constrained least-squares fit
   (OptEwaldSum.cc:27-173).  The reference hands C-ordered blitz arrays to the
   column-major LAPACK routines, so they factor and solve the transposes of
   a and mat; so does this. */
static void ew_fit(ewald_t* e, const double* a_cell) {
  const int D = e->D, np = EW_NPOLY;
  const double kcut = e->kcut, khalo = 4 * e->kcut, rcut = e->rcut;
  const double kcut2 = kcut * kcut, khalo2 = khalo * khalo;
  double deltak[3], dk2[3];
  for (int i = 0; i < D; ++i) { deltak[i] = 2 * EW_PI / a_cell[i]; dk2[i] = deltak[i] * deltak[i]; }
  double A[EW_NPOLY * EW_NPOLY] = {0}, b[EW_NPOLY] = {0};
  for (int i = 0; i < khalo / deltak[0]; ++i) {
    double kx2 = i * i * dk2[0];
    int idx = (i == 0) ? 1 : 2;
    for (int j = 0; j < sqrt(khalo2 - kx2) / deltak[1]; ++j) {
      double kx2ky2 = kx2 + j * j * dk2[1];
      int idxy = idx * ((j == 0) ? 1 : 2);
      /* 2D: one pass with k2 = kx2ky2; 3D: the kz column of the octant */
      const int kzn = (D == 3) ? -1 : 1;
      for (int k = 0; (D == 3) ? (k < sqrt(khalo2 - kx2ky2) / deltak[2]) : (k < kzn); ++k) {
        double k2 = (D == 3) ? kx2ky2 + k * k * dk2[2] : kx2ky2;
        int idegen = (D == 3) ? idxy * ((k == 0) ? 1 : 2) : idxy;
        if (!(k2 > kcut2)) continue;
        double kk = sqrt(k2), v, fm1, ft[EW_NPOLY];
        if (D == 3) {
          ew_series3d(kk, rcut, &fm1, ft);
          v = -4. * EW_PI / (kk * kk) + fm1;
        } else {
          v = -2. * EW_PI / kk;
          if (kk * rcut > 30.) {
            v = 0;
            for (int p = 0; p < np; ++p) ft[p] = 0.;
          } else {
            ew_series2d(kk, rcut, &fm1, ft);
            v += fm1;
          }
        }
        for (int p = 0; p < np; ++p) {
          b[p] += idegen * v * ft[p];
          for (int q = 0; q < np; ++q) A[p * np + q] += idegen * ft[p] * ft[q];
        }
      }
    }
  }
  /* the matrix LAPACK sees: A^T */
  double At[EW_NPOLY * EW_NPOLY];
  for (int p = 0; p < np; ++p) for (int q = 0; q < np; ++q) At[p * np + q] = A[q * np + p];
  int piv[EW_NPOLY];
  ew_lu(At, np, piv);
  double c[EW_NPOLY], d[EW_NPOLY], t0[EW_NPOLY], tc[EW_NPOLY], td[EW_NPOLY], t[EW_NPOLY];
  for (int i = 0; i < np; ++i) { c[i] = pow(rcut, 2 * i); d[i] = (2 * i) * pow(rcut, 2 * i - 1); }
  memcpy(t0, b, sizeof t0);
  ew_lu_solve(At, np, piv, t0);
  double f0 = ew_frcut(e, t0, 0), df0 = ew_frcut(e, t0, 1);
  double lamc = 10 * rcut, lamd = -rcut * rcut;
  for (int i = 0; i < np; ++i) { tc[i] = b[i] + lamc * c[i]; td[i] = b[i] + lamd * d[i]; }
  ew_lu_solve(At, np, piv, tc);
  ew_lu_solve(At, np, piv, td);
  double mat[4], rhs[2];
  mat[0] = (ew_frcut(e, tc, 0) - f0) / lamc;   /* mat(0,0) */
  mat[1] = (ew_frcut(e, td, 0) - f0) / lamd;   /* mat(0,1) */
  mat[2] = (ew_frcut(e, tc, 1) - df0) / lamc;  /* mat(1,0) */
  mat[3] = (ew_frcut(e, td, 1) - df0) / lamd;  /* mat(1,1) */
  rhs[0] = 1. / rcut - f0;
  rhs[1] = -1. / (rcut * rcut) - df0;
  double mt[4] = {mat[0], mat[2], mat[1], mat[3]};
  int piv2[2];
  ew_lu(mt, 2, piv2);
  ew_lu_solve(mt, 2, piv2, rhs);
  lamc = rhs[0]; lamd = rhs[1];
  for (int i = 0; i < np; ++i) t[i] = b[i] + lamc * c[i] + lamd * d[i];
  ew_lu_solve(At, np, piv, t);
  memcpy(e->coef, t, sizeof t);
}

/* EwaldSum constructor's k-vectors (EwaldSum.cc:16-88), setLongRangeArray and
   evalSelfEnergy (EwaldSum.cc:90-97,180-185) with the particles' charges, as
   CoulombAction's constructor leaves them (CoulombAction.cc:28-44) */
static void ewald_init(ewald_t* e, const cell_t* cell, int N, const double* q,
                       double rcut, double kcut, int trad, double kappa) {
  const int D = cell->D;
  e->D = D; e->N = N;
  e->trad = trad; e->kappa = kappa;
  e->rcut = rcut; e->kcut = kcut;
  for (int i = 0; i < 3; ++i) { e->ikmax[i] = 0; e->dk[i] = 0; }
  for (int i = 0; i < D; ++i) {
    e->ikmax[i] = (int)(cell->a[i] * kcut / (2 * EW_PI));
    e->dk[i] = 2 * EW_PI * cell->b[i];
  }
  double vol = 1;
  for (int i = 0; i < D; ++i) vol *= cell->a[i];
  e->oneOver2V = 0.5 / vol;
  if (!trad) ew_fit(e, cell->a);
  int cap = 64;
  e->kv = (int*)malloc(sizeof(int) * 3 * cap);
  e->nk = 0;
  const int kzmax = (D == 3) ? e->ikmax[2] : 0;
  for (int kx = 0; kx <= e->ikmax[0]; ++kx) {
    double kx2 = kx * kx * e->dk[0] * e->dk[0];
    for (int ky = ((kx == 0) ? 0 : -e->ikmax[1]); ky <= e->ikmax[1]; ++ky) {
      double ky2 = ky * ky * e->dk[1] * e->dk[1];
      int kzlo = (D == 3) ? ((kx == 0 && ky == 0) ? 0 : -kzmax) : 0;
      for (int kz = kzlo; kz <= kzmax; ++kz) {
        double k2 = (D == 3) ? kx2 + ky2 + kz * kz * e->dk[2] * e->dk[2] : kx2 + ky * ky * e->dk[1] * e->dk[1];
        if (k2 < kcut * kcut && k2 != 0) {
          if (e->nk == cap) { cap *= 2; e->kv = (int*)realloc(e->kv, sizeof(int) * 3 * cap); }
          e->kv[3 * e->nk] = kx; e->kv[3 * e->nk + 1] = ky; e->kv[3 * e->nk + 2] = kz;
          e->nk++;
        }
      }
    }
  }
  e->vk = (double*)malloc(sizeof(double) * (e->nk ? e->nk : 1));
  for (int k = 0; k < e->nk; ++k) {
    const int* v = &e->kv[3 * k];
    /* kvec2 as the constructor accumulates it */
    double k2 = v[0] * v[0] * e->dk[0] * e->dk[0] + v[1] * v[1] * e->dk[1] * e->dk[1];
    if (D == 3) k2 += v[2] * v[2] * e->dk[2] * e->dk[2];
    e->vk[k] = ew_fk(e, sqrt(k2));
  }
  double Q = 0, Q2 = 0;
  for (int i = 0; i < N; ++i) { e->q[i] = q[i]; Q += q[i]; Q2 += q[i] * q[i]; }
  const double fr0 = trad ? 2 * kappa / sqrt(EW_PI) : e->coef[0];   /* evalFR0 */
  e->self = -0.5 * Q2 * fr0 + e->oneOver2V * ew_fk0(e) * Q * Q;
}

static void ewald_free(ewald_t* e) { free(e->kv); free(e->vk); e->kv = 0; e->vk = 0; }

/* EwaldSum::evalLongRange (EwaldSum.cc:114-178): the k-space sum of the
   configuration r [N][D], with e^{i k r} built by the same recursions */
static double ewald_long_range(const ewald_t* e, const double* r) {
  const int D = e->D, N = e->N;
  const int nx = e->ikmax[0] + 1, ny = 2 * e->ikmax[1] + 1, nz = (D == 3) ? 2 * e->ikmax[2] + 1 : 1;
  double* ex = (double*)malloc(sizeof(double) * 2 * nx * N);
  double* ey = (double*)malloc(sizeof(double) * 2 * ny * N);
  double* ez = (double*)malloc(sizeof(double) * 2 * nz * N);
#define EX(k, i) (ex + 2 * ((size_t)(k) * N + (i)))
#define EY(k, i) (ey + 2 * ((size_t)((k) + e->ikmax[1]) * N + (i)))
#define EZ(k, i) (ez + 2 * ((size_t)((k) + e->ikmax[2]) * N + (i)))
  for (int i = 0; i < N; ++i) {
    EX(0, i)[0] = 1; EX(0, i)[1] = 0;
    EY(0, i)[0] = 1; EY(0, i)[1] = 0;
    if (D == 3) { EZ(0, i)[0] = 1; EZ(0, i)[1] = 0; }
  }
  if (e->ikmax[0] > 0)
    for (int i = 0; i < N; ++i) { double a = e->dk[0] * r[i * D]; EX(1, i)[0] = cos(a); EX(1, i)[1] = sin(a); }
  if (e->ikmax[1] > 0)
    for (int i = 0; i < N; ++i) {
      double a = e->dk[1] * r[i * D + 1];
      EY(1, i)[0] = cos(a); EY(1, i)[1] = sin(a);
      EY(-1, i)[0] = EY(1, i)[0]; EY(-1, i)[1] = -EY(1, i)[1];
    }
  if (D == 3 && e->ikmax[2] > 0)
    for (int i = 0; i < N; ++i) {
      double a = e->dk[2] * r[i * D + 2];
      EZ(1, i)[0] = cos(a); EZ(1, i)[1] = sin(a);
      EZ(-1, i)[0] = EZ(1, i)[0]; EZ(-1, i)[1] = -EZ(1, i)[1];
    }
#define CMUL(o, p, q) do { double _r = (p)[0] * (q)[0] - (p)[1] * (q)[1]; \
                           double _i = (p)[0] * (q)[1] + (p)[1] * (q)[0]; (o)[0] = _r; (o)[1] = _i; } while (0)
  for (int k = 2; k <= e->ikmax[0]; ++k)
    for (int i = 0; i < N; ++i) CMUL(EX(k, i), EX(k - 1, i), EX(1, i));
  for (int k = 2; k <= e->ikmax[1]; ++k)
    for (int i = 0; i < N; ++i) { CMUL(EY(k, i), EY(k - 1, i), EY(1, i)); CMUL(EY(-k, i), EY(-k + 1, i), EY(-1, i)); }
  if (D == 3)
    for (int k = 2; k <= e->ikmax[2]; ++k)
      for (int i = 0; i < N; ++i) { CMUL(EZ(k, i), EZ(k - 1, i), EZ(1, i)); CMUL(EZ(-k, i), EZ(-k + 1, i), EZ(-1, i)); }
  double sum = 0;
  for (int k = 0; k < e->nk; ++k) {
    const int* v = &e->kv[3 * k];
    double cs[2] = {0, 0};
    for (int j = 0; j < N; ++j) {
      /* q_j * eikx * eiky (* eikz), left to right */
      double a[2] = {e->q[j] * EX(v[0], j)[0], e->q[j] * EX(v[0], j)[1]}, t[2], u[2];
      CMUL(t, a, EY(v[1], j));
      if (D == 3) { CMUL(u, t, EZ(v[2], j)); } else { u[0] = t[0]; u[1] = t[1]; }
      cs[0] += u[0];
      cs[1] += u[1];
    }
    sum += 2 * e->vk[k] * (cs[0] * cs[0] + cs[1] * cs[1]);
  }
#undef CMUL
#undef EX
#undef EY
#undef EZ
  free(ex); free(ey); free(ez);
  return sum * e->oneOver2V + e->self;
}

/* CoulombAction::u (src/action/CoulombAction.cc:209-248), no screening
   shift; with an Ewald sum u_0 loses tau q1 q2 f(r) (:240-242) */
static double coulomb_u(double r, int order, double mu, double q1q2, double tau, const ewald_t* ew) {
  r = sqrt(r * r + 0.0);
  double reff = 2.0 * mu * q1q2 * r;
  double stau = q1q2 * sqrt(2.0 * mu * tau);
  c1d_t c;
  c1d_init(&c, stau);
  double u = c3d(&c, order, reff);
  if (order == 0 && ew) u -= tau * q1q2 * ew_fr(ew, r);
  return u;
}

/* CoulombAction::utau (src/action/CoulombAction.cc:250-259) */
static double coulomb_utau(double r, int order, double mu, double q1q2, double tau, const ewald_t* ew) {
  const double eps = 1e-7;
  double up = coulomb_u(r, order, mu, q1q2, tau * (1.0 + eps), ew);
  double um = coulomb_u(r, order, mu, q1q2, tau * (1.0 - eps), ew);
  return (up - um) / (2.0 * eps * tau);
}

/* ----------------------------------------------------------- the chain */

typedef struct {
  int s1, s2, norder, ngpts;
  double rgridinv, logrratioinv;
  double* u;   /* [ngpts][norder+1] */
  double* ut;  /* [ngpts][norder+1] */
} pair_t;

typedef struct {
  int sp, model, ifirst, ns, nodal;
  double mass, omega, T, coshwt, sinhwt, c, scale;
  pg_t pg[3];
  double *dm, *newdm, *dist, *newdist;  /* dist [S][2][ns] */
  double* inv;                           /* [S][ns*ns] latest inverse per slice */
  /* FreeParticleNodes::MatrixUpdate (useUpdates): the moving particles' new
     matrix columns phi [S][ns][PQ_MAX_MOVING], the updated inverses newinv
     [S][ns*ns], the moving particles of this species (index1 - ifirst and
     their positions mindex1 among the moving particles) and
     isNewMatrixUpdated */
  int upd;
  double *phi, *newinv;
  int nm, index1[PQ_MAX_MOVING], mindex1[PQ_MAX_MOVING];
  int new_updated;
  int not_my_species;
  long nerror;
} fn_t;

struct oq_chain {
  pq_chain_desc d;
  int N, D, M, L, S, nsp;
  int ifirst[PQ_MAX_SPECIES];
  int spec[MAXN];
  double lambda[MAXN];
  double tau, T;
  cell_t cell;
  /* periodic Gaussians of spring (level, species, dim), mover */
  int npg_levels;
  pg_t* spg; unsigned char* spg_on;   /* [(L+2)][nsp][3] */
  pg_t* mpg; unsigned char* mpg_on;
  double* beads;
  int npair;
  pair_t pair[16];
  int has_ewald;                   /* CoulombAction useEwald (optEwald) */
  ewald_t ew;
  double ceps;                     /* CoulombAction epsilon */
  double* ewr;                     /* [N][D] configuration scratch of the Ewald sum */
  int nfn;
  fn_t fn[PQ_MAX_NODES];
  double *sec1, *sec2, *mov;
  int idx[PQ_MAX_MOVING];
  int cyc[PQ_MAX_MOVING];          /* permutation of the moving particles */
  /* SerialPaths permutation / inversePermutation: bead (i, t + M) is bead
     (perm[i], t) (src/base/SerialPaths.cc:68-117) */
  int perm[MAXN], invperm[MAXN];
  int secperm[MAXN];               /* section permutation of section 1 */
  /* WalkingChooser tables t(i,j) and cump(i,j) of the active Sample, rows
     filled on first use (src/advancer/WalkingChooser.cc:89-118) */
  double *wt, *wcump;
  unsigned char* wrow;
  pg_t wpg[PQ_MAX_SPECIES][3];
  /* the section group being run (set_group) */
  int gnsample, gnsection;
  const pq_sample_desc* gsample;
  /* MultiLevelSampler::factor, the movers' sigma scale, and the mover's
     forwardProb of its last call (FreeMover.h:33-35, 45) */
  double fac, fwd;
  double* rej;                     /* delayed rejection: the rejected moving beads */
  oq_rng rng;
  pq_chain_stats st;
};

static int pg_index(const oq_chain* ch, int level, int sp, int d) {
  return (level * ch->nsp + sp) * 3 + d;
}

/* the sweep's g-th <ChooseSection nlevel> group: its level count, section
   length and Samples, and the WalkingChooser Gaussians of its Samples,
   pg(d) = PeriodicGaussian(m/(2 tau 2^L), a_d) (WalkingChooser.cc:20-25),
   which each ChooseSection's Samples build for their own nlevel */
static void set_group(oq_chain* ch, int g) {
  const pq_chain_desc* d = &ch->d;
  ch->L = g ? d->group[g - 1].nlevel : d->nlevel;
  ch->S = (1 << ch->L) + 1;
  ch->gnsample = g ? d->group[g - 1].nsample : d->nsample;
  ch->gsample = g ? d->group[g - 1].sample : d->sample;
  ch->gnsection = g ? d->group[g - 1].nsection : d->nsection;
  for (int s = 0; s < ch->nsp; ++s)
    for (int k = 0; k < ch->D; ++k)
      pg_init(&ch->wpg[s][k], d->species[s].mass / (2 * ch->tau * (1 << ch->L)), d->cell[k]);
}

oq_chain* oq_create(const pq_chain_desc* desc) {
  /* a Sample moves 1..PQ_MAX_MOVING distinct particles; more than its
     particle pool holds would make the reference's choosers draw forever */
  int ntot = 0, maxcount = 1;
  for (int s = 0; s < desc->nspecies; ++s) {
    ntot += desc->species[s].count;
    if (desc->species[s].count > maxcount) maxcount = desc->species[s].count;
  }
  if (desc->ngroup < 0 || desc->ngroup > PQ_MAX_GROUPS - 1) return NULL;
  int lmax = desc->nlevel;
  for (int g = 0; g <= desc->ngroup; ++g) {
    const int ns = g ? desc->group[g - 1].nsample : desc->nsample;
    const pq_sample_desc* smp = g ? desc->group[g - 1].sample : desc->sample;
    if (g && desc->group[g - 1].nlevel > lmax) lmax = desc->group[g - 1].nlevel;
    for (int k = 0; k < ns; ++k) {
      const pq_sample_desc* sm = &smp[k];
      if (sm->chooser < PQ_CHOOSER_SPECIES || sm->chooser > PQ_CHOOSER_ALL_RANDOM) return NULL;
      if (sm->nmoving < 1 || sm->nmoving > PQ_MAX_MOVING) return NULL;
      const int pool = (sm->chooser == PQ_CHOOSER_ALL || sm->chooser == PQ_CHOOSER_ALL_RANDOM)
                           ? ntot : desc->species[sm->species].count;
      if (sm->nmoving > pool) return NULL;
    }
  }
  const int smax = (1 << lmax) + 1;
  oq_chain* ch = (oq_chain*)calloc(1, sizeof(oq_chain));
  if (ch) ch->fac = 1.0;
  ch->d = *desc;
  ch->D = desc->ndim;
  ch->M = desc->nslice;
  ch->L = desc->nlevel;
  ch->S = (1 << ch->L) + 1;
  ch->nsp = desc->nspecies;
  ch->tau = desc->tau;
  ch->T = 1.0 / (desc->tau * desc->nslice);
  int n = 0;
  for (int s = 0; s < ch->nsp; ++s) {
    ch->ifirst[s] = n;
    for (int i = 0; i < desc->species[s].count; ++i) {
      ch->spec[n] = s;
      ch->lambda[n] = 0.5 / desc->species[s].mass;
      ++n;
    }
  }
  ch->N = n;
  cell_init(&ch->cell, desc->cell, ch->D);
  /* SpringAction pg(level,species,dim): alpha = m/(2 tau 2^l) (SpringAction.cc:31-49);
     FreeMover / FreeMoverPBC: alpha = m/(tau 2^l), levels 0..L+1 (FreeMoverPBC.cc:37-51) */
  ch->npg_levels = lmax + 2;
  int npg = ch->npg_levels * ch->nsp * 3;
  ch->spg = (pg_t*)calloc(npg, sizeof(pg_t));
  ch->mpg = (pg_t*)calloc(npg, sizeof(pg_t));
  ch->spg_on = (unsigned char*)calloc(npg, 1);
  ch->mpg_on = (unsigned char*)calloc(npg, 1);
  for (int l = 0; l < ch->npg_levels; ++l)
    for (int s = 0; s < ch->nsp; ++s)
      for (int d = 0; d < ch->D; ++d) {
        double m = desc->species[s].mass, len = desc->cell[d];
        double as = m / (2 * ch->tau * (1 << l));
        double am = m / (ch->tau * (1 << l));
        int k = pg_index(ch, l, s, d);
        if (l <= lmax && oq_pg_nterms(as, len) < 16) { ch->spg_on[k] = 1; pg_init(&ch->spg[k], as, len); }
        if (oq_pg_nterms(am, len) < 16) { ch->mpg_on[k] = 1; pg_init(&ch->mpg[k], am, len); }
      }
  ch->beads = (double*)calloc((size_t)ch->M * ch->N * ch->D, sizeof(double));
  /* CoulombAction constructor (src/action/CoulombAction.cc:45-90): one
     PairAction per interacting species pair; rmin/rmax/ngpts defaults are
     written back and inherited by later pairs. */
  ch->ewr = (double*)calloc((size_t)ch->N * ch->D, sizeof(double));
  if (desc->coulomb.enabled) {
    double eps = desc->coulomb.epsilon == 0 ? 1 : desc->coulomb.epsilon;
    ch->ceps = eps;
    /* the Ewald sum exists before the pair tables that subtract f(r)
       (CoulombAction.cc:28-44); its charges are the particles' own */
    if (desc->coulomb.use_ewald && (ch->D == 2 || ch->D == 3)) {
      double rc = desc->coulomb.ewald_rcut, q[MAXN];
      if (rc == 0.) rc = desc->cell[0] / 2.;
      for (int i = 0; i < ch->N; ++i) q[i] = desc->species[ch->spec[i]].charge;
      double kappa = desc->coulomb.ewald_kappa;
      if (desc->coulomb.ewald_type == 1 && kappa == 0) {
        /* the parser's default (ActionParser.cc:137-146) */
        double lside = 1000000000000000.0;
        for (int d = 0; d < ch->D; ++d) if (lside > desc->cell[d]) lside = desc->cell[d];
        kappa = sqrt(pow(3.6 * ch->N, 1.0 / 6.0) * sqrt(3.1415926535897931) / lside);
      }
      ewald_init(&ch->ew, &ch->cell, ch->N, q, rc, desc->coulomb.ewald_kcut, desc->coulomb.ewald_type == 1, kappa);
      ch->has_ewald = 1;
    }
    double rmin = desc->coulomb.rmin, rmax = desc->coulomb.rmax;
    int ngpts = desc->coulomb.ngpts;
    for (int i = 0; i < ch->nsp; ++i)
      for (int j = i; j < ch->nsp; ++j) {
        const pq_species_desc *s1 = &desc->species[i], *s2 = &desc->species[j];
        double q1q2 = s1->charge * s2->charge / eps;
        if (q1q2 != 0 && (i != j || s1->count > 1)) {
          double mu = 1. / (1. / s1->mass + 1. / s2->mass);
          if (rmin == 0) rmin = 0.005 / (2 * ((mu > 500) ? 1 : mu) * fabs(q1q2));
          if (rmax == 0) {
            double a2 = 0;
            for (int d = 0; d < ch->D; ++d) a2 += desc->cell[d] * desc->cell[d];
            rmax = sqrt(a2);
          }
          if (ngpts == 0) ngpts = 500;
          pair_t* p = &ch->pair[ch->npair++];
          p->s1 = i; p->s2 = j;
          p->norder = (mu > 500) ? 0 : desc->coulomb.norder;
          p->ngpts = ngpts;
          /* PairAction(EmpiricalPairAction) (src/action/PairAction.cc:161-185) */
          p->rgridinv = 1.0 / rmin;
          p->logrratioinv = (ngpts - 1) / log(rmax / rmin);
          int K1 = p->norder + 1;
          p->u = (double*)calloc((size_t)ngpts * K1, sizeof(double));
          p->ut = (double*)calloc((size_t)ngpts * K1, sizeof(double));
          for (int ig = 0; ig < ngpts; ++ig) {
            double r = (1. / p->rgridinv) * exp(ig / p->logrratioinv);
            for (int k = 0; k < K1; ++k) {
              p->u[ig * K1 + k] = coulomb_u(r, k, mu, q1q2, ch->tau, ch->has_ewald ? &ch->ew : NULL);
              p->ut[ig * K1 + k] = coulomb_utau(r, k, mu, q1q2, ch->tau, ch->has_ewald ? &ch->ew : NULL);
            }
          }
        }
      }
  }
  /* FixedNodeAction + node model per <FixedNodeAction> */
  ch->nfn = desc->nnodes;
  for (int a = 0; a < ch->nfn; ++a) {
    const pq_node_desc* nd = &desc->nodes[a];
    fn_t* f = &ch->fn[a];
    f->sp = nd->species;
    f->model = nd->model;
    f->ifirst = ch->ifirst[nd->species];
    f->ns = desc->species[nd->species].count;
    f->nodal = nd->nodal_action;
    f->mass = desc->species[nd->species].mass;
    f->omega = nd->omega;
    f->T = (nd->temperature == 0) ? ch->T : nd->temperature;
    f->scale = 1.0;
    /* SHONodes constructor (src/fixednode/SHONodes.cc:22-35) */
    f->coshwt = cosh(f->omega * 0.5 / f->T);
    f->sinhwt = sinh(f->omega * 0.5 / f->T);
    f->c = f->mass * 0.5 * f->omega / f->sinhwt;
    /* FreeParticleNodes pg[d] = PeriodicGaussian(mass*T, a_d) (FreeParticleNodes.cc:54-58) */
    for (int d = 0; d < ch->D; ++d) pg_init(&f->pg[d], f->mass * f->T, desc->cell[d]);
    f->dm = (double*)calloc(smax, sizeof(double));
    f->newdm = (double*)calloc(smax, sizeof(double));
    f->dist = (double*)calloc((size_t)smax * 2 * f->ns, sizeof(double));
    f->newdist = (double*)calloc((size_t)smax * 2 * f->ns, sizeof(double));
    f->inv = (double*)calloc((size_t)smax * f->ns * f->ns, sizeof(double));
    f->upd = nd->use_updates != 0;
    if (f->upd) {
      f->phi = (double*)calloc((size_t)smax * f->ns * PQ_MAX_MOVING, sizeof(double));
      f->newinv = (double*)calloc((size_t)smax * f->ns * f->ns, sizeof(double));
    }
  }
  size_t secsz = (size_t)smax * ch->N * ch->D;
  ch->sec1 = (double*)calloc(secsz, sizeof(double));
  ch->sec2 = (double*)calloc(secsz, sizeof(double));
  ch->mov = (double*)calloc((size_t)smax * PQ_MAX_MOVING * ch->D, sizeof(double));
  ch->rej = (double*)calloc((size_t)smax * PQ_MAX_MOVING * ch->D, sizeof(double));
  for (int i = 0; i < ch->N; ++i) ch->perm[i] = ch->invperm[i] = i;
  ch->wt = (double*)calloc((size_t)maxcount * maxcount, sizeof(double));
  ch->wcump = (double*)calloc((size_t)maxcount * maxcount, sizeof(double));
  ch->wrow = (unsigned char*)calloc(maxcount, 1);
  set_group(ch, 0);
  oq_rng_glibc(ch, 0);
  return ch;
}

void oq_destroy(oq_chain* ch) {
  if (!ch) return;
  for (int p = 0; p < ch->npair; ++p) { free(ch->pair[p].u); free(ch->pair[p].ut); }
  for (int a = 0; a < ch->nfn; ++a) {
    free(ch->fn[a].dm); free(ch->fn[a].newdm); free(ch->fn[a].dist);
    free(ch->fn[a].newdist); free(ch->fn[a].inv);
    free(ch->fn[a].phi); free(ch->fn[a].newinv);
  }
  free(ch->spg); free(ch->mpg); free(ch->spg_on); free(ch->mpg_on);
  free(ch->beads); free(ch->sec1); free(ch->sec2); free(ch->mov); free(ch->rej);
  free(ch->wt); free(ch->wcump); free(ch->wrow);
  if (ch->has_ewald) ewald_free(&ch->ew);
  free(ch->ewr);
  free(ch);
}

void oq_set_beads(oq_chain* ch, const double* b) {
  memcpy(ch->beads, b, sizeof(double) * ch->M * ch->N * ch->D);
}

/* Paths permutation as ReadPaths sets it (src/algorithm/PathReader.cc:60-75);
   returns -1 unless p is a permutation of 0..N-1 */
int oq_set_permutation(oq_chain* ch, const int* p) {
  unsigned char seen[MAXN] = {0};
  for (int i = 0; i < ch->N; ++i) {
    if (p[i] < 0 || p[i] >= ch->N || seen[p[i]]) return -1;
    seen[p[i]] = 1;
  }
  for (int i = 0; i < ch->N; ++i) ch->perm[i] = p[i];
  for (int i = 0; i < ch->N; ++i) ch->invperm[p[i]] = i;
  return 0;
}
void oq_get_permutation(const oq_chain* ch, int* p) {
  for (int i = 0; i < ch->N; ++i) p[i] = ch->perm[i];
}
void oq_get_beads(const oq_chain* ch, double* b) {
  memcpy(b, ch->beads, sizeof(double) * ch->M * ch->N * ch->D);
}
int oq_npair(const oq_chain* ch) { return ch->npair; }

int oq_pair_table(const oq_chain* ch, int ip, int* meta, double* grid, double* u, double* ut) {
  if (ip < 0 || ip >= ch->npair) return -1;
  const pair_t* p = &ch->pair[ip];
  meta[0] = p->s1; meta[1] = p->s2; meta[2] = p->norder; meta[3] = p->ngpts;
  grid[0] = p->rgridinv; grid[1] = p->logrratioinv;
  size_t n = (size_t)p->ngpts * (p->norder + 1);
  if (u) memcpy(u, p->u, n * sizeof(double));
  if (ut) memcpy(ut, p->ut, n * sizeof(double));
  return 0;
}

void oq_rng_glibc(oq_chain* ch, unsigned int seed) {
  memset(&ch->rng, 0, sizeof(ch->rng));
  ch->rng.mode = 0;
  glibc_seed(&ch->rng, seed);
}
void oq_rng_philox(oq_chain* ch, uint64_t seed, uint32_t stream) {
  memset(&ch->rng, 0, sizeof(ch->rng));
  ch->rng.mode = 1; ch->rng.seed = seed; ch->rng.stream = stream;
}
void oq_rng_inject(oq_chain* ch, const double* u, int64_t n) {
  memset(&ch->rng, 0, sizeof(ch->rng));
  ch->rng.mode = 2; ch->rng.inj = u; ch->rng.ninj = n;
}
double oq_rng_next(oq_chain* ch) { return rng_next(&ch->rng); }
uint64_t oq_rng_count(const oq_chain* ch) { return ch->rng.count; }

int oq_rng_exhausted(const oq_chain* ch) {
  return ch->rng.mode == 2 && ch->rng.count > (uint64_t)ch->rng.ninj;
}

/* ---------------------------------------------------------- actions */

#define SEC(b, t, i, d) ((b)[((size_t)(t) * ch->N + (i)) * ch->D + (d)])
#define MOV(b, t, m, d) ((b)[((size_t)(t) * nmov + (m)) * ch->D + (d)])

/* SpringAction::getActionDifference (src/action/SpringAction.cc:56-98) */
static double spring_diff(oq_chain* ch, const double* sec, const double* mov,
                          const int* idx, int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  double dA = 0;
  for (int t = s; t < ch->S; t += s) {
    for (int m = 0; m < nmov; ++m) {
      int i = idx[m], sp = ch->spec[i];
      if (ch->d.species[sp].is_static) continue;
      double inv2Sigma2 = 0.25 / (ch->lambda[i] * ch->tau * s);
      double delta[3];
      for (int d = 0; d < D; ++d) delta[d] = MOV(mov, t, m, d) - MOV(mov, t - s, m, d);
      pbc(&ch->cell, delta);
      double temp = 1.0;
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->spg_on[k]) temp *= pg_eval(&ch->spg[k], delta[d]);
        else dA += delta[d] * delta[d] * inv2Sigma2;
      }
      temp = 1.0 / temp;
      for (int d = 0; d < D; ++d) delta[d] = SEC(sec, t, i, d) - SEC(sec, t - s, i, d);
      pbc(&ch->cell, delta);
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->spg_on[k]) temp *= pg_eval(&ch->spg[k], delta[d]);
        else dA -= delta[d] * delta[d] * inv2Sigma2;
      }
      dA += log(temp);
    }
  }
  return dA;
}

/* SHOAction::getActionDifference (src/action/SHOAction.cc:19-65) */
static double sho_diff(oq_chain* ch, const pq_sho_desc* a, const double* sec,
                       const double* mov, const int* idx, int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  const int ifirst = ch->ifirst[a->species], np = ch->d.species[a->species].count;
  const int ndim = a->ndim ? a->ndim : D;
  const double omega = a->omega, m = ch->d.species[a->species].mass;
  double dA = 0;
  double wt = omega * ch->tau * s;
  double sinhwt = sinh(wt), coshwt = cosh(wt);
  double div1 = 1.0 / (2.0 * sinhwt), div2 = 1.0 / (2.0 * ch->tau * s);
  for (int t = s; t < ch->S; t += s) {
    for (int mm = 0; mm < nmov; ++mm) {
      int i = idx[mm];
      if (i < ifirst || i >= ifirst + np) continue;
      for (int pass = 0; pass < 2; ++pass) {
        double r1[3], r0[3];
        for (int d = 0; d < D; ++d) {
          if (pass == 0) { r1[d] = MOV(mov, t, mm, d) - a->center[d]; r0[d] = MOV(mov, t - s, mm, d) - a->center[d]; }
          else { r1[d] = SEC(sec, t, i, d) - a->center[d]; r0[d] = SEC(sec, t - s, i, d) - a->center[d]; }
        }
        double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
        for (int d = D - ndim; d < D; ++d) {
          double dl = r1[d] - r0[d];
          r0r0 += r0[d] * r0[d]; r0r1 += r0[d] * r1[d]; r1r1 += r1[d] * r1[d]; delta2 += dl * dl;
        }
        double v = m * omega * ((r1r1 + r0r0) * coshwt - 2.0 * r0r1) * div1 - m * delta2 * div2;
        if (pass == 0) dA += v; else dA -= v;
      }
    }
  }
  return dA;
}

/* PairAction grid lookups u00 / uk0 (src/action/PairAction.cc:601-621) */
static void grid_locate(const pair_t* p, double q, int* ig, double* x) {
  q *= p->rgridinv;
  double xx = log(q) * p->logrratioinv;
  int i = (int)xx;
  if (q < 1) { i = 0; xx = 0; }
  else if (i > p->ngpts - 2) { i = p->ngpts - 2; xx = 1; }
  else xx -= i;
  *ig = i; *x = xx;
}

static double u00(const pair_t* p, double r) {
  int i; double x;
  grid_locate(p, r, &i, &x);
  int K1 = p->norder + 1;
  return (1 - x) * p->u[i * K1] + x * p->u[(i + 1) * K1];
}

static double uk0(const pair_t* p, double q, double s2) {
  int i; double x;
  grid_locate(p, q, &i, &x);
  int K1 = p->norder + 1;
  double action = 0;
  for (int k = p->norder; k >= 0; k--) {
    action *= s2;
    action += (1 - x) * p->u[i * K1 + k] + x * p->u[(i + 1) * K1 + k];
  }
  return action;
}

/* uk0CalcDerivatives, the u and utau parts (src/action/PairAction.cc:640-657) */
static void uk0_u_utau(const pair_t* p, double q, double s2, double* u, double* ut) {
  int i; double x;
  grid_locate(p, q, &i, &x);
  int K1 = p->norder + 1;
  double a = 0, b = 0;
  for (int k = p->norder; k >= 0; k--) {
    a *= s2; a += (1 - x) * p->u[i * K1 + k] + x * p->u[(i + 1) * K1 + k];
    b *= s2; b += (1 - x) * p->ut[i * K1 + k] + x * p->ut[(i + 1) * K1 + k];
  }
  *u = a; *ut = b;
}

static double dot3(const double* a, const double* b, int D) {
  double s = 0;
  for (int d = 0; d < D; ++d) s += a[d] * b[d];
  return s;
}

/* PairAction::getActionDifference without exchange (src/action/PairAction.cc:219-360) */
static double pair_diff(oq_chain* ch, const pair_t* p, const double* sec,
                        const double* mov, const int* idx, int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  const int if1 = ch->ifirst[p->s1], n1 = ch->d.species[p->s1].count;
  const int if2 = ch->ifirst[p->s2], n2 = ch->d.species[p->s2].count;
  double dA = 0;
  for (int im = 0; im < nmov; ++im) {
    int i = idx[im], jb, je;
    if (i >= if1 && i < if1 + n1) { jb = if2; je = if2 + n2; }
    else if (i >= if2 && i < if2 + n2) { jb = if1; je = if1 + n1; }
    else continue;
    for (int j = jb; j < je; ++j) {
      int isMoving = 0, jm = 0;
      for (int k = 0; k < nmov; ++k) if (j == idx[k]) { isMoving = 1; jm = k; break; }
      if (isMoving && i <= j) continue;
      double prevDelta[3], prevMovingDelta[3];
      for (int d = 0; d < D; ++d) prevDelta[d] = SEC(sec, 0, i, d) - SEC(sec, 0, j, d);
      pbc(&ch->cell, prevDelta);
      for (int d = 0; d < D; ++d) prevMovingDelta[d] = prevDelta[d];
      double prevR = sqrt(dot3(prevDelta, prevDelta, D));
      double prevMovingR = prevR;
      for (int t = s; t < ch->S; t += s) {
        double delta[3], svec[3];
        for (int d = 0; d < D; ++d)
          delta[d] = MOV(mov, t, im, d) - (isMoving ? MOV(mov, t, jm, d) : SEC(sec, t, j, d));
        pbc(&ch->cell, delta);
        double r = sqrt(dot3(delta, delta, D));
        double q = 0.5 * (r + prevMovingR);
        double action = 0.;
        if (level < 3 && p->norder > 0) {
          for (int d = 0; d < D; ++d) svec[d] = delta[d] - prevMovingDelta[d];
          double s2 = dot3(svec, svec, D) / (q * q);
          action += uk0(p, q, s2) * s;
        } else {
          action += u00(p, q) * s;
        }
        dA += action;
        for (int d = 0; d < D; ++d) prevMovingDelta[d] = delta[d];
        prevMovingR = r;
        for (int d = 0; d < D; ++d) delta[d] = SEC(sec, t, i, d) - SEC(sec, t, j, d);
        pbc(&ch->cell, delta);
        r = sqrt(dot3(delta, delta, D));
        q = 0.5 * (r + prevR);
        action = 0.;
        if (level < 3 && p->norder > 0) {
          for (int d = 0; d < D; ++d) svec[d] = delta[d] - prevDelta[d];
          double s2 = dot3(svec, svec, D) / (q * q);
          action += uk0(p, q, s2) * s;
        } else {
          action += u00(p, q) * s;
        }
        dA -= action;
        for (int d = 0; d < D; ++d) prevDelta[d] = delta[d];
        prevR = r;
      }
    }
  }
  return dA;
}

/* CompositeAction::getActionDifference (src/action/CompositeAction.cc:30-37)
   and CoulombAction::getActionDifference (src/action/CoulombAction.cc:100-105) */
double oq_action_difference(oq_chain* ch, const double* sec, const double* mov,
                            const int* idx, int nmov, int level, int mask) {
  double diff = 0;
  if ((mask & 1) && ch->d.spring) diff += spring_diff(ch, sec, mov, idx, nmov, level);
  if (mask & 2)
    for (int a = 0; a < ch->d.nsho; ++a) diff += sho_diff(ch, &ch->d.sho[a], sec, mov, idx, nmov, level);
  if ((mask & 12) && (ch->npair || ch->has_ewald)) {
    double u = 0;
    if (mask & 4)
      for (int p = 0; p < ch->npair; ++p) u += pair_diff(ch, &ch->pair[p], sec, mov, idx, nmov, level);
    /* the long-range Ewald action at the lowest level, slices 1..2^L-1
       (CoulombAction.cc:107-119); bit 4 of mask takes CoulombAction whole,
       bit 8 alone its Ewald part */
    if (ch->has_ewald && level == 0 && (mask & 12)) {
      const size_t sl = (size_t)ch->N * ch->D;
      for (int t = 1; t < ch->S - 1; ++t) {
        memcpy(ch->ewr, &sec[t * sl], sl * sizeof(double));
        u -= ewald_long_range(&ch->ew, ch->ewr) * ch->tau / ch->ceps;
        for (int m = 0; m < nmov; ++m)
          for (int d = 0; d < ch->D; ++d) ch->ewr[idx[m] * ch->D + d] = MOV(mov, t, m, d);
        u += ewald_long_range(&ch->ew, ch->ewr) * ch->tau / ch->ceps;
      }
    }
    diff += u;
  }
  return diff;
}

/* path access for -M <= t < 2M: slices before 0 belong to
   inversePermutation[i], slices from M on to permutation[i]
   (SerialPaths::operator(), src/base/SerialPaths.cc:68-104) */
static const double* bead(const oq_chain* ch, int i, int t) {
  if (t < 0) { i = ch->invperm[i]; t += ch->M; }
  else if (t >= ch->M) { i = ch->perm[i]; t -= ch->M; }
  return &ch->beads[((size_t)t * ch->N + i) * ch->D];
}

/* EwaldSum::evalLongRange of path slice 0 <= t < M (Paths::getSlice) */
static double slice_long_range(const oq_chain* ch, int t) {
  return ewald_long_range(&ch->ew, &ch->beads[(size_t)t * ch->N * ch->D]);
}

/* getBeadAction utau of SpringAction (SpringAction.cc:112-123), SHOAction
   (SHOAction.cc:71-105), PairAction (PairAction.cc:509-576), summed as
   CompositeAction::getBeadAction (CompositeAction.cc:57-67). */
double oq_bead_utau(const oq_chain* ch, int ipart, int islice, int mask) {
  const int D = ch->D;
  double utau = 0;
  if ((mask & 1) && ch->d.spring && !ch->d.species[ch->spec[ipart]].is_static) {
    double delta[3];
    const double *a = bead(ch, ipart, islice), *b = bead(ch, ipart, islice - 1);
    for (int d = 0; d < D; ++d) delta[d] = a[d] - b[d];
    pbc(&ch->cell, delta);
    double ui = D / (2 * ch->tau) - dot3(delta, delta, D) / (4.0 * ch->lambda[ipart] * ch->tau * ch->tau);
    utau += ui;
  }
  if (mask & 2) {
    for (int s = 0; s < ch->d.nsho; ++s) {
      const pq_sho_desc* a = &ch->d.sho[s];
      int ifirst = ch->ifirst[a->species], np = ch->d.species[a->species].count;
      double ui = 0;
      if (!(ipart < ifirst || ipart >= ifirst + np)) {
        int ndim = a->ndim ? a->ndim : D;
        double omega = a->omega, m = ch->d.species[a->species].mass;
        double wt = omega * ch->tau, sinhwt = sinh(wt), coshwt = cosh(wt);
        double cschwt = 1.0 / sinhwt, cothwt = coshwt * cschwt, tin = 1.0 / ch->tau;
        const double *p1 = bead(ch, ipart, islice), *p0 = bead(ch, ipart, islice - 1);
        double r1[3], r0[3], delta[3];
        for (int d = 0; d < D; ++d) { r1[d] = p1[d] - a->center[d]; r0[d] = p0[d] - a->center[d]; delta[d] = p1[d] - p0[d]; }
        pbc(&ch->cell, delta);
        double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
        for (int d = D - ndim; d < D; ++d) {
          r0r0 += r0[d] * r0[d]; r0r1 += r0[d] * r1[d]; r1r1 += r1[d] * r1[d]; delta2 += delta[d] * delta[d];
        }
        ui = 0.5 * ndim * omega * cothwt
             + 0.5 * m * omega * omega * ((r1r1 + r0r0) * (1 - cothwt * cothwt) + 2.0 * r0r1 * coshwt * cschwt * cschwt)
             + 0.5 * m * delta2 * tin * tin - 0.5 * ndim * tin;
      }
      utau += ui;
    }
  }
  if ((mask & 4) && (ch->npair || ch->has_ewald)) {
    double uc = 0;
    for (int pp = 0; pp < ch->npair; ++pp) {
      const pair_t* p = &ch->pair[pp];
      int if1 = ch->ifirst[p->s1], n1 = ch->d.species[p->s1].count;
      int if2 = ch->ifirst[p->s2], n2 = ch->d.species[p->s2].count;
      int jb, je;
      double up = 0;
      if (ipart >= if1 && ipart < if1 + n1) { jb = if2; je = if2 + n2; }
      else if (ipart >= if2 && ipart < if2 + n2) { jb = if1; je = if1 + n1; }
      else { uc += up; continue; }
      for (int j = jb; j < je; ++j) {
        if (ipart == j) continue;
        double delta[3], prevDelta[3], svec[3];
        const double *ai = bead(ch, ipart, islice), *aj = bead(ch, j, islice);
        const double *bi = bead(ch, ipart, islice - 1), *bj = bead(ch, j, islice - 1);
        for (int d = 0; d < D; ++d) delta[d] = ai[d] - aj[d];
        pbc(&ch->cell, delta);
        double r = sqrt(dot3(delta, delta, D));
        for (int d = 0; d < D; ++d) prevDelta[d] = bi[d] - bj[d];
        pbc(&ch->cell, prevDelta);
        double prevR = sqrt(dot3(prevDelta, prevDelta, D));
        double q = 0.5 * (r + prevR);
        for (int d = 0; d < D; ++d) svec[d] = delta[d] - prevDelta[d];
        double s2 = dot3(svec, svec, D) / (q * q);
        double v, vtau;
        uk0_u_utau(p, q, s2, &v, &vtau);
        up += 0.5 * vtau;
      }
      uc += up;
    }
    /* the long-range part rides on particle 0 (CoulombAction.cc:164-170) */
    if (ch->has_ewald && ipart == 0) uc += slice_long_range(ch, islice) / ch->ceps;
    utau += uc;
  }
  return utau;
}

/* -------------------------------------------------------- node models */

/* Unblocked LU with partial pivoting (as LAPACK dgetf2) on a row-major n x n
   matrix; returns info (0 ok, k+1 for the first zero pivot) and the pivot
   sign product; then the inverse A^-1 = U^-1 L^-1 P written row-major. */
static int lu_det_inverse(double* a, int n, double* det, double* inv) {
  int ipiv[PQ_MAX_NODE_PART];
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(a[k * n + k]);
    for (int r = k + 1; r < n; ++r)
      if (fabs(a[r * n + k]) > amax) { amax = fabs(a[r * n + k]); p = r; }
    ipiv[k] = p;
    if (a[p * n + k] != 0.0) {
      if (p != k)
        for (int c = 0; c < n; ++c) { double t = a[k * n + c]; a[k * n + c] = a[p * n + c]; a[p * n + c] = t; }
      for (int r = k + 1; r < n; ++r) a[r * n + k] /= a[k * n + k];
    } else if (info == 0) {
      info = k + 1;
    }
    for (int r = k + 1; r < n; ++r)
      for (int c = k + 1; c < n; ++c) a[r * n + c] -= a[r * n + k] * a[k * n + c];
  }
  double d = 1;
  for (int i = 0; i < n; ++i) { d *= a[i * n + i]; d *= (i == ipiv[i]) ? 1 : -1; }
  *det = d;
  if (info) return info;
  for (int j = 0; j < n; ++j) {
    double x[PQ_MAX_NODE_PART];
    for (int r = 0; r < n; ++r) x[r] = (r == j) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) if (ipiv[k] != k) { double t = x[k]; x[k] = x[ipiv[k]]; x[ipiv[k]] = t; }
    for (int r = 0; r < n; ++r) for (int c = 0; c < r; ++c) x[r] -= a[r * n + c] * x[c];
    for (int r = n - 1; r >= 0; --r) {
      for (int c = r + 1; c < n; ++c) x[r] -= a[r * n + c] * x[c];
      x[r] /= a[r * n + r];
    }
    for (int r = 0; r < n; ++r) inv[r * n + j] = x[r];
  }
  return 0;
}

/* density-matrix element rho(i,j): row i from r2 (section +beta/2), column j
   from r1 (SHONodes.cc:58-67, FreeParticleNodes.cc:79-90) */
static double node_elem(const oq_chain* ch, const fn_t* f, const double* r1,
                        const double* r2, int i, int j) {
  const int D = ch->D;
  const double* rj = &r1[(size_t)(f->ifirst + j) * D];
  const double* ri = &r2[(size_t)(f->ifirst + i) * D];
  if (f->model == PQ_NODES_SHO) {
    double rj2 = dot3(rj, rj, D), ri2 = dot3(ri, ri, D), rirj = dot3(ri, rj, D);
    return f->scale * exp(-f->c * ((ri2 + rj2) * f->coshwt - 2.0 * rirj));
  }
  double delta[3];
  for (int d = 0; d < D; ++d) delta[d] = rj[d] - ri[d];
  pbc(&ch->cell, delta);
  double ear2 = f->scale;
  for (int d = 0; d < D; ++d) ear2 *= pg_eval(&f->pg[d], delta[d]);
  return ear2;
}

/* NodeModel::evaluate for SHONodes (SHONodes.cc:51-128) and FreeParticleNodes
   (FreeParticleNodes.cc:72-157): returns err, det in *det; the inverse of the
   slice is kept for evaluateDistance. */
static int node_evaluate(oq_chain* ch, fn_t* f, const double* r1, const double* r2,
                         int islice, int scaleMag, double* detOut) {
  const int n = f->ns;
  double a[PQ_MAX_NODE_PART * PQ_MAX_NODE_PART];
  int err = 0, pass = 0;
  double det = 0;
  do {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) a[i * n + j] = node_elem(ch, f, r1, r2, i, j);
    int info = lu_det_inverse(a, n, &det, &f->inv[(size_t)islice * n * n]);
    err = info != 0;
    if (err) f->nerror++;
    if (scaleMag && (!err && (fabs(det) < 1e-50 || fabs(det) > 1e50))) {
      if (fabs(det) < 1e-250) f->scale *= 2;
      else f->scale *= pow(fabs(det), -1. / n);
    }
    if (++pass >= 200) break;
  } while (scaleMag && (err || (fabs(det) < 1e-50 || fabs(det) > 1e50)));
  *detOut = det;
  return err;
}

/* evaluateDistance: SHONodes.cc:130-170, FreeParticleNodes.cc:184-234 */
static void node_distance(oq_chain* ch, fn_t* f, const double* r1, const double* r2,
                          int islice, double* d1, double* d2) {
  const int n = f->ns, D = ch->D;
  const double* inv = &f->inv[(size_t)islice * n * n];
  for (int pass = 0; pass < 2; ++pass) {
    for (int a = 0; a < n; ++a) {
      double lg[3] = {0, 0, 0};
      for (int b = 0; b < n; ++b) {
        int jp = pass == 0 ? a : b, ip = pass == 0 ? b : a;
        const double* rj = &r1[(size_t)(f->ifirst + jp) * D];
        const double* ri = &r2[(size_t)(f->ifirst + ip) * D];
        double grad[3];
        double w = inv[jp * n + ip];
        if (f->model == PQ_NODES_SHO) {
          double ri2 = dot3(ri, ri, D), rj2 = dot3(rj, rj, D), rirj = dot3(ri, rj, D);
          double e = exp(-f->c * ((ri2 + rj2) * f->coshwt - 2 * rirj));
          for (int d = 0; d < D; ++d) {
            double g = pass == 0 ? 2 * f->c * (ri[d] - f->coshwt * rj[d]) * e
                                 : 2 * f->c * (rj[d] - f->coshwt * ri[d]) * e;
            lg[d] += w * f->scale * g;
          }
        } else {
          double delta[3], value[3];
          for (int d = 0; d < D; ++d) delta[d] = pass == 0 ? rj[d] - ri[d] : ri[d] - rj[d];
          pbc(&ch->cell, delta);
          for (int d = 0; d < D; ++d) {
            double gr;
            value[d] = pg_eval3(&f->pg[d], delta[d], &gr, 0);
            grad[d] = gr / (value[d] + 1e-300);
          }
          for (int k = 0; k < D; ++k) for (int d = 0; d < D; ++d) grad[d] *= value[k];
          for (int d = 0; d < D; ++d) lg[d] += w * grad[d] * f->scale;
        }
      }
      double dist = sqrt(2 * f->mass / ((dot3(lg, lg, D) + 1e-15) * ch->tau));
      if (pass == 0) d1[a] = dist; else d2[a] = dist;
    }
  }
}

int oq_node_evaluate(oq_chain* ch, int inode, const double* r1, const double* r2,
                     int islice, int scaleMag, double* det) {
  return node_evaluate(ch, &ch->fn[inode], r1, r2, islice, scaleMag, det);
}
void oq_node_distance(oq_chain* ch, int inode, const double* r1, const double* r2,
                      int islice, double* d1, double* d2) {
  node_distance(ch, &ch->fn[inode], r1, r2, islice, d1, d2);
}
double oq_node_scale(const oq_chain* ch, int inode) { return ch->fn[inode].scale; }

/* ------------------------------------ FreeParticleNodes::MatrixUpdate */

/* determinant of the k x k row-major matrix a by LU with partial pivoting
   (dgetf2 order, as lu_det_inverse); a is overwritten */
static double small_det(double* a, int k) {
  double det = 1;
  for (int c = 0; c < k; ++c) {
    int p = c;
    double amax = fabs(a[c * k + c]);
    for (int r = c + 1; r < k; ++r)
      if (fabs(a[r * k + c]) > amax) { amax = fabs(a[r * k + c]); p = r; }
    if (a[p * k + c] != 0.0) {
      if (p != c)
        for (int q = 0; q < k; ++q) { double t = a[c * k + q]; a[c * k + q] = a[p * k + q]; a[p * k + q] = t; }
      for (int r = c + 1; r < k; ++r) a[r * k + c] /= a[c * k + c];
    }
    for (int r = c + 1; r < k; ++r)
      for (int q = c + 1; q < k; ++q) a[r * k + q] -= a[r * k + c] * a[c * k + q];
    det *= a[c * k + c];
    det *= (p == c) ? 1 : -1;
  }
  return det;
}

/* MatrixUpdate::evaluateChange (FreeParticleNodes.cc:894-948): the moving
   particles' new columns phi(k, jm) = scale prod_d PG(x_moving - x2_k) at
   slice t, b(i, jm) = sum_k inv(i, k) phi(k, jm) on the moving rows, and the
   determinant of that nm x nm block, the ratio det(new)/det(old). */
static double upd_change(oq_chain* ch, fn_t* f, const double* mov, const int* idx, int nmov, int t) {
  const int n = f->ns, D = ch->D;
  const size_t slice = (size_t)ch->N * D;
  f->new_updated = 0;
  f->nm = 0;
  for (int m = 0; m < nmov; ++m)
    if (idx[m] >= f->ifirst && idx[m] < f->ifirst + n) {
      f->index1[f->nm] = idx[m] - f->ifirst;
      f->mindex1[f->nm] = m;
      f->nm++;
    }
  const int nm = f->nm;
  double* phi = &f->phi[(size_t)t * n * PQ_MAX_MOVING];
  const double* inv = &f->inv[(size_t)t * n * n];
  double b[PQ_MAX_MOVING * PQ_MAX_MOVING];
  for (int jm = 0; jm < nm; ++jm) {
    const double* rm = &mov[((size_t)t * nmov + f->mindex1[jm]) * D];
    for (int k = 0; k < n; ++k) {
      const double* r2 = &ch->sec2[t * slice + (size_t)(f->ifirst + k) * D];
      double delta[3];
      for (int d = 0; d < D; ++d) delta[d] = rm[d] - r2[d];
      pbc(&ch->cell, delta);
      double ear2 = f->scale;
      for (int d = 0; d < D; ++d) ear2 *= pg_eval(&f->pg[d], delta[d]);
      phi[k * PQ_MAX_MOVING + jm] = ear2;
    }
    for (int im = 0; im < nm; ++im) {
      const int ip = f->index1[im];
      double v = 0;
      for (int k = 0; k < n; ++k) v += phi[k * PQ_MAX_MOVING + jm] * inv[ip * n + k];
      b[im * nm + jm] = v;
    }
  }
  return small_det(b, nm);
}

/* the Sherman-Morrison row updates of evaluateNewInverse and of
   acceptLastMove (FreeParticleNodes.cc:950-974, 1026-1057) on m (n x n),
   one moving particle after the other */
static void upd_rows(const fn_t* f, double* m, const double* phi) {
  const int n = f->ns;
  for (int jm = 0; jm < f->nm; ++jm) {
    const int jp = f->index1[jm];
    double bjjinv = 0;
    for (int k = 0; k < n; ++k) bjjinv += phi[k * PQ_MAX_MOVING + jm] * m[jp * n + k];
    bjjinv = 1. / bjjinv;
    for (int k = 0; k < n; ++k) m[jp * n + k] *= bjjinv;
    for (int ip = 0; ip < n; ++ip) {
      if (ip == jp) continue;
      double bij = 0;
      for (int k = 0; k < n; ++k) bij += phi[k * PQ_MAX_MOVING + jm] * m[ip * n + k];
      for (int k = 0; k < n; ++k) m[ip * n + k] -= m[jp * n + k] * bij;
    }
  }
}

/* MatrixUpdate::evaluateNewInverse (FreeParticleNodes.cc:950-974) */
static void upd_new_inverse(oq_chain* ch, fn_t* f, int t) {
  const int n = f->ns;
  double* m = &f->newinv[(size_t)t * n * n];
  memcpy(m, &f->inv[(size_t)t * n * n], sizeof(double) * n * n);
  upd_rows(f, m, &f->phi[(size_t)t * n * PQ_MAX_MOVING]);
  f->new_updated = 1;
}

/* MatrixUpdate::evaluateNewDistance (FreeParticleNodes.cc:976-1024) with the
   updated inverse.  Unlike evaluateDistance it leaves the scale out of the
   log gradient, and in its first loop (d1) each gradient component is
   multiplied by its own dimension's value only (grad[i] *= value[i]); the
   second loop (d2) multiplies the whole vector by every value. */
static void upd_distance(oq_chain* ch, fn_t* f, const double* r1, const double* r2, int t,
                         double* d1, double* d2) {
  const int n = f->ns, D = ch->D;
  const double* m = &f->newinv[(size_t)t * n * n];
  for (int pass = 0; pass < 2; ++pass)
    for (int a = 0; a < n; ++a) {
      double lg[3] = {0, 0, 0};
      for (int b = 0; b < n; ++b) {
        const int jp = pass == 0 ? a : b, ip = pass == 0 ? b : a;
        const double* rj = &r1[(size_t)(f->ifirst + jp) * D];
        const double* ri = &r2[(size_t)(f->ifirst + ip) * D];
        double delta[3], value[3], grad[3];
        for (int d = 0; d < D; ++d) delta[d] = pass == 0 ? rj[d] - ri[d] : ri[d] - rj[d];
        pbc(&ch->cell, delta);
        for (int d = 0; d < D; ++d) {
          double gr;
          value[d] = pg_eval3(&f->pg[d], delta[d], &gr, 0);
          grad[d] = gr / (value[d] + 1e-300);
        }
        if (pass == 0) {
          for (int d = 0; d < D; ++d) grad[d] *= value[d];
        } else {
          for (int k = 0; k < D; ++k) for (int d = 0; d < D; ++d) grad[d] *= value[k];
        }
        const double w = m[jp * n + ip];
        for (int d = 0; d < D; ++d) lg[d] += w * grad[d];
      }
      const double dist = sqrt(2 * f->mass / ((dot3(lg, lg, D) + 1e-15) * ch->tau));
      if (pass == 0) d1[a] = dist; else d2[a] = dist;
    }
}

/* MatrixUpdate::acceptLastMove (FreeParticleNodes.cc:1026-1057): slices
   1..S-2 take the updated inverses, or are updated in place when no level-0
   nodal action made them; the last slice keeps its inverse */
static void upd_accept(oq_chain* ch, fn_t* f) {
  const int n = f->ns;
  for (int t = 1; t < ch->S - 1; ++t) {
    double* m = &f->inv[(size_t)t * n * n];
    if (f->new_updated) memcpy(m, &f->newinv[(size_t)t * n * n], sizeof(double) * n * n);
    else upd_rows(f, m, &f->phi[(size_t)t * n * PQ_MAX_MOVING]);
  }
}

/* FixedNodeAction::initialize (src/fixednode/FixedNodeAction.cc:331-357) */
void oq_fn_initialize(oq_chain* ch, const double* sec1, const double* sec2) {
  const size_t slice = (size_t)ch->N * ch->D;
  if (sec1 != ch->sec1) memcpy(ch->sec1, sec1, sizeof(double) * ch->S * slice);
  if (sec2 != ch->sec2) memcpy(ch->sec2, sec2, sizeof(double) * ch->S * slice);
  for (int a = 0; a < ch->nfn; ++a) {
    fn_t* f = &ch->fn[a];
    const int ns = f->ns;
    for (int t = 0; t < ch->S; ++t) {
      const double* r1 = &ch->sec1[t * slice];
      const double* r2 = &ch->sec2[t * slice];
      double det;
      int err = node_evaluate(ch, f, r1, r2, t, 1, &det);
      f->dm[t] = det;
      if (err || f->dm[t] * f->dm[0] <= 1e-200) { f->nerror++; ch->st.node_errors++; }
      if (f->nodal) node_distance(ch, f, r1, r2, t, &f->dist[(t * 2 + 0) * ns], &f->dist[(t * 2 + 1) * ns]);
    }
    f->newdm[0] = f->dm[0];
    for (int k = 0; k < 2 * ns; ++k) f->newdist[k] = f->dist[k];
  }
}

/* FixedNodeAction::getActionDifference (src/fixednode/FixedNodeAction.cc:39-131),
   summed over FixedNodeActions as CompositeDoubleAction. */
double oq_fn_action_difference(oq_chain* ch, const double* mov, const int* idx,
                               int nmov, int level) {
  const int D = ch->D, N = ch->N;
  const size_t slice = (size_t)N * D;
  double total = 0;
  double r1[MAXN * 3];
  for (int a = 0; a < ch->nfn; ++a) {
    fn_t* f = &ch->fn[a];
    const int ns = f->ns;
    double dA = 0;
    int mine = 0;
    for (int m = 0; m < nmov; ++m)
      if (idx[m] >= f->ifirst && idx[m] < f->ifirst + ns) { mine = 1; break; }
    if (!mine) { f->not_my_species = 1; total += 0; continue; }
    f->not_my_species = 0;
    const int nStride = 1 << (level + 1);
    int crossed = 0;
    for (int t = nStride / 2; t < ch->S; t += nStride) {
      if (f->upd) {
        /* newDMValue = evaluateChange * dmValue (FixedNodeAction.cc:61-63) */
        f->newdm[t] = upd_change(ch, f, mov, idx, nmov, t) * f->dm[t];
      } else {
        memcpy(r1, &ch->sec1[t * slice], sizeof(double) * slice);
        for (int m = 0; m < nmov; ++m)
          for (int d = 0; d < D; ++d) r1[idx[m] * D + d] = mov[((size_t)t * nmov + m) * D + d];
        double det;
        int err = node_evaluate(ch, f, r1, &ch->sec2[t * slice], t, 0, &det);
        if (err) { crossed = 1; break; }
        f->newdm[t] = det;
      }
      if (f->newdm[t] * f->dm[0] <= 1e-200) { crossed = 1; break; }
    }
    if (crossed) { dA = 2e100; total += dA; continue; }
    if (f->nodal && level == 0) {
      for (int t = 1; t < ch->S; ++t) {
        memcpy(r1, &ch->sec1[t * slice], sizeof(double) * slice);
        for (int m = 0; m < nmov; ++m)
          for (int d = 0; d < D; ++d) r1[idx[m] * D + d] = mov[((size_t)t * nmov + m) * D + d];
        if (f->upd) {
          /* evaluateNewInverse + evaluateNewDistance (FixedNodeAction.cc:79-91) */
          upd_new_inverse(ch, f, t);
          upd_distance(ch, f, r1, &ch->sec2[t * slice], t, &f->newdist[(t * 2 + 0) * ns],
                       &f->newdist[(t * 2 + 1) * ns]);
        } else {
          node_distance(ch, f, r1, &ch->sec2[t * slice], t, &f->newdist[(t * 2 + 0) * ns],
                        &f->newdist[(t * 2 + 1) * ns]);
        }
        const double* od = f->dist; const double* nd = f->newdist;
        if (ch->d.nodes[a].many_body_distance) {
          /* useManyBodyDistance (FixedNodeAction.cc:99-116): the sums run
             over all N particles, the other species' at distance 200 */
          double s[2][4] = {{0}};
          for (int i = 0; i < N; ++i)
            for (int p = 0; p < 2; ++p) {
              const int own = i >= f->ifirst && i < f->ifirst + ns;
              const int k = i - f->ifirst;
              const double o1 = own ? od[(t * 2 + p) * ns + k] : 200;
              const double o0 = own ? od[((t - 1) * 2 + p) * ns + k] : 200;
              const double n1 = own ? nd[(t * 2 + p) * ns + k] : 200;
              const double n0 = own ? nd[((t - 1) * 2 + p) * ns + k] : 200;
              s[p][0] += 1. / (o1 * o1); s[p][1] += 1. / (o0 * o0);
              s[p][2] += 1. / (n1 * n1); s[p][3] += 1. / (n0 * n0);
            }
          for (int p = 0; p < 2; ++p)
            dA += log((1 - exp(-1. / sqrt(s[p][0] * s[p][1]))) / (1 - exp(-1. / sqrt(s[p][2] * s[p][3]))));
          continue;
        }
        for (int i = 0; i < ns; ++i) {
          dA += log((1 - exp(-od[(t * 2 + 0) * ns + i] * od[((t - 1) * 2 + 0) * ns + i]))
                    / (1 - exp(-nd[(t * 2 + 0) * ns + i] * nd[((t - 1) * 2 + 0) * ns + i])));
          dA += log((1 - exp(-od[(t * 2 + 1) * ns + i] * od[((t - 1) * 2 + 1) * ns + i]))
                    / (1 - exp(-nd[(t * 2 + 1) * ns + i] * nd[((t - 1) * 2 + 1) * ns + i])));
        }
      }
    }
    total += dA;
  }
  return total;
}

void oq_fn_state(const oq_chain* ch, int inode, int which, double* det, double* dist) {
  const fn_t* f = &ch->fn[inode];
  if (det) memcpy(det, which ? f->newdm : f->dm, sizeof(double) * ch->S);
  if (dist) memcpy(dist, which ? f->newdist : f->dist, sizeof(double) * ch->S * 2 * f->ns);
}

/* FixedNodeAction::acceptLastMove (src/fixednode/FixedNodeAction.cc:359-367) */
void oq_fn_accept(oq_chain* ch) {
  for (int a = 0; a < ch->nfn; ++a) {
    fn_t* f = &ch->fn[a];
    if (f->not_my_species) continue;
    memcpy(f->dm, f->newdm, sizeof(double) * ch->S);
    memcpy(f->dist, f->newdist, sizeof(double) * ch->S * 2 * f->ns);
    if (f->upd) upd_accept(ch, f);
  }
}

/* ------------------------------------------------------------- movers */

/* FreeMover::makeMove (src/advancer/mover/FreeMover.cc:59-124) */
static double free_mover(oq_chain* ch, const double* sec, double* mov, const int* idx,
                         int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  double g[PQ_MAX_MOVING * 3];
  double lnT = 0;
  ch->fwd = 0;
  for (int t = s; t < ch->S - s; t += 2 * s) {
    gauss_rand(&ch->rng, g, nmov * D);
    for (int m = 0; m < nmov; ++m) {
      int i = idx[m], sp = ch->spec[i];
      double sigma = ch->fac * sqrt(ch->lambda[i] * ch->tau * s);
      double inv2Sigma2 = 0.5 / (sigma * sigma);
      double mid[3], delta[3];
      for (int d = 0; d < D; ++d) mid[d] = MOV(mov, t + s, m, d) - MOV(mov, t - s, m, d);
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += MOV(mov, t - s, m, d); }
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) delta[d] = g[m * D + d] * sigma;
      pbc(&ch->cell, delta);
      double nb[3];
      for (int d = 0; d < D; ++d) nb[d] = mid[d] + delta[d];
      pbc(&ch->cell, nb);
      for (int d = 0; d < D; ++d) MOV(mov, t, m, d) = nb[d];
      double temp = 1.0;
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->mpg_on[k]) temp *= pg_eval(&ch->mpg[k], delta[d]);
        else {
          double tmp = delta[d] * delta[d] * inv2Sigma2;
          ch->fwd += tmp;
          lnT += tmp;
        }
      }
      temp = 1.0 / temp;
      ch->fwd += log(temp);
      for (int d = 0; d < D; ++d) mid[d] = SEC(sec, t + s, i, d) - SEC(sec, t - s, i, d);
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += SEC(sec, t - s, i, d); }
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) delta[d] = SEC(sec, t, i, d) - mid[d];
      pbc(&ch->cell, delta);
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->mpg_on[k]) temp *= pg_eval(&ch->mpg[k], delta[d]);
        else lnT -= delta[d] * delta[d] * inv2Sigma2;
      }
      lnT += log(temp);
    }
  }
  return lnT;
}

/* FreeMoverPBC::makeMove (src/advancer/mover/FreeMoverPBC.cc:59-162) */
static double free_mover_pbc(oq_chain* ch, const double* sec, double* mov, const int* idx,
                             int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  double g[PQ_MAX_MOVING * 3];
  double lnT = 0.;
  for (int t = s; t < ch->S - s; t += 2 * s) {
    gauss_rand(&ch->rng, g, nmov * D);
    for (int m = 0; m < nmov; ++m) {
      int i = idx[m], sp = ch->spec[i];
      double sigma = ch->fac * sqrt(ch->lambda[i] * ch->tau * s);
      double inv2Sigma2 = 0.5 / (sigma * sigma);
      double mid[3], mid2[3];
      for (int d = 0; d < D; ++d) mid[d] = MOV(mov, t + s, m, d) - MOV(mov, t - s, m, d);
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) mid[d] *= 0.5;
      for (int d = 0; d < D; ++d) mid2[d] = mid[d] + 0.5 * ch->cell.a[d];
      pbc(&ch->cell, mid2);
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        double w1, w2;
        if (ch->mpg_on[k]) { w1 = pg_eval(&ch->mpg[k], mid[d]); w2 = pg_eval(&ch->mpg[k], mid2[d]); }
        else { w1 = exp(-inv2Sigma2 * mid[d] * mid[d]); w2 = exp(-inv2Sigma2 * mid2[d] * mid2[d]); }
        if (rng_next(&ch->rng) > w1 / (w1 + w2)) mid[d] = mid2[d];
      }
      for (int d = 0; d < D; ++d) mid[d] += MOV(mov, t - s, m, d);
      double nb[3];
      for (int d = 0; d < D; ++d) nb[d] = mid[d] + g[m * D + d] * sigma;
      pbc(&ch->cell, nb);
      for (int d = 0; d < D; ++d) MOV(mov, t, m, d) = nb[d];
      double dpn[3], dnn[3], dnew[3], dpo[3], dno[3], dold[3];
      for (int d = 0; d < D; ++d) {
        dpn[d] = MOV(mov, t, m, d) - MOV(mov, t - s, m, d);
        dnn[d] = MOV(mov, t, m, d) - MOV(mov, t + s, m, d);
        dnew[d] = MOV(mov, t + s, m, d) - MOV(mov, t - s, m, d);
        dpo[d] = SEC(sec, t, i, d) - SEC(sec, t - s, i, d);
        dno[d] = SEC(sec, t, i, d) - SEC(sec, t + s, i, d);
        dold[d] = SEC(sec, t + s, i, d) - SEC(sec, t - s, i, d);
      }
      pbc(&ch->cell, dpn); pbc(&ch->cell, dnn); pbc(&ch->cell, dnew);
      pbc(&ch->cell, dpo); pbc(&ch->cell, dno); pbc(&ch->cell, dold);
      double temp = 1.;
      for (int d = 0; d < D; ++d) {
        int k1 = pg_index(ch, level + 1, sp, d), k2 = pg_index(ch, level + 2, sp, d);
        if (level + 1 < ch->npg_levels && ch->mpg_on[k1]) {
          const pg_t* p = &ch->mpg[k1];
          temp *= pg_eval(p, dpo[d]); temp *= pg_eval(p, dno[d]);
          temp /= pg_eval(p, dpn[d]); temp /= pg_eval(p, dnn[d]);
        } else {
          lnT -= dpo[d] * dpo[d] * inv2Sigma2 * 0.5;
          lnT -= dno[d] * dno[d] * inv2Sigma2 * 0.5;
          lnT += dpn[d] * dpn[d] * inv2Sigma2 * 0.5;
          lnT += dnn[d] * dnn[d] * inv2Sigma2 * 0.5;
        }
        if (level + 2 < ch->npg_levels && ch->mpg_on[k2]) {
          const pg_t* p = &ch->mpg[k2];
          temp /= pg_eval(p, dold[d]); temp *= pg_eval(p, dnew[d]);
        } else {
          lnT += dold[d] * dold[d] * inv2Sigma2 * 0.25;
          lnT -= dnew[d] * dnew[d] * inv2Sigma2 * 0.25;
        }
      }
      lnT += log(temp);
    }
  }
  return lnT;
}

double oq_mover(oq_chain* ch, const double* sec, double* mov, const int* idx,
                int nmov, int level) {
  if (ch->d.mover == PQ_MOVER_FREE) return free_mover(ch, sec, mov, idx, nmov, level);
  return free_mover_pbc(ch, sec, mov, idx, nmov, level);
}

/* FreeMover::makeDelayedMove (src/advancer/mover/FreeMover.cc:126-184): for
   every slice of the level, ln(T_old/T_new) of proposing the rejected beads
   rej ([S][nmov][D]) from the neighbours in mov, against mov's beads from
   rej's neighbours; forwardProb takes minus the rejected beads' Gaussian
   exponents plus the log of their periodic-Gaussian weights.  No draws. */
static double free_delayed_mover(oq_chain* ch, const double* rej, const double* mov,
                                 const int* idx, int nmov, int level) {
  const int s = 1 << level, D = ch->D;
  double lnT = 0;
  ch->fwd = 0;
  for (int t = s; t < ch->S - s; t += 2 * s) {
    for (int m = 0; m < nmov; ++m) {
      int i = idx[m], sp = ch->spec[i];
      double sigma = ch->fac * sqrt(ch->lambda[i] * ch->tau * s);
      double inv2Sigma2 = 0.5 / (sigma * sigma);
      double temp = 1.0;
      double mid[3], delta[3];
      for (int d = 0; d < D; ++d) mid[d] = MOV(mov, t + s, m, d) - MOV(mov, t - s, m, d);
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += MOV(mov, t - s, m, d); }
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) delta[d] = MOV(rej, t, m, d) - mid[d];
      pbc(&ch->cell, delta);
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->mpg_on[k]) temp *= pg_eval(&ch->mpg[k], delta[d]);
        else {
          double tmp = delta[d] * delta[d] * inv2Sigma2;
          ch->fwd -= tmp;
          lnT += tmp;
        }
      }
      ch->fwd += log(temp);
      temp = 1.0 / temp;
      /* reverse: mov's bead about the rejected beads' midpoint */
      for (int d = 0; d < D; ++d) mid[d] = MOV(rej, t + s, m, d) - MOV(rej, t - s, m, d);
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) { mid[d] *= 0.5; mid[d] += MOV(rej, t - s, m, d); }
      pbc(&ch->cell, mid);
      for (int d = 0; d < D; ++d) delta[d] = MOV(mov, t, m, d) - mid[d];
      pbc(&ch->cell, delta);
      for (int d = 0; d < D; ++d) {
        int k = pg_index(ch, level, sp, d);
        if (ch->mpg_on[k]) temp *= pg_eval(&ch->mpg[k], delta[d]);
        else lnT -= delta[d] * delta[d] * inv2Sigma2;
      }
      lnT += log(temp);
    }
  }
  return lnT;
}

/* ------------------------------------------------------------ sampler */

/* SerialPaths::getBeads (src/base/SerialPaths.cc:119-132): slices before 0
   are read through inversePermutation, slices from M on through permutation */
static void get_beads(oq_chain* ch, int t0, double* sec) {
  const int N = ch->N, D = ch->D, M = ch->M;
  for (int k = 0; k < ch->S; ++k) {
    int t = k + t0;
    const int* map = NULL;
    if (t < 0) { t += M; map = ch->invperm; }
    else if (t >= M) { t -= M; map = ch->perm; }
    for (int i = 0; i < N; ++i) {
      const int src = map ? map[i] : i;
      memcpy(&sec[((size_t)k * N + i) * D], &ch->beads[((size_t)t * N + src) * D], sizeof(double) * D);
    }
  }
}

static int is_identity(const int* p, int n) {
  for (int i = 0; i < n; ++i)
    if (p[i] != i) return 0;
  return 1;
}

/* Permutation::prepend / append (src/util/Permutation.cc:23-43):
   prepend(q): p(i) <- p(q(i));  append(q): p(i) <- q(p(i)) */
static void perm_prepend(int* p, const int* q, int n) {
  int tmp[MAXN];
  for (int i = 0; i < n; ++i) tmp[i] = p[q[i]];
  memcpy(p, tmp, sizeof(int) * n);
}
static void perm_append(int* p, const int* q, int n) {
  int tmp[MAXN];
  for (int i = 0; i < n; ++i) tmp[i] = q[p[i]];
  memcpy(p, tmp, sizeof(int) * n);
}

void oq_perm_compose(int* p, const int* q, int n, int append) {
  if (append) perm_append(p, q, n);
  else perm_prepend(p, q, n);
}

/* SerialPaths::putBeads (src/base/SerialPaths.cc:144-185): the section's
   slices go back (before 0: skipped, from M on: to particle permutation[i]);
   then, for a non-identity section permutation sp, the slices after the
   section are relabelled, bead(i, t) <- bead(temp[i], t), temp = sp, or
   permutation o sp o inversePermutation when the section ran past M, and the
   path permutation becomes permutation o temp. */
static void put_beads(oq_chain* ch, int t0, const double* sec, const int* sp) {
  const int N = ch->N, D = ch->D, M = ch->M, S = ch->S;
  for (int k = 0; k < S; ++k) {
    int t = k + t0;
    if (t < 0) continue;
    if (t < M) {
      memcpy(&ch->beads[(size_t)t * N * D], &sec[(size_t)k * N * D], sizeof(double) * N * D);
    } else {
      t %= M;
      for (int i = 0; i < N; ++i)
        memcpy(&ch->beads[((size_t)t * N + ch->perm[i]) * D], &sec[((size_t)k * N + i) * D], sizeof(double) * D);
    }
  }
  if (is_identity(sp, N)) return;
  int ifollowing = t0 + S;
  int temp[MAXN];
  memcpy(temp, sp, sizeof(int) * N);
  if (ifollowing > M) {
    ifollowing -= M;
    perm_prepend(temp, ch->invperm, N);
    perm_append(temp, ch->perm, N);
  }
  double* buf = (double*)malloc(sizeof(double) * N * D);
  for (int t = ifollowing; t < M; ++t) {
    double* sl = &ch->beads[(size_t)t * N * D];
    for (int i = 0; i < N; ++i) memcpy(&buf[i * D], &sl[temp[i] * D], sizeof(double) * D);
    memcpy(sl, buf, sizeof(double) * N * D);
  }
  free(buf);
  perm_prepend(ch->perm, temp, N);
  for (int i = 0; i < N; ++i) ch->invperm[ch->perm[i]] = i;
}

/* SerialPaths::putDoubleBeads (src/base/SerialPaths.cc:187-206) with an
   unsampled section 2 (DoubleMLSampler without "both": its permutation p2
   stays the identity, so p1 p2 p1^-1 is the identity and section 2's beads
   are relabelled by p1 when section 1 goes first). */
static void put_double_beads(oq_chain* ch, int t0, int t0b) {
  const int N = ch->N, D = ch->D, S = ch->S;
  int ident[MAXN];
  for (int i = 0; i < N; ++i) ident[i] = i;
  if (t0 < ch->M / 2 - 1) {
    put_beads(ch, t0, ch->sec1, ch->secperm);
    if (!is_identity(ch->secperm, N)) {
      /* Beads::permute (src/base/Beads.h:254-261): coord(i,t) <- coord(p[i],t) */
      double* buf = (double*)malloc(sizeof(double) * N * D);
      for (int k = 0; k < S; ++k) {
        double* sl = &ch->sec2[(size_t)k * N * D];
        for (int i = 0; i < N; ++i) memcpy(&buf[i * D], &sl[ch->secperm[i] * D], sizeof(double) * D);
        memcpy(sl, buf, sizeof(double) * N * D);
      }
      free(buf);
      put_beads(ch, t0b, ch->sec2, ident);
    }
  } else {
    /* section 2 is unchanged and its permutation the identity: putting it
       back first rewrites the same values */
    put_beads(ch, t0, ch->sec1, ch->secperm);
  }
}

/* SpeciesParticleChooser::chooseParticles (src/advancer/SpeciesParticleChooser.cc:19-31) */
static void choose_particles(oq_chain* ch, int sp, int nmov, int* idx) {
  int npart = ch->d.species[sp].count, ifirst = ch->ifirst[sp];
  for (int im = 0; im < nmov;) {
    int i = (int)(npart * rng_next(&ch->rng) + ifirst);
    if (i == npart + ifirst) i = npart + ifirst - 1;
    for (int jm = 0; jm <= im; ++jm) {
      if (jm == im) { idx[im++] = i; break; }
      if (idx[jm] == i) break;
    }
  }
}

/* SimpleParticleChooser::chooseParticles (src/advancer/SimpleParticleChooser.cc:15-27) */
static void choose_any(oq_chain* ch, int nmov, int* idx) {
  const int N = ch->N;
  for (int im = 0; im < nmov;) {
    int i = (int)(N * rng_next(&ch->rng));
    if (i == N) continue;
    for (int jm = 0; jm <= im; ++jm) {
      if (jm == im) { idx[im++] = i; break; }
      if (idx[jm] == i) break;
    }
  }
}

/* RandomPermutationChooser::choosePermutation (src/advancer/RandomPermutationChooser.cc:12-24) */
static void random_permutation(oq_chain* ch, int nsize, int* p) {
  for (int ifrom = 0; ifrom < nsize;) {
    int ito = (int)(nsize * rng_next(&ch->rng));
    if (ito == nsize) ito = nsize - 1;
    for (int jfrom = 0; jfrom <= ifrom; ++jfrom) {
      if (jfrom == ifrom) { p[ifrom++] = ito; break; }
      if (p[jfrom] == ito) break;
    }
  }
}

/* row i of WalkingChooser::init's tables (src/advancer/WalkingChooser.cc:89-118):
   t(i,j) = prod_d PG(pbc(r_i(first) - r_j(last))_d), cump(i,j) = sum_{k<=j} t(i,k) */
static void walk_row(oq_chain* ch, int sp, int i) {
  const int npart = ch->d.species[sp].count, ifirst = ch->ifirst[sp], D = ch->D;
  if (ch->wrow[i]) return;
  double* t = &ch->wt[(size_t)i * npart];
  double* cump = &ch->wcump[(size_t)i * npart];
  for (int j = 0; j < npart; ++j) {
    double delta[3];
    for (int d = 0; d < D; ++d)
      delta[d] = SEC(ch->sec1, 0, i + ifirst, d) - SEC(ch->sec1, ch->S - 1, j + ifirst, d);
    pbc(&ch->cell, delta);
    t[j] = 1;
    for (int d = 0; d < D; ++d) t[j] *= pg_eval(&ch->wpg[sp][d], delta[d]);
  }
  cump[0] = t[0];
  for (int j = 1; j < npart; ++j) cump[j] = cump[j - 1] + t[j];
  ch->wrow[i] = 1;
}

/* WalkingChooser::iSearch (src/advancer/WalkingChooser.cc:120-142) */
static int walk_search(const double* cump, int npart, double x) {
  if (x < cump[0]) return 0;
  if (cump[npart - 2] < x) return npart - 1;
  int ilo = 0, ihi = npart - 2, mid = 0;
  for (int k = 0; k < npart - 1; k++) {
    if (ihi == ilo + 1) return ihi;
    mid = (ilo + ihi) / 2;
    if (cump[mid] > x) ihi = mid;
    else ilo = mid;
  }
  return mid;
}

/* WalkingChooser::choosePermutation (src/advancer/WalkingChooser.cc:39-87):
   a walk of nsize distinct particles, each next one drawn with probability
   t(prev, j)/h(prev); the moving particles form the cycle m -> m+1 (mod
   nsize) set in the constructor (:26-28).  lnTranProb = log prod_m
   t(i_m,i_m)/t(i_m,i_m+1); the move goes ahead iff a last uniform is below
   the ratio of reverse to forward selection probabilities. */
static int walk_choose(oq_chain* ch, int sp, int nsize, int* idx, double* lnTP) {
  const int npart = ch->d.species[sp].count, ifirst = ch->ifirst[sp];
  memset(ch->wrow, 0, npart);   /* init(): a new table per Sample */
  double tranProb = 0, revTranProb = 0;
  for (int ip = 0; ip < nsize; ++ip) {
    if (ip == 0) {
      idx[0] = (int)(npart * rng_next(&ch->rng));
      if (idx[0] == npart) idx[0] = npart - 1;
    } else {
      const int prev = idx[ip - 1];
      walk_row(ch, sp, prev);
      const double* t = &ch->wt[(size_t)prev * npart];
      const double h = ch->wcump[(size_t)prev * npart + npart - 1];
      const double x = h * rng_next(&ch->rng);
      idx[ip] = walk_search(&ch->wcump[(size_t)prev * npart], npart, x);
      for (int jp = 0; jp < ip; ++jp)
        if (idx[ip] == idx[jp]) return 0;
      revTranProb += h / t[prev];
      tranProb += h / t[idx[ip]];
    }
  }
  const int last = idx[nsize - 1];
  walk_row(ch, sp, last);
  const double h = ch->wcump[(size_t)last * npart + npart - 1];
  revTranProb += h / ch->wt[(size_t)last * npart + last];
  tranProb += h / ch->wt[(size_t)last * npart + idx[0]];
  const double accept = revTranProb / tranProb;
  double prob = 1;
  for (int m = 0; m < nsize; ++m) {
    walk_row(ch, sp, idx[m]);
    const double* t = &ch->wt[(size_t)idx[m] * npart];
    prob *= t[idx[m]] / t[idx[(m + 1) % nsize]];
  }
  *lnTP = log(prob);
  for (int m = 0; m < nsize; ++m) idx[m] += ifirst;
  return rng_next(&ch->rng) < accept;
}

/* Permutation and particle choice of a Sample, in the order of
   MultiLevelSampler::run (src/advancer/MultiLevelSampler.cc:68-80) and
   DoubleMLSampler::run (src/advancer/DoubleMLSampler.cc:76-81):
   choosePermutation, then chooseParticles.  Fills idx (the moving particles),
   cyc (the permutation: moving particle m ends on idx[cyc[m]]'s last bead)
   and the log transition probability; returns 0 when the permutation chooser
   turns the move down (isNewPerm false).
   - SPECIES: PermutationChooser (identity, no draw,
     src/advancer/PermutationChooser.cc:13-15), then SpeciesParticleChooser.
   - WALKING: WalkingChooser (walk_choose); its chooseParticles draws nothing
     (src/advancer/WalkingChooser.cc:36-37).
   - ALL / ALL_RANDOM: PermutationChooser, or RandomPermutationChooser of
     nmoving elements, then SimpleParticleChooser over all particles. */
static int choose(oq_chain* ch, const pq_sample_desc* s, int* idx, int* cyc, double* lnTP) {
  const int kind = s->chooser, nmov = s->nmoving;
  *lnTP = 0;
  for (int m = 0; m < nmov; ++m) cyc[m] = m;
  if (kind == PQ_CHOOSER_WALKING) {
    for (int m = 0; m < nmov; ++m) cyc[m] = (m + 1) % nmov;
    return walk_choose(ch, s->species, nmov, idx, lnTP);
  }
  if (kind == PQ_CHOOSER_ALL || kind == PQ_CHOOSER_ALL_RANDOM) {
    if (kind == PQ_CHOOSER_ALL_RANDOM) random_permutation(ch, nmov, cyc);
    choose_any(ch, nmov, idx);
    return 1;
  }
  choose_particles(ch, s->species, nmov, idx);
  return 1;
}

/* MultiLevelSampler::run/tryMove (src/advancer/MultiLevelSampler.cc:68-177) and
   DoubleMLSampler::run/tryMove (src/advancer/DoubleMLSampler.cc:66-245), with
   delayed rejection at level L-1 when the Sample asks for it (Free mover);
   section 2 is not sampled.  The moving beads start
   from the chosen particles' first beads and end on the permuted particles'
   last beads; an accepted move writes them back (last slice included) and
   appends the permutation to the section permutation. */
static int sample_once(oq_chain* ch, const pq_sample_desc* smp) {
  const int D = ch->D, S = ch->S, L = ch->L;
  const int nmov = smp->nmoving;
  int* idx = ch->idx;
  int* cyc = ch->cyc;
  double lnTP;
  if (!choose(ch, smp, idx, cyc, &lnTP)) return 0;
  double* mov = ch->mov;
  for (int m = 0; m < nmov; ++m)
    for (int d = 0; d < D; ++d) {
      MOV(mov, 0, m, d) = SEC(ch->sec1, 0, idx[m], d);
      MOV(mov, S - 1, m, d) = SEC(ch->sec1, S - 1, idx[cyc[m]], d);
    }
  const int dr = smp->delayed_rejection != 0;
  const double fac = smp->sigma_factor == 0 ? 1.0 : smp->sigma_factor;
  const double nfac = smp->new_sigma_factor == 0 ? 0.75 : smp->new_sigma_factor;
  double old = lnTP;
  for (int l = L; l >= 0; --l) {
    ch->st.tries[l]++;
    if (l < L) ch->st.bead_updates += (uint64_t)nmov << (L - 1 - l);
    ch->fac = fac;
    double lnT = oq_mover(ch, ch->sec1, mov, idx, nmov, l);
    double dA = oq_action_difference(ch, ch->sec1, mov, idx, nmov, l, 7);
    if (ch->nfn) dA += oq_fn_action_difference(ch, mov, idx, nmov, l);
    double piRatio = -dA + old;
    double acceptProb = exp(lnT + piRatio);
    if (dr && l == L - 1) {
      if (rng_next(&ch->rng) > acceptProb) {
        /* MultiLevelSampler.cc:116-143 / DoubleMLSampler.cc:150-206 */
        double accRatioA2B = acceptProb;
        double transA2B = ch->fwd;
        memcpy(ch->rej, mov, sizeof(double) * S * nmov * D);
        ch->fac = nfac;
        double lnTC = oq_mover(ch, ch->sec1, mov, idx, nmov, l);
        ch->st.bead_updates += (uint64_t)nmov << (L - 1 - l);
        dA = oq_action_difference(ch, ch->sec1, mov, idx, nmov, l, 7);
        if (ch->nfn) dA += oq_fn_action_difference(ch, mov, idx, nmov, l);
        double piRatioCoA = -dA + old;
        double accRatioA2C = exp(lnTC + piRatioCoA);
        ch->fac = fac;
        double lnTC2B = free_delayed_mover(ch, ch->rej, mov, idx, nmov, l);
        double transC2B = ch->fwd;
        double accRatioC2B = exp(lnTC2B - piRatioCoA + piRatio);
        accRatioC2B = (accRatioC2B >= 1) ? 1 : accRatioC2B;
        double accRatioA2B2C = accRatioA2C * exp(transC2B + transA2B)
                               * (1 - accRatioC2B) / (1 - accRatioA2B);
        dA = dA - transC2B - transA2B - log((1 - accRatioC2B) / (1 - accRatioA2B));
        if (rng_next(&ch->rng) > accRatioA2B2C) return 0;
      }
    } else if (rng_next(&ch->rng) > acceptProb) return 0;
    old = dA;
    ch->st.accepts[l]++;
  }
  if (ch->nfn) oq_fn_accept(ch);
  for (int t = 0; t < S; ++t)
    for (int m = 0; m < nmov; ++m)
      for (int d = 0; d < D; ++d) SEC(ch->sec1, t, idx[m], d) = MOV(mov, t, m, d);
  int temp[PQ_MAX_MOVING];
  for (int m = 0; m < nmov; ++m) temp[m] = ch->secperm[idx[m]];
  for (int m = 0; m < nmov; ++m) ch->secperm[idx[m]] = temp[cyc[m]];
  ch->st.accepted_bead_updates += (uint64_t)nmov * ((1u << L) - 1);
  return 1;
}

/* MultiLevelSampler::run / DoubleMLSampler::run (MultiLevelSampler.cc:68-98,
   DoubleMLSampler.cc:66-120): nrepeat tries on one section; the chooser's
   tables are rebuilt after an accepted try (init()), which the per-try
   tables of walk_choose equal, the section being unchanged otherwise */
static void sample(oq_chain* ch, const pq_sample_desc* smp) {
  const int nrep = smp->nrepeat > 0 ? smp->nrepeat : 1;
  for (int r = 0; r < nrep; ++r) sample_once(ch, smp);
}

/* SectionChooser::run (src/advancer/SectionChooser.cc:30-48) and
   DoubleSectionChooser::run (src/advancer/DoubleSectionChooser.cc:33-65) */
static void run_section(oq_chain* ch) {
  const int M = ch->M;
  int t0, t0b = 0;
  for (int i = 0; i < ch->N; ++i) ch->secperm[i] = i;
  if (ch->nfn) {
    double x = rng_next(&ch->rng) * (1 - 1e-8);
    int ilo = -1, ihi = M / 2 - 1;
    t0 = ilo + (int)((ihi + 1 - ilo) * x);
    if (t0 > ihi) t0 = ihi;
    if (rng_next(&ch->rng) >= 0.5) t0 = (t0 + M / 2 + 1) % M - 1;
    t0b = (t0 + M / 2 + 1) % M - 1;
    get_beads(ch, t0, ch->sec1);
    get_beads(ch, t0b, ch->sec2);
    oq_fn_initialize(ch, ch->sec1, ch->sec2);
  } else {
    double x = rng_next(&ch->rng) * (1 - 1e-8);
    int ilo = -1, ihi = M - 1;
    t0 = ilo + (int)((ihi + 1 - ilo) * x);
    if (t0 > ihi) t0 = ihi;
    get_beads(ch, t0, ch->sec1);
  }
  for (int s = 0; s < ch->gnsample; ++s) sample(ch, &ch->gsample[s]);
  if (ch->nfn) put_double_beads(ch, t0, t0b);
  else put_beads(ch, t0, ch->sec1, ch->secperm);
}

void oq_run_sweeps(oq_chain* ch, int nsweep) {
  for (int k = 0; k < nsweep; ++k)
    for (int g = 0; g <= ch->d.ngroup; ++g) {
      set_group(ch, g);
      for (int s = 0; s < ch->gnsection; ++s) run_section(ch);
    }
  set_group(ch, 0);
  ch->st.draws = ch->rng.count;
  ch->st.sweeps += (uint64_t)nsweep;
}

void oq_get_stats(const oq_chain* ch, pq_chain_stats* st) {
  *st = ch->st;
  st->draws = ch->rng.count;
}

/* ThermoEnergyEstimator over SerialPaths::sumOverLinks
   (src/estimator/ThermoEnergyEstimator.cc:32-54, src/base/SerialPaths.cc:48-62,
   src/stats/SimpleScalarAccumulator.cpp:25-38).  FixedNodeAction::getBeadAction
   adds utau = 0 when useDistDerivative is off (FixedNodeAction.cc:306-313). */
double oq_thermo_energy(const oq_chain* ch) {
  double value = 0;
  for (int t = 0; t < ch->M; ++t)
    for (int i = 0; i < ch->N; ++i) value += oq_bead_utau(ch, i, t, 7);
  return value / ch->M;
}

/* PairAction::uk0CalcDerivatives without z (src/action/PairAction.cc:640-657):
   the link action v, its tau derivative, dv/dq and dv/ds2 */
static void uk0_derivs(const pair_t* p, double q, double s2, double* u, double* ut, double* uq, double* us2) {
  q *= p->rgridinv;
  double x = log(q) * p->logrratioinv;
  int i = (int)x;
  if (q < 1) { i = 0; x = 0; }
  else if (i > p->ngpts - 2) { i = p->ngpts - 2; x = 1; }
  else x -= i;
  const int K1 = p->norder + 1;
  double a = 0, b = 0, c = 0, e = 0;
  for (int k = p->norder; k >= 0; k--) {
    a *= s2; a += (1 - x) * p->u[i * K1 + k] + x * p->u[(i + 1) * K1 + k];
    b *= s2; b += (1 - x) * p->ut[i * K1 + k] + x * p->ut[(i + 1) * K1 + k];
    c *= s2; c += p->u[(i + 1) * K1 + k] - p->u[i * K1 + k];
  }
  c *= p->logrratioinv * p->rgridinv / q;
  for (int k = p->norder; k > 0; k--) {
    e *= s2; e += k * ((1 - x) * p->u[i * K1 + k] + x * p->u[(i + 1) * K1 + k]);
  }
  *u = a; *ut = b; *uq = c; *us2 = e;
}

/* CompositeAction::getBeadAction (CompositeAction.cc:57-67) over
   SpringAction (SpringAction.cc:112-123: utau and forces, no u), SHOAction
   (SHOAction.cc:71-105) and PairAction (PairAction.cc:509-599, no exchange):
   u, utau, and the forces fm (from the link into slice t) and fp (from the
   link out of it) on bead (ipart, islice). */
void oq_bead_action(const oq_chain* ch, int ipart, int islice, double* u_out, double* utau_out,
                    double* fm, double* fp) {
  const int D = ch->D;
  double u = 0, utau = 0;
  for (int d = 0; d < D; ++d) fm[d] = fp[d] = 0;
  const double* b1 = bead(ch, ipart, islice);
  const double* b0 = bead(ch, ipart, islice - 1);
  const double* b2 = bead(ch, ipart, islice + 1);
  double dm[3], dp[3];
  for (int d = 0; d < D; ++d) { dm[d] = b1[d] - b0[d]; dp[d] = b1[d] - b2[d]; }
  pbc(&ch->cell, dm);
  pbc(&ch->cell, dp);
  if (ch->d.spring && !ch->d.species[ch->spec[ipart]].is_static) {
    const double lam = ch->lambda[ipart], tau = ch->tau;
    for (int d = 0; d < D; ++d) { fm[d] -= dm[d] / (2 * lam * tau); fp[d] -= dp[d] / (2 * lam * tau); }
    utau += D / (2 * tau) - dot3(dm, dm, D) / (4.0 * lam * tau * tau);
  }
  for (int s = 0; s < ch->d.nsho; ++s) {
    const pq_sho_desc* a = &ch->d.sho[s];
    const int ifirst = ch->ifirst[a->species], np = ch->d.species[a->species].count;
    if (ipart < ifirst || ipart >= ifirst + np) continue;
    const int ndim = a->ndim ? a->ndim : D;
    const double omega = a->omega, m = ch->d.species[a->species].mass, tau = ch->tau;
    const double wt = omega * tau, sinhwt = sinh(wt), coshwt = cosh(wt);
    const double cschwt = 1.0 / sinhwt, cothwt = coshwt * cschwt, tin = 1.0 / tau;
    double r1[3], r0[3], r2[3];
    for (int d = 0; d < D; ++d) { r1[d] = b1[d] - a->center[d]; r0[d] = b0[d] - a->center[d]; r2[d] = b2[d] - a->center[d]; }
    double r0r0 = 0, r0r1 = 0, r1r1 = 0, delta2 = 0;
    for (int d = D - ndim; d < D; ++d) {
      r0r0 += r0[d] * r0[d]; r0r1 += r0[d] * r1[d]; r1r1 += r1[d] * r1[d]; delta2 += dm[d] * dm[d];
    }
    for (int d = D - ndim; d < D; ++d) {
      fm[d] -= m * omega * (r1[d] * coshwt - r0[d]) * cschwt - m * dm[d] * tin;
      fp[d] -= m * omega * (r1[d] * coshwt - r2[d]) * cschwt - m * dp[d] * tin;
    }
    utau += 0.5 * ndim * omega * cothwt
            + 0.5 * m * omega * omega * ((r1r1 + r0r0) * (1 - cothwt * cothwt) + 2.0 * r0r1 * coshwt * cschwt * cschwt)
            + 0.5 * m * delta2 * tin * tin - 0.5 * ndim * tin;
    u += 0.5 * ndim * log(2 * 3.14159265358979 * sinhwt / (m * omega)) +
         m * omega * ((r1r1 + r0r0) * coshwt - 2.0 * r0r1) * cschwt * 0.5 -
         0.5 * log(2.0 * 3.14159265358979 * tau / m) - 0.5 * m * delta2 / tau;
  }
  for (int pp = 0; pp < ch->npair; ++pp) {
    const pair_t* p = &ch->pair[pp];
    const int if1 = ch->ifirst[p->s1], n1 = ch->d.species[p->s1].count;
    const int if2 = ch->ifirst[p->s2], n2 = ch->d.species[p->s2].count;
    int jb, je;
    if (ipart >= if1 && ipart < if1 + n1) { jb = if2; je = if2 + n2; }
    else if (ipart >= if2 && ipart < if2 + n2) { jb = if1; je = if1 + n1; }
    else continue;
    for (int j = jb; j < je; ++j) {
      if (ipart == j) continue;
      double delta[3], prev[3], next[3], svec[3];
      const double* aj = bead(ch, j, islice);
      const double* bj = bead(ch, j, islice - 1);
      const double* cj = bead(ch, j, islice + 1);
      for (int d = 0; d < D; ++d) { delta[d] = b1[d] - aj[d]; prev[d] = b0[d] - bj[d]; next[d] = b2[d] - cj[d]; }
      pbc(&ch->cell, delta);
      pbc(&ch->cell, prev);
      const double r = sqrt(dot3(delta, delta, D));
      const double prevR = sqrt(dot3(prev, prev, D));
      double q = 0.5 * (r + prevR);
      for (int d = 0; d < D; ++d) svec[d] = delta[d] - prev[d];
      double s2 = dot3(svec, svec, D) / (q * q);
      double v, vtau, vq, vs2;
      uk0_derivs(p, q, s2, &v, &vtau, &vq, &vs2);
      u += 0.5 * v;
      utau += 0.5 * vtau;
      for (int d = 0; d < D; ++d)
        fm[d] -= vq * delta[d] / (2 * r) + vs2 * (2 * svec[d] / (q * q) - s2 * delta[d] / (q * r));
      pbc(&ch->cell, next);
      const double nextR = sqrt(dot3(next, next, D));
      q = 0.5 * (r + nextR);
      for (int d = 0; d < D; ++d) svec[d] = delta[d] - next[d];
      s2 = dot3(svec, svec, D) / (q * q);
      uk0_derivs(p, q, s2, &v, &vtau, &vq, &vs2);
      for (int d = 0; d < D; ++d)
        fp[d] -= vq * delta[d] / (2 * r) + vs2 * (2 * svec[d] / (q * q) - s2 * delta[d] / (q * r));
    }
  }
  /* CoulombAction's long-range part: u and utau of particle 0, no force
     (CoulombAction.cc:164-170) */
  if (ch->has_ewald && ipart == 0) {
    const double lr = slice_long_range(ch, islice) / ch->ceps;
    u += lr * ch->tau;
    utau += lr;
  }
  *u_out = u;
  *utau_out = utau;
}

/* VirialEnergyEstimator (src/estimator/VirialEnergyEstimator.cc:27-78) over
   SerialPaths::sumOverLinks with firstSlice 0, static particles skipped; the
   FixedNodeActions' getBeadAction gives zero forces and, without
   useDistDerivative, zero utau (FixedNodeAction.cc:254-329), so they add
   nothing.  Returns the energy of the current paths (energy / nslice). */
double oq_virial_energy(const oq_chain* ch, int nwindow) {
  const int D = ch->D, M = ch->M, N = ch->N;
  const double tau = ch->tau;
  if (nwindow < 1) nwindow = 1;
  double* r0 = (double*)calloc((size_t)N * 3, sizeof(double));
  double* rav = (double*)calloc((size_t)N * 3, sizeof(double));
  double* fav = (double*)calloc((size_t)N * 3, sizeof(double));
  double energy = 0;
  for (int t = 0; t < M; ++t)
    for (int i = 0; i < N; ++i) {
      double* R0 = &r0[i * 3]; double* RA = &rav[i * 3]; double* FA = &fav[i * 3];
      const int stat = ch->d.species[ch->spec[i]].is_static;
      if (t % nwindow == 0 && !stat) {
        energy += (1 / (2 * tau)) * dot3(FA, R0, D);
        for (int d = 0; d < 3; ++d) R0[d] = RA[d] = FA[d] = 0;
      }
      const int W = (t / nwindow < (M - 1) / nwindow) ? nwindow : (M - 1) % nwindow + 1;
      const double *e = bead(ch, i, t), *s = bead(ch, i, t - 1);
      double r[3];
      for (int d = 0; d < D; ++d) r[d] = e[d] - s[d];
      pbc(&ch->cell, r);
      for (int d = 0; d < D; ++d) {
        RA[d] += r[d];
        R0[d] += (W - t % nwindow) * r[d] / W;
      }
      double u, utau, fm[3], fp[3], f[3];
      oq_bead_action(ch, i, t, &u, &utau, fm, fp);
      for (int d = 0; d < D; ++d) { f[d] = fm[d] + fp[d]; FA[d] += f[d]; }
      energy += utau;
      if (!stat) energy += -(1. - 1. / W) * (D / (2 * tau)) - (1 / (2 * tau)) * dot3(f, RA, D);
    }
  for (int i = 0; i < N; ++i)
    if (!ch->d.species[ch->spec[i]].is_static) energy += (1 / (2 * tau)) * dot3(&fav[i * 3], &r0[i * 3], D);
  free(r0); free(rav); free(fav);
  return energy / M;
}

/* CoulombEnergyEstimator::handleLink (src/estimator/CoulombEnergyEstimator.cc:30-40)
   over SerialPaths::sumOverLinks (slice 0 first, particles in order), then
   SimpleScalarAccumulator::storeValue's division by M */
double oq_coulomb_energy(const oq_chain* ch, double epsilon) {
  const int D = ch->D;
  if (epsilon == 0) epsilon = 1;
  double value = 0;
  for (int t = 0; t < ch->M; ++t)
    for (int i = 0; i < ch->N; ++i) {
      double energy = 0.0;
      const double* end = bead(ch, i, t);
      for (int j = 0; j < i; ++j) {
        double delta[3];
        const double* rj = bead(ch, j, t);
        for (int d = 0; d < D; ++d) delta[d] = end[d] - rj[d];
        pbc(&ch->cell, delta);
        energy += ch->d.species[ch->spec[i]].charge * ch->d.species[ch->spec[j]].charge /
                  (epsilon * sqrt(dot3(delta, delta, D)));
      }
      value += energy;
    }
  return value / ch->M;
}

/* EwaldCoulombEstimator with its OptEwaldSum (src/estimator/EwaldCoulombEstimator.cc:63-140,
   CoulombEnergyEstimator useEwald="true", EstimatorParser.cc:178-190): the
   short-range part q_i q_j (1/(r eps) - f(r)/eps) inside rcut from a 1001-point
   grid of -f/eps, whose interpolation weight is r - ig dr (not divided by dr,
   as the reference has it), and the k-space sum on particle 0 of every slice */
double oq_ewald_coulomb_energy(const oq_chain* ch, double epsilon, double rcut, double kcut) {
  return oq_ewald_coulomb_energy_type(ch, epsilon, rcut, kcut, 0, 0.0);
}

/* the estimator's default kappa of its TradEwaldSum (EstimatorParser.cc:194-203):
   sqrt((3.6 N)^(1/6) sqrt(pi) / Lside), Lside the shortest cell side */
double oq_trad_estimator_kappa(const oq_chain* ch) {
  double lside = 1000000000000000.0;
  for (int d = 0; d < ch->D; ++d)
    if (lside > ch->cell.a[d]) lside = ch->cell.a[d];
  return sqrt(pow(3.6 * ch->N, 1.0 / 6.0) * sqrt(3.1415926535897931) / lside);
}

/* EwaldCoulombEstimator with an OptEwaldSum (trad 0) or, for ewaldType
   "tradEwald" with ewaldImages 1, a TradEwaldSum of kappa (0: the parser's
   default) (EwaldCoulombEstimator.cc:17-61,105-140) */
double oq_ewald_coulomb_energy_type(const oq_chain* ch, double epsilon, double rcut, double kcut,
                                    int trad, double kappa) {
  const int D = ch->D, N = ch->N;
  if (epsilon == 0) epsilon = 1;
  if (trad && kappa == 0) kappa = oq_trad_estimator_kappa(ch);
  double q[MAXN];
  for (int i = 0; i < N; ++i) q[i] = ch->d.species[ch->spec[i]].charge;
  ewald_t e;
  ewald_init(&e, &ch->cell, N, q, rcut, kcut, trad, kappa);
  const int nradial = 1001;
  const double dr = rcut / 1000, drinv = 1. / dr;
  /* one entry past the reference's 1001: r just below rcut can round to
     igrid = 1000, whose weight on vgrid(1001) is ~0 (the reference reads one
     past its array there); -f(1001 dr)/eps keeps that read in bounds */
  double* vgrid = (double*)malloc(sizeof(double) * (nradial + 1));
  vgrid[0] = -(trad ? 2 * kappa / sqrt(EW_PI) : e.coef[0]) / epsilon;   /* evalFR0 */
  for (int i = 1; i <= nradial; ++i) vgrid[i] = -ew_fr(&e, i * dr) / epsilon;
  double value = 0;
  for (int t = 0; t < ch->M; ++t)
    for (int i = 0; i < N; ++i) {
      double energy = 0.0;
      const double* end = bead(ch, i, t);
      for (int j = 0; j < i; ++j) {
        double delta[3];
        const double* rj = bead(ch, j, t);
        for (int d = 0; d < D; ++d) delta[d] = end[d] - rj[d];
        pbc(&ch->cell, delta);
        double r = sqrt(dot3(delta, delta, D));
        if (r < rcut) {
          int ig = (int)(r * drinv);
          double x = r - ig * dr;
          energy += q[i] * q[j] * (1. / (r * epsilon) + (1 - x) * vgrid[ig] + x * vgrid[ig + 1]);
        }
      }
      if (i == 0) energy += ewald_long_range(&e, &ch->beads[(size_t)t * N * D]) / epsilon;
      value += energy;
    }
  free(vgrid);
  ewald_free(&e);
  return value / ch->M;
}

/* the OptEwaldSum of a chain's CoulombAction: its fitted coefficients
   coef[8], the number of k-vectors, the self energy and, for 0 <= ik < nk,
   kv[3] and vk; returns 0 without an Ewald sum */
int oq_ewald_info(const oq_chain* ch, double* coef, int* nk, double* self, int ik, int* kv, double* vk) {
  if (!ch->has_ewald) return 0;
  if (coef) memcpy(coef, ch->ew.coef, sizeof ch->ew.coef);
  if (nk) *nk = ch->ew.nk;
  if (self) *self = ch->ew.self;
  if (ik >= 0 && ik < ch->ew.nk) {
    if (kv) for (int d = 0; d < 3; ++d) kv[d] = ch->ew.kv[3 * ik + d];
    if (vk) *vk = ch->ew.vk[ik];
  }
  return 1;
}

/* EwaldSum::evalLongRange of a configuration r [N][D] with the chain's sum */
double oq_ewald_long_range(const oq_chain* ch, const double* r) {
  return ch->has_ewald ? ewald_long_range(&ch->ew, r) : 0.0;
}

/* the f(r) of the chain's OptEwaldSum (OptEwaldSum.cc:178-190) */
double oq_ewald_fr(const oq_chain* ch, double r) {
  return ch->has_ewald ? ew_fr(&ch->ew, r) : 0.0;
}

/* PairCFEstimator<1>::handleLink with PairRadial (src/estimator/PairCFEstimator.h:72-93,
   src/util/PairDistance.h:22-39): for i of species[0] and each partner species
   species[1..nspec-1] in the order listed, every j != i of that species on the
   same slice: ibin = int((|pbc(r_i - r_j)| - rmin) * nbin/(rmax - rmin)),
   dropped when d < rmin or ibin > nbin - 1; integer counts summed over slices
   (the reference's float temp, exact below 2^24, before its /nslice) */
void oq_pair_cf_histogram(const oq_chain* ch, int nspec, const int* species, double rmin,
                          double rmax, int nbin, uint64_t* counts) {
  const int D = ch->D;
  double deltaInv = nbin / (rmax - rmin);
  if (nspec < 1) return;
  int if1 = ch->ifirst[species[0]], n1 = ch->d.species[species[0]].count;
  for (int t = 0; t < ch->M; ++t)
    for (int i = if1; i < if1 + n1; ++i)
      for (int ks = 1; ks < nspec; ++ks) {
        const int if2 = ch->ifirst[species[ks]], n2 = ch->d.species[species[ks]].count;
        for (int j = if2; j < if2 + n2; ++j) {
          if (i == j) continue;
          const double *a = bead(ch, i, t), *b = bead(ch, j, t);
          double delta[3];
          for (int d = 0; d < D; ++d) delta[d] = a[d] - b[d];
          pbc(&ch->cell, delta);
          double r2 = 0;
          for (int d = 0; d < D; ++d) r2 += delta[d] * delta[d] * 1.0;
          double dd = sqrt(r2);
          int ibin = (int)((dd - rmin) * deltaInv);
          if (dd < rmin || ibin > nbin - 1) continue;
          counts[ibin]++;
        }
      }
}

void oq_pair_histogram(const oq_chain* ch, int s1, int s2, double rmin,
                       double rmax, int nbin, uint64_t* counts) {
  const int sp[2] = {s1, s2};
  oq_pair_cf_histogram(ch, 2, sp, rmin, rmax, nbin, counts);
}

uint64_t oq_run_batch(const pq_chain_desc* descs, int nchain, double* beads,
                      int nsweep, uint64_t seed, uint32_t first_stream,
                      int nthreads, double* energy) {
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (nchain + 1));
  off[0] = 0;
  for (int c = 0; c < nchain; ++c) {
    int n = 0;
    for (int s = 0; s < descs[c].nspecies; ++s) n += descs[c].species[s].count;
    off[c + 1] = off[c] + (int64_t)descs[c].nslice * n * descs[c].ndim;
  }
  uint64_t total = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
#endif
  for (int c = 0; c < nchain; ++c) {
    oq_chain* ch = oq_create(&descs[c]);
    oq_set_beads(ch, beads + off[c]);
    oq_rng_philox(ch, seed, first_stream + (uint32_t)c);
    oq_run_sweeps(ch, nsweep);
    total += ch->st.bead_updates;
    if (energy) energy[c] = oq_thermo_energy(ch);
    oq_get_beads(ch, beads + off[c]);
    oq_destroy(ch);
  }
  free(off);
  return total;
}

/* oq_run_batch with a Measure after every sweep: the thermal energy (summed
   into energy_sum[c]) and the pair-correlation counts of species[0] against
   species[1..nspec-1] (added to counts[c][nbin]) */
uint64_t oq_run_batch_measured(const pq_chain_desc* descs, int nchain, double* beads, int nsweep,
                               uint64_t seed, uint32_t first_stream, int nthreads, int nspec,
                               const int* species, double rmin, double rmax, int nbin,
                               double* energy_sum, uint64_t* counts) {
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (nchain + 1));
  off[0] = 0;
  for (int c = 0; c < nchain; ++c) {
    int n = 0;
    for (int s = 0; s < descs[c].nspecies; ++s) n += descs[c].species[s].count;
    off[c + 1] = off[c] + (int64_t)descs[c].nslice * n * descs[c].ndim;
  }
  uint64_t total = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
#endif
  for (int c = 0; c < nchain; ++c) {
    oq_chain* ch = oq_create(&descs[c]);
    oq_set_beads(ch, beads + off[c]);
    oq_rng_philox(ch, seed, first_stream + (uint32_t)c);
    double e = 0;
    for (int k = 0; k < nsweep; ++k) {
      oq_run_sweeps(ch, 1);
      e += oq_thermo_energy(ch);
      if (nspec > 0 && counts) oq_pair_cf_histogram(ch, nspec, species, rmin, rmax, nbin, counts + (size_t)c * nbin);
    }
    total += ch->st.bead_updates;
    if (energy_sum) energy_sum[c] = e;
    oq_get_beads(ch, beads + off[c]);
    oq_destroy(ch);
  }
  free(off);
  return total;
}

/* Loop/Measure/Collect: nblock blocks of nmeasure (sweep, Measure) pairs; the
   block value is the mean of its measurements (SimpleScalarAccumulator). */
void oq_energy_blocks(oq_chain* ch, int nblock, int nmeasure, double* block) {
  for (int b = 0; b < nblock; ++b) {
    double sum = 0;
    for (int k = 0; k < nmeasure; ++k) {
      oq_run_sweeps(ch, 1);
      sum += oq_thermo_energy(ch);
    }
    block[b] = sum / nmeasure;
  }
}
