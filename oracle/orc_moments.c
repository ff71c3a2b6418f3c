/*
 * orc_moments.c -- restatement of the reference moment model (TEST
 * INFRASTRUCTURE ONLY).
 *
 *   get_rate / size_scaling ..... scripts/model.jl:1-27
 *   get_vary_map ................ scripts/model.jl:30-43
 *   labelling ................... scripts/model.jl:58-64
 *   momentodes / f .............. scripts/model.jl:67-86
 *   model (CVODE_BDF solve) ..... scripts/model.jl:89-96 -> adaptive
 *                                 Dormand-Prince 5(4) at tight tolerance,
 *                                 split at every rate discontinuity; method 1
 *                                 solves the stiff gene block in closed form
 *   periodic_boundary ........... scripts/model.jl:98-111
 *   transient_phase ............. scripts/model.jl:114-142
 *   trajectories ................ scripts/model.jl:146-174
 *   syntheticdata ............... scripts/model.jl:176-187
 *   run_part_sim ................ scripts/recover_statistics.jl:2-11
 *   downsample .................. scripts/model.jl:221-239
 *   abc_sim statistics .......... scripts/abc_simulation.jl:23-46
 *   weighted_cov ................ scripts/data_summary_statistics.jl:179-181
 *
 * Powers use the engine's own exp10 / exp / log (as the GPU moment kernels
 * do), so the two integrate with identical operations and step sequences.
 *
 * The reference solves with SUNDIALS CVODE at default tolerances
 * (reltol 1e-3, abstol 1e-6, third-party, unpinned version); this restatement
 * integrates to rtol ~1e-10, so fixture comparisons carry CVODE's error.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Julia mod(x, y) for floats: floored remainder */
static double jmod(double x, double y) {
  double r = fmod(x, y);
  if (r != 0.0 && ((r < 0.0) != (y < 0.0))) r += y;
  return r;
}

typedef struct { int n_kon, n_koff, n_al, n_ga; int o[4]; int o_lam; } vmap_t;

static vmap_t vmap_of(int model, int n_steps) {
  vmap_t v;
  int flag[4] = {0, 0, 0, 0};
  if (model == 3) flag[0] = 1;
  if (model == 4) flag[2] = 1;
  if (model == 5) flag[3] = 1;
  int k = 0, cnt[4];
  for (int i = 0; i < 4; ++i) {
    v.o[i] = k;
    cnt[i] = flag[i] ? n_steps : 1;
    k += cnt[i];
  }
  v.n_kon = cnt[0]; v.n_koff = cnt[1]; v.n_al = cnt[2]; v.n_ga = cnt[3];
  v.o_lam = k;
  return v;
}

/* rates p[4] (kon, koff, alpha, gamma) at time t; the bin and the window are
 * taken at tb (a point strictly inside the current smooth piece). */
static void rates_at(int model, const double* th, double cycle, int n_steps, double t,
                     double tb, double* p) {
  vmap_t v = vmap_of(model, n_steps);
  int scaling = model != 2;
  double m = jmod(tb, cycle);
  int bin = 0;
  if (m != 0.0) {
    double h = cycle / n_steps;
    bin = (int)ceil(m / h) - 1;
    if (bin < 0) bin = 0;
    if (bin >= n_steps) bin = n_steps - 1;
  }
  int cnt[4] = {v.n_kon, v.n_koff, v.n_al, v.n_ga};
  for (int i = 0; i < 4; ++i) {
    double lp = th[v.o[i] + (cnt[i] > 1 ? bin : 0)];
    double val = orc_exp10(lp);
    if (i == 2 && scaling) {
      /* size_scaling: 1 + mod(t,T)/T inside the cycle that holds tb */
      double cs = floor(tb / cycle) * cycle;
      double a = t - cs;
      val = val * (1.0 + a / cycle);
    }
    p[i] = val;
  }
}

static void rhs(const double* p, double lam, const double* y, double* dy) {
  dy[0] = p[0] * (1 - y[0]) - p[1] * y[0];
  dy[1] = (1 - lam) * p[2] * y[0] - p[3] * y[1];
  dy[2] = lam * p[2] * y[0] - p[3] * y[2];
  dy[3] = p[1] * y[0] + p[0] * (1 - y[0]) - 2 * (p[0] + p[1]) * y[3];
  dy[4] = p[2] * (1 - lam) * y[3] - (p[0] + p[1] + p[3]) * y[4];
  dy[5] = p[2] * lam * y[3] - (p[0] + p[1] + p[3]) * y[5];
  dy[6] = p[2] * (1 - lam) * y[0] + p[3] * y[1] + 2 * p[2] * (1 - lam) * y[4] - 2 * p[3] * y[6];
  dy[7] = p[2] * lam * y[4] + p[2] * (1 - lam) * y[5] - 2 * p[3] * y[7];
  dy[8] = p[2] * lam * y[0] + p[3] * y[2] + 2 * p[2] * lam * y[5] - 2 * p[3] * y[8];
}

static double lam_at(const double* th, int model, int n_steps, double texp, double pulse,
                     double tb) {
  vmap_t v = vmap_of(model, n_steps);
  if (tb >= texp && tb <= texp + pulse) return orc_exp10(th[v.o_lam]);
  return 0.0;
}

void orc_moment_rhs(int model, const double* theta, double cycle, int n_steps, double texp,
                    double pulse, double t, const double* y, double* dy) {
  double p[4];
  rates_at(model, theta, cycle, n_steps, t, t, p);
  rhs(p, lam_at(theta, model, n_steps, texp, pulse, t), y, dy);
}

/* Dormand-Prince 5(4) on one smooth piece [a,b]. */
static int64_t dopri_piece(int model, const double* th, double cycle, int n_steps,
                           double lam, double tb, double a, double b, double* y,
                           double rtol, double atol) {
  static const double c2 = 1.0 / 5, c3 = 3.0 / 10, c4 = 4.0 / 5, c5 = 8.0 / 9;
  static const double a21 = 1.0 / 5;
  static const double a31 = 3.0 / 40, a32 = 9.0 / 40;
  static const double a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
  static const double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561,
                      a54 = -212.0 / 729;
  static const double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247,
                      a64 = 49.0 / 176, a65 = -5103.0 / 18656;
  static const double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192,
                      b5 = -2187.0 / 6784, b6 = 11.0 / 84;
  static const double e1 = 71.0 / 57600, e3 = -71.0 / 16695, e4 = 71.0 / 1920,
                      e5 = -17253.0 / 339200, e6 = 22.0 / 525, e7 = -1.0 / 40;
  double t = a, len = b - a;
  if (len <= 0) return 0;
  double p[4];
  double h = len < 1e-2 ? len : 1e-2;
  double k1[9], k2[9], k3[9], k4[9], k5[9], k6[9], k7[9], yt[9], yn[9];
  int64_t steps = 0;
  rates_at(model, th, cycle, n_steps, t, tb, p);
  rhs(p, lam, y, k1);
  while (t < b) {
    if (t + h > b) h = b - t;
    int i;
    for (i = 0; i < 9; ++i) yt[i] = y[i] + h * (a21 * k1[i]);
    rates_at(model, th, cycle, n_steps, t + c2 * h, tb, p); rhs(p, lam, yt, k2);
    for (i = 0; i < 9; ++i) yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
    rates_at(model, th, cycle, n_steps, t + c3 * h, tb, p); rhs(p, lam, yt, k3);
    for (i = 0; i < 9; ++i) yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
    rates_at(model, th, cycle, n_steps, t + c4 * h, tb, p); rhs(p, lam, yt, k4);
    for (i = 0; i < 9; ++i)
      yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    rates_at(model, th, cycle, n_steps, t + c5 * h, tb, p); rhs(p, lam, yt, k5);
    for (i = 0; i < 9; ++i)
      yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    rates_at(model, th, cycle, n_steps, t + h, tb, p); rhs(p, lam, yt, k6);
    for (i = 0; i < 9; ++i)
      yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
    rates_at(model, th, cycle, n_steps, t + h, tb, p); rhs(p, lam, yn, k7);
    double err = 0.0;
    for (i = 0; i < 9; ++i) {
      double ei = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]);
      double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
      double r = fabs(ei) / sc;
      if (r > err) err = r;
    }
    steps++;
    if (err <= 1.0) {
      t += h;
      memcpy(y, yn, sizeof yn);
      memcpy(k1, k7, sizeof k7);
      double fac = err == 0.0 ? 5.0 : 0.9 * orc_exp(-0.2 * orc_log(err));
      if (fac > 5.0) fac = 5.0;
      h *= fac;
    } else {
      double fac = 0.9 * orc_exp(-0.2 * orc_log(err));
      if (fac < 0.2) fac = 0.2;
      h *= fac;
    }
    if (h < 1e-14 * (fabs(t) + 1.0)) h = 1e-14 * (fabs(t) + 1.0);
  }
  return steps;
}

/* ---- method 1: the gene block in closed form ----------------------------
 * On a smooth piece [a, b] kon, koff, gamma and the labelled fraction lam are
 * constant and alpha(a + s) = al0 + al1 s (al1 = 0 without size scaling).
 * The gene block -- E g (0), Var g (3), Cov gu (4), Cov gl (5) -- is a
 * lower-triangular linear system with decay rates K, 2K and nu = K + gamma
 * (K = kon + koff), the stiff part of the moment ODE.  With g_inf = kon / K
 * and D = E g(a) - g_inf:
 *   E g(s)    = g_inf + D e^{-Ks}
 *   Var g(s)  = Pv + A e^{-Ks} + B e^{-2Ks},  Pv = g_inf koff / K,
 *               A = (koff - kon) D / K,  B = Var g(a) - Pv - A
 *   Cov gu(s) = Cov gu(a) e^{-nu s} + (1 - lam) J(s),  Cov gl likewise with lam,
 *   J(s)      = int_0^s e^{-nu (s - r)} (al0 + al1 r) Var g(r) dr
 *             = sum over (C, mu) in {(Pv, 0), (A, K), (B, 2K)} of C (al0 E1 + al1 E2),
 *   E1 = int_0^s e^{-nu (s-r)} e^{-mu r} dr        = (e^{-mu s} - e^{-nu s}) / d,
 *   E2 = int_0^s e^{-nu (s-r)} r e^{-mu r} dr      = (e^{-mu s} (d s - 1) + e^{-nu s}) / d^2,
 * d = nu - mu (nu, gamma, gamma - K), as products with 1 / d formed once per
 * piece.  For |d s| < 1/2 the same integrals are
 * e^{-mu s} s phi1(d s) and e^{-mu s} s^2 phi2(d s), with the series
 * phi1(x) = sum_k (-x)^k / (k+1)!, phi2(x) = sum_k (-x)^k / (k+2)!, k < 16.
 * The other five moments (E u, E l, Var u, Cov ul, Var l: decay gamma and
 * 2 gamma) are integrated by the same Dormand-Prince 5(4) as method 0, with
 * the gene block evaluated exactly at every stage time; the error norm runs
 * over those five.  The explicit method's step is then no longer held below
 * 3.3 / (2K). */
static const double kInvFact[17] = {
    0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
    0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
    0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
    0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45,
    0x1.952c77030ad4ap-49};  /* 1/1! .. 1/17! */

typedef struct {
  double K, gam, nu, ginf, D, Pv, A, B, al0, al1, cu, cl, c4, c5;
  double r0, rK, r2;  /* 1 / d for the three sources: 1 / nu, 1 / gamma, 1 / (gamma - K) */
} gblock_t;

/* E1, E2 for decay nu and source decay mu, d = nu - mu, emu = e^{-mu s}, enu = e^{-nu s} */
static void gb_e12(double d, double rd, double s, double emu, double enu, double* E1,
                   double* E2) {
  const double x = d * s;
  if (fabs(x) < 0.5) {
    const double y = -x;
    double p1 = kInvFact[15], p2 = kInvFact[16];
    for (int k = 14; k >= 0; --k) {
      p1 = p1 * y + kInvFact[k];
      p2 = p2 * y + kInvFact[k + 1];
    }
    *E1 = emu * (s * p1);
    *E2 = emu * ((s * s) * p2);
  } else {
    *E1 = (emu - enu) * rd;
    *E2 = (emu * (x - 1.0) + enu) * (rd * rd);
  }
}

/* p: kon, koff, alpha(a), gamma; al1: d alpha / dt; y: the state at a */
static void gb_setup(gblock_t* q, const double* p, double al1, double lam, const double* y) {
  q->K = p[0] + p[1];
  q->gam = p[3];
  q->nu = q->K + q->gam;
  q->ginf = p[0] / q->K;
  q->D = y[0] - q->ginf;
  q->Pv = q->ginf * (p[1] / q->K);
  q->A = ((p[1] - p[0]) * q->D) / q->K;
  q->B = (y[3] - q->Pv) - q->A;
  q->al0 = p[2];
  q->al1 = al1;
  q->cu = 1 - lam;
  q->cl = lam;
  q->c4 = y[4];
  q->c5 = y[5];
  q->r0 = 1.0 / q->nu;
  q->rK = 1.0 / q->gam;
  q->r2 = 1.0 / (q->gam - q->K);  /* unused (the series) when gamma = K */
}

/* the gene block at a + s into y[0], y[3], y[4], y[5] */
static void gb_eval(const gblock_t* q, double s, double* y) {
  const double e1 = orc_exp(-(q->K * s));
  const double eg = orc_exp(-(q->gam * s));
  const double e2 = e1 * e1;
  const double en = e1 * eg;
  double E10, E20, E1K, E2K, E12, E22;
  gb_e12(q->nu, q->r0, s, 1.0, en, &E10, &E20);
  gb_e12(q->gam, q->rK, s, e1, en, &E1K, &E2K);
  gb_e12(q->gam - q->K, q->r2, s, e2, en, &E12, &E22);
  const double J = q->al0 * ((q->Pv * E10 + q->A * E1K) + q->B * E12) +
                   q->al1 * ((q->Pv * E20 + q->A * E2K) + q->B * E22);
  y[0] = q->ginf + q->D * e1;
  y[3] = (q->Pv + q->A * e1) + q->B * e2;
  y[4] = q->c4 * en + q->cu * J;
  y[5] = q->c5 * en + q->cl * J;
}

/* Method 1 on one smooth piece [a, b]: Dormand-Prince 5(4) over the five
 * moments of Z, the gene block exact. */
static int64_t egb_piece(int model, const double* th, double cycle, int n_steps,
                         double lam, double tb, double a, double b, double* y,
                         double rtol, double atol) {
  static const int Z[5] = {1, 2, 6, 7, 8};
  static const double c2 = 1.0 / 5, c3 = 3.0 / 10, c4 = 4.0 / 5, c5 = 8.0 / 9;
  static const double a21 = 1.0 / 5;
  static const double a31 = 3.0 / 40, a32 = 9.0 / 40;
  static const double a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
  static const double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561,
                      a54 = -212.0 / 729;
  static const double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247,
                      a64 = 49.0 / 176, a65 = -5103.0 / 18656;
  static const double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192,
                      b5 = -2187.0 / 6784, b6 = 11.0 / 84;
  static const double e1 = 71.0 / 57600, e3 = -71.0 / 16695, e4 = 71.0 / 1920,
                      e5 = -17253.0 / 339200, e6 = 22.0 / 525, e7 = -1.0 / 40;
  double t = a, len = b - a;
  if (len <= 0) return 0;
  double p[4], pb[4];
  rates_at(model, th, cycle, n_steps, a, tb, p);
  rates_at(model, th, cycle, n_steps, floor(tb / cycle) * cycle, tb, pb);
  const double al1 = model != 2 ? pb[2] / cycle : 0.0;
  gblock_t q;
  gb_setup(&q, p, al1, lam, y);
  double h = len < 1e-2 ? len : 1e-2;
  double k1[9], k2[9], k3[9], k4[9], k5[9], k6[9], k7[9], yt[9], yn[9];
  int64_t steps = 0;
  memcpy(yt, y, sizeof yt);
  gb_eval(&q, t - a, yt);
  rhs(p, lam, yt, k1);
  while (t < b) {
    if (t + h > b) h = b - t;
    int j, i;
    double ts;
    for (j = 0; j < 5; ++j) { i = Z[j]; yt[i] = y[i] + h * (a21 * k1[i]); }
    ts = t + c2 * h; gb_eval(&q, ts - a, yt);
    rates_at(model, th, cycle, n_steps, ts, tb, p); rhs(p, lam, yt, k2);
    for (j = 0; j < 5; ++j) { i = Z[j]; yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]); }
    ts = t + c3 * h; gb_eval(&q, ts - a, yt);
    rates_at(model, th, cycle, n_steps, ts, tb, p); rhs(p, lam, yt, k3);
    for (j = 0; j < 5; ++j) { i = Z[j]; yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]); }
    ts = t + c4 * h; gb_eval(&q, ts - a, yt);
    rates_at(model, th, cycle, n_steps, ts, tb, p); rhs(p, lam, yt, k4);
    for (j = 0; j < 5; ++j) {
      i = Z[j];
      yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    }
    ts = t + c5 * h; gb_eval(&q, ts - a, yt);
    rates_at(model, th, cycle, n_steps, ts, tb, p); rhs(p, lam, yt, k5);
    for (j = 0; j < 5; ++j) {
      i = Z[j];
      yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    }
    ts = t + h; gb_eval(&q, ts - a, yt);
    rates_at(model, th, cycle, n_steps, ts, tb, p); rhs(p, lam, yt, k6);
    memcpy(yn, yt, sizeof yn);  /* the gene block at t + h */
    for (j = 0; j < 5; ++j) {
      i = Z[j];
      yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
    }
    rhs(p, lam, yn, k7);
    double err = 0.0;
    for (j = 0; j < 5; ++j) {
      i = Z[j];
      double ei = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]);
      double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
      double r = fabs(ei) / sc;
      if (r > err) err = r;
    }
    steps++;
    if (err <= 1.0) {
      t += h;
      memcpy(y, yn, sizeof yn);
      memcpy(k1, k7, sizeof k7);
      double fac = err == 0.0 ? 5.0 : 0.9 * orc_exp(-0.2 * orc_log(err));
      if (fac > 5.0) fac = 5.0;
      h *= fac;
    } else {
      double fac = 0.9 * orc_exp(-0.2 * orc_log(err));
      if (fac < 0.2) fac = 0.2;
      h *= fac;
    }
    if (h < 1e-14 * (fabs(t) + 1.0)) h = 1e-14 * (fabs(t) + 1.0);
  }
  return steps;
}

int64_t orc_moment_solve(int model, const double* th, double cycle, int n_steps, double texp,
                         double pulse, double tmin, double tmax, double* y, double rtol,
                         double atol, int method) {
  /* breakpoints: every bin boundary k*cycle/n_steps and the window edges */
  double bp[4096];
  int nbp = 0;
  double h = cycle / n_steps;
  double kstart = floor(tmin / h) + 1.0;
  for (double k = kstart; k * h < tmax && nbp < 4000; k += 1.0) {
    double bt = k * h;
    if (bt > tmin) bp[nbp++] = bt;
  }
  double we = texp + pulse;
  if (texp > tmin && texp < tmax) bp[nbp++] = texp;
  if (we > tmin && we < tmax) bp[nbp++] = we;
  /* insertion sort (small) */
  for (int i = 1; i < nbp; ++i) {
    double x = bp[i]; int j = i;
    while (j > 0 && bp[j - 1] > x) { bp[j] = bp[j - 1]; --j; }
    bp[j] = x;
  }
  int64_t steps = 0;
  double a = tmin;
  for (int i = 0; i <= nbp; ++i) {
    double b = i < nbp ? bp[i] : tmax;
    if (b <= a) continue;
    double tb = 0.5 * (a + b);
    double lam = lam_at(th, model, n_steps, texp, pulse, tb);
    steps += method == 1 ? egb_piece(model, th, cycle, n_steps, lam, tb, a, b, y, rtol, atol)
                         : dopri_piece(model, th, cycle, n_steps, lam, tb, a, b, y, rtol, atol);
    a = b;
  }
  return steps;
}

void orc_periodic_boundary(const double* e, double* v) {
  memcpy(v, e, 9 * sizeof(double));
  v[4] = e[4] / 2; v[5] = e[5] / 2;
  v[6] = e[6] / 4 + e[1] / 4;
  v[8] = e[8] / 4 + e[2] / 4;
  v[7] = e[7] / 4;
  v[1] = e[1] / 2; v[2] = e[2] / 2;
}

static void cycle_map(int model, const double* th, const orc_moment_cfg* c, const double* in,
                      double* out) {
  memcpy(out, in, 9 * sizeof(double));
  orc_moment_solve(model, th, c->cycle, c->n_steps, -1.0, 0.1, 0.0, c->cycle, out, c->rtol,
                   c->atol, c->method);
}

/* Solve A x = b (9x9) by Gaussian elimination with partial pivoting. */
static void solve9(double A[9][9], double* b) {
  for (int col = 0; col < 9; ++col) {
    int piv = col;
    for (int r = col + 1; r < 9; ++r) if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (piv != col) {
      for (int k = 0; k < 9; ++k) { double t = A[col][k]; A[col][k] = A[piv][k]; A[piv][k] = t; }
      double t = b[col]; b[col] = b[piv]; b[piv] = t;
    }
    for (int r = col + 1; r < 9; ++r) {
      double f = A[r][col] / A[col][col];
      for (int k = col; k < 9; ++k) A[r][k] -= f * A[col][k];
      b[r] -= f * b[col];
    }
  }
  for (int r = 8; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < 9; ++k) s -= A[r][k] * b[k];
    b[r] = s / A[r][r];
  }
}

/* The cycle map is affine, y -> M y + v (the moment ODE is linear): v is the
 * map of the zero state and column i of M the map of unit vector i minus v.
 * Both transient rules work on (M, v), so they need 10 integrations that
 * do not depend on each other, whatever the number of iterations. */
static void cycle_affine(int model, const double* th, const orc_moment_cfg* c, double M[9][9],
                         double* v) {
  double zero[9] = {0};
  cycle_map(model, th, c, zero, v);
  for (int i = 0; i < 9; ++i) {
    double e[9] = {0}, o[9];
    e[i] = 1.0;
    cycle_map(model, th, c, e, o);
    for (int r = 0; r < 9; ++r) M[r][i] = o[r] - v[r];
  }
}

static void affine_apply(double M[9][9], const double* v, const double* y, double* out) {
  for (int r = 0; r < 9; ++r) {
    double acc = v[r];
    for (int i = 0; i < 9; ++i) acc = acc + M[r][i] * y[i];
    out[r] = acc;
  }
}

static int transient_from(double M[9][9], const double* v, const orc_moment_cfg* c, double* ss) {
  if (c->transient_mode == 1) {
    /* exact periodic steady state: the partition P is linear, so the state
     * after division solves (I - P M) x = P v */
    double Pv[9], A[9][9];
    orc_periodic_boundary(v, Pv);
    for (int i = 0; i < 9; ++i) {
      double col[9], pc[9];
      for (int r = 0; r < 9; ++r) col[r] = M[r][i];
      orc_periodic_boundary(col, pc);
      for (int r = 0; r < 9; ++r) A[r][i] = (r == i ? 1.0 : 0.0) - pc[r];
    }
    memcpy(ss, Pv, sizeof Pv);
    solve9(A, ss);
    return 0;
  }
  /* reference rule, model.jl:114-142, on the affine map */
  static const int check_idx[5] = {0, 1, 3, 4, 6};
  double e1[9], e2[9], iv[9];
  affine_apply(M, v, c->iv, e1);
  memcpy(e2, e1, sizeof e1);
  int k = 0, conv = 0;
  const int max_iter = 100;
  while (!conv && k <= max_iter) {
    k++;
    orc_periodic_boundary(e1, iv);
    affine_apply(M, v, iv, e2);
    int nz = 0, ok = 0;
    for (int q = 0; q < 5; ++q) {
      int i = check_idx[q];
      if (e1[i] > 0.0) {
        nz++;
        if (fabs((e1[i] - e2[i]) / e1[i]) <= 0.01) ok++;
      }
    }
    if (ok == nz) conv = 1;
    memcpy(e1, e2, sizeof e2);
  }
  orc_periodic_boundary(e2, ss);
  return k;
}

int orc_transient(int model, const double* th, const orc_moment_cfg* c, double* ss) {
  double M[9][9], v[9];
  cycle_affine(model, th, c, M, v);
  return transient_from(M, v, c, ss);
}

/* Test hook: the reference rule by direct integration of every cycle (as
 * model.jl:128-141 runs it), for checking the affine form above. */
int orc_transient_direct(int model, const double* th, const orc_moment_cfg* c, double* ss) {
  static const int check_idx[5] = {0, 1, 3, 4, 6};
  double e1[9], e2[9], iv[9];
  cycle_map(model, th, c, c->iv, e1);
  memcpy(e2, e1, sizeof e1);
  int k = 0, conv = 0;
  while (!conv && k <= 100) {
    k++;
    orc_periodic_boundary(e1, iv);
    cycle_map(model, th, c, iv, e2);
    int nz = 0, ok = 0;
    for (int q = 0; q < 5; ++q) {
      int i = check_idx[q];
      if (e1[i] > 0.0) {
        nz++;
        if (fabs((e1[i] - e2[i]) / e1[i]) <= 0.01) ok++;
      }
    }
    if (ok == nz) conv = 1;
    memcpy(e1, e2, sizeof e2);
  }
  orc_periodic_boundary(e2, ss);
  return k;
}

/* trajectories, model.jl:146-174.  A whole cycle that starts on a cycle
 * boundary and does not meet the labelling window [texp, texp + pulse] has
 * the rates of [0, T] shifted by whole cycles and no labelling, so it is the
 * affine cycle map (M, v) itself; with M = NULL every cycle is integrated. */
static int64_t trajectory(int model, const double* th, const orc_moment_cfg* c, const double* iv,
                          double M[9][9], const double* v, double age, double pulse, double texp,
                          double* out5) {
  double T = c->cycle, t0 = c->t0, we = texp + pulse;
  int nsols = (age > 0 && age < T) ? (int)floor((age - t0) / T) + 1 : (int)floor((age - t0) / T);
  double tau = t0, y[9], tmp[9];
  int64_t st = 0;
  memcpy(y, iv, sizeof y);
  for (int k = 1; k <= nsols; ++k) {
    double tf = tau < 0 ? tau + T : age;
    if (k > 1) { orc_periodic_boundary(y, tmp); memcpy(y, tmp, sizeof y); }
    if (M && tau < 0 && jmod(tau, T) == 0.0 && (we <= tau || texp >= tf)) {
      affine_apply(M, v, y, tmp);
      memcpy(y, tmp, sizeof y);
    } else {
      st += orc_moment_solve(model, th, T, c->n_steps, texp, pulse, tau, tf, y, c->rtol, c->atol,
                             c->method);
    }
    tau = tf;
  }
  out5[0] = y[1]; out5[1] = y[2]; out5[2] = y[6]; out5[3] = y[7]; out5[4] = y[8];
  return st;
}

static void recover(int model, const double* th, const orc_moment_cfg* c, double* out,
                    int direct, int64_t* traj_steps) {
  double M[9][9], v[9], ss[9];
  cycle_affine(model, th, c, M, v);
  transient_from(M, v, c, ss);
  for (int j = 0; j < c->n_conds; ++j)
    for (int a = 0; a < c->n_ages; ++a) {
      double age = c->ages[a];
      double texp = age - c->pulse[j] - c->chase[j];
      const int64_t st = trajectory(model, th, c, ss, direct ? NULL : M, v, age, c->pulse[j],
                                    texp, out + (j * c->n_ages + a) * 5);
      if (traj_steps) traj_steps[j * c->n_ages + a] = st;
    }
}

void orc_recover(int model, const double* th, const orc_moment_cfg* c, double* out) {
  recover(model, th, c, out, 0, NULL);
}

/* Test hook: orc_recover, also returning the integration steps of each
 * trajectory, [cond * n_ages + age]. */
void orc_recover_steps(int model, const double* th, const orc_moment_cfg* c, double* out,
                       int64_t* traj_steps) {
  recover(model, th, c, out, 0, traj_steps);
}

/* Test hook: orc_recover integrating every cycle of every trajectory. */
void orc_recover_direct(int model, const double* th, const orc_moment_cfg* c, double* out) {
  recover(model, th, c, out, 1, NULL);
}

void orc_downsample(double* s, int n_ages, const double* mb, const double* m2b,
                    const double* vb) {
  for (int i = 0; i < n_ages; ++i) {
    double* r = s + 5 * i;
    double mu = r[0], ml = r[1], vu = r[2], cv = r[3], vl = r[4];
    r[0] = mu * mb[i];
    r[1] = ml * mb[i];
    r[2] = ((mb[i] - m2b[i]) * mu) + (vb[i] * (mu * mu + vu)) + (mb[i] * mb[i] * vu);
    r[4] = ((mb[i] - m2b[i]) * ml) + (vb[i] * (ml * ml + vl)) + (mb[i] * mb[i] * vl);
    r[3] = (vb[i] * (mu * ml + cv)) + (mb[i] * mb[i] * cv);
  }
}

static double wsum(const double* w, const double* s, int col, int n) {
  double acc = w[0] * s[col];
  for (int b = 1; b < n; ++b) acc = acc + w[b] * s[5 * b + col];
  return acc;
}

static double wcov_col(const double* s, int cx, int cy, const double* w, int n) {
  double mx = wsum(w, s, cx, n), my = wsum(w, s, cy, n);
  double acc = w[0] * ((s[cx] - mx) * (s[cy] - my));
  for (int b = 1; b < n; ++b) acc = acc + w[b] * ((s[5 * b + cx] - mx) * (s[5 * b + cy] - my));
  return acc;
}

void orc_moment_stats(const double* s, int n_ages, int n_conds, const double* w,
                      int pulse_cond, int chase_cond, double* out) {
  int nb = n_ages, J = n_conds;
  for (int j = 0; j < J; ++j) {
    const double* sj = s + j * nb * 5;
    const double* wj = w + nb * j;
    double su = wsum(wj, sj, 0, nb), sl = wsum(wj, sj, 1, nb);
    out[4 * nb + j] = sl / (su + sl);
    double stds = sqrt(fabs((wsum(wj, sj, 2, nb) + wcov_col(sj, 0, 0, wj, nb)) *
                            (wsum(wj, sj, 4, nb) + wcov_col(sj, 1, 1, wj, nb))));
    out[4 * nb + J + j] = wsum(wj, sj, 3, nb) / stds;
    out[4 * nb + 2 * J + j] = wcov_col(sj, 0, 1, wj, nb) / stds;
    int slot = j == pulse_cond ? 0 : (j == chase_cond ? 1 : -1);
    if (slot >= 0) {
      for (int b = 0; b < nb; ++b) {
        const double* r = sj + 5 * b;
        double m = r[0] + r[1];
        double eps = m == 0.0 ? 0.0001 : 0.0;
        out[2 * slot * nb + b] = m;
        out[(2 * slot + 1) * nb + b] = ((r[2] + 2 * r[3]) + r[4]) / (r[0] + r[1] + eps);
      }
    }
  }
}

/* abc_sim on the moment model for a batch (the CPU baseline of the GPU
 * moment path): per particle orc_recover, downsampling of each condition with
 * the capture moments of its (age, pulse/chase) group, then orc_moment_stats.
 * theta is n x cols column-major with leading dimension ld; cap[(age * 2 +
 * group) * 3 + {0,1,2}] = E b, E b^2, Var b or NULL; stats is n x K and
 * moments (optional) n x (n_conds * n_ages * 5), both column-major. */
int orc_moment_sim(int model, int64_t n, const double* theta, int64_t ld,
                   const orc_moment_cfg* c, const int32_t* is_chase, const double* cap,
                   const double* w, int pulse_cond, int chase_cond, int n_threads,
                   double* stats, double* moments) {
  if (model < 1 || model > 5 || n < 0 || ld < n) return -1;
  const int A = c->n_ages, J = c->n_conds, K = 4 * A + 3 * J;
  const int cols = (model >= 3 ? c->n_steps : 1) + 4;
  const size_t per = (size_t)J * A * 5;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t i = 0; i < n; ++i) {
    double th[32], out[4 * 16 + 3 * 32];
    double* s = (double*)malloc(sizeof(double) * per);
    for (int k = 0; k < cols; ++k) th[k] = theta[i + ld * k];
    orc_recover(model, th, c, s);
    if (cap) {
      for (int j = 0; j < J; ++j) {
        const int g = is_chase ? (is_chase[j] != 0) : 0;
        double mb[16], m2[16], vb[16];
        for (int a = 0; a < A; ++a) {
          mb[a] = cap[(a * 2 + g) * 3 + 0];
          m2[a] = cap[(a * 2 + g) * 3 + 1];
          vb[a] = cap[(a * 2 + g) * 3 + 2];
        }
        orc_downsample(s + (size_t)j * A * 5, A, mb, m2, vb);
      }
    }
    orc_moment_stats(s, A, J, w, pulse_cond, chase_cond, out);
    for (int k = 0; k < K; ++k) stats[i + n * k] = out[k];
    if (moments)
      for (size_t k = 0; k < per; ++k) moments[i + n * (int64_t)k] = s[k];
    free(s);
  }
  return 0;
}
