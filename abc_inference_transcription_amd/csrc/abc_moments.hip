// abc_moments.hip -- the reference's own simulator on the GPU: the closed
// 9-dimensional linear moment model of the telegraph gene and its unlabelled
// and labelled mRNA (scripts/model.jl), batched over particles.
//
//   rates / size scaling ......... scripts/model.jl:1-27
//   vary map ..................... scripts/model.jl:30-43
//   labelling window ............. scripts/model.jl:58-64
//   moment equations ............. scripts/model.jl:67-86
//   ODE solve .................... scripts/model.jl:89-96 (CVODE_BDF there;
//                                  Dormand-Prince 5(4) here, split at every
//                                  rate discontinuity)
//   periodic boundary ............ scripts/model.jl:98-111
//   transient phase .............. scripts/model.jl:114-142
//   trajectories ................. scripts/model.jl:146-174
//   synthetic data ............... scripts/model.jl:176-187, 191-199
//   downsample ................... scripts/model.jl:221-239
//   abc_sim statistics ........... scripts/abc_simulation.jl:23-46
//   run_part_sim ................. scripts/recover_statistics.jl:2-11
//
// Both integrating kernels are held to 256 VGPRs (two waves per SIMD; the
// trajectory kernel spills 52 bytes per lane): 12 % faster at 65,536
// particles than the unconstrained allocation, which runs one wave per SIMD.
//
// Work decomposition.  moment_cyclemap_kernel: one lane per (particle, start
// vector) integrates the 10 cycle maps that fix the affine cycle map;
// moment_transient_kernel: one lane per particle finds the periodic state
// from it.  moment_traj_kernel: one lane
// per (particle, condition, age) integrates that state from t0 to the age;
// the items of a particle are adjacent, so the lanes of a wavefront share
// their rates (and their stiffness, which sets the step count).
// moment_stats_kernel: one lane per particle downsamples the moments and forms
// the statistics.
//
// Every operation is an IEEE-rounded +, -, *, / or sqrt in the order written,
// the powers use the engine's own exp10 / exp / log, and the library is built
// with -ffp-contract=off: the step sequence, and so every result bit, is that
// of the CPU restatement compiled the same way.
#include "abc_device.h"
#include "abc_kernels.h"

using namespace abcd;

namespace {

// Dormand-Prince 5(4) tableau (Dormand & Prince 1980), first-same-as-last.
constexpr double kC2 = 1.0 / 5, kC3 = 3.0 / 10, kC4 = 4.0 / 5, kC5 = 8.0 / 9;
constexpr double kA21 = 1.0 / 5;
constexpr double kA31 = 3.0 / 40, kA32 = 9.0 / 40;
constexpr double kA41 = 44.0 / 45, kA42 = -56.0 / 15, kA43 = 32.0 / 9;
constexpr double kA51 = 19372.0 / 6561, kA52 = -25360.0 / 2187, kA53 = 64448.0 / 6561,
                 kA54 = -212.0 / 729;
constexpr double kA61 = 9017.0 / 3168, kA62 = -355.0 / 33, kA63 = 46732.0 / 5247,
                 kA64 = 49.0 / 176, kA65 = -5103.0 / 18656;
constexpr double kB1 = 35.0 / 384, kB3 = 500.0 / 1113, kB4 = 125.0 / 192,
                 kB5 = -2187.0 / 6784, kB6 = 11.0 / 84;
constexpr double kE1 = 71.0 / 57600, kE3 = -71.0 / 16695, kE4 = 71.0 / 1920,
                 kE5 = -17253.0 / 339200, kE6 = 22.0 / 525, kE7 = -1.0 / 40;

// Column offsets of (kon, koff, alpha, gamma) in a reference parameter row
// and how many bins each spans (model.jl:30-43 with the flags of
// abc_simulation.jl:78); lambda follows them.
struct VaryMap {
  int o[4], cnt[4], o_lam;
};

ABC_DEV VaryMap vary_map(int model, int nb) {
  VaryMap v;
  const int flag[4] = {model == 3, 0, model == 4, model == 5};
  int k = 0;
  for (int i = 0; i < 4; ++i) {
    v.o[i] = k;
    v.cnt[i] = flag[i] ? nb : 1;
    k += v.cnt[i];
  }
  v.o_lam = k;
  return v;
}

// Julia's mod(x, y) for floats: the floored remainder.
ABC_DEV double jmod(double x, double y) {
  double r = fmod(x, y);
  if (r != 0.0 && ((r < 0.0) != (y < 0.0))) r += y;
  return r;
}

// A smooth piece of the integration: the rates at its interior point tb.
struct Piece {
  double base[4];  // 10^theta of kon, koff, alpha, gamma in tb's bin
  double lam;      // labelled fraction (0 outside the window)
  double cs;       // start of the cycle holding tb (size scaling)
};

struct Particle {
  const double* th;  // parameter row, column stride ld
  int64_t ld;
  VaryMap v;
  int scaling;
};

ABC_DEV double theta_at(const Particle& P, int col) { return P.th[(int64_t)col * P.ld]; }

ABC_DEV Piece make_piece(const Particle& P, const MomCfgDev& cfg, double texp, double pulse,
                         double tb) {
  Piece pc;
  const double m = jmod(tb, cfg.cycle);
  int bin = 0;
  if (m != 0.0) {
    const double h = cfg.cycle / cfg.n_steps;
    bin = (int)ceil(m / h) - 1;
    if (bin < 0) bin = 0;
    if (bin >= cfg.n_steps) bin = cfg.n_steps - 1;
  }
  for (int i = 0; i < 4; ++i)
    pc.base[i] = ex10(theta_at(P, P.v.o[i] + (P.v.cnt[i] > 1 ? bin : 0)));
  pc.lam = (tb >= texp && tb <= texp + pulse) ? ex10(theta_at(P, P.v.o_lam)) : 0.0;
  pc.cs = floor(tb / cfg.cycle) * cfg.cycle;
  return pc;
}

// rates at time t inside the piece: alpha carries the size scaling 1 + a/T
ABC_DEV void rates(const Piece& pc, const Particle& P, double cycle, double t, double* p) {
  p[0] = pc.base[0];
  p[1] = pc.base[1];
  double al = pc.base[2];
  if (P.scaling) {
    const double a = t - pc.cs;
    al = al * (1.0 + a / cycle);
  }
  p[2] = al;
  p[3] = pc.base[3];
}

// y = (E g, E u, E l, Var g, Cov gu, Cov gl, Var u, Cov ul, Var l)
ABC_DEV void rhs(const double* p, double lam, const double* y, double* dy) {
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

ABC_DEV void eval(const Piece& pc, const Particle& P, double cycle, double t, const double* y,
                  double* dy) {
  double p[4];
  rates(pc, P, cycle, t, p);
  rhs(p, pc.lam, y, dy);
}

// Adaptive Dormand-Prince 5(4) over one smooth piece [a, b]; returns the
// attempted steps.
ABC_DEV int64_t dopri_piece(const Piece& pc, const Particle& P, const MomCfgDev& cfg, double a,
                            double b, double* y, const LogEntry* ltab) {
  const double len = b - a;
  if (len <= 0) return 0;
  const double cyc = cfg.cycle, rtol = cfg.rtol, atol = cfg.atol;
  double t = a;
  double h = len < 1e-2 ? len : 1e-2;
  double k1[9], k2[9], k3[9], k4[9], k5[9], k6[9], k7[9], yt[9], yn[9];
  int64_t steps = 0;
  eval(pc, P, cyc, t, y, k1);
  while (t < b) {
    if (t + h > b) h = b - t;
#pragma unroll
    for (int i = 0; i < 9; ++i) yt[i] = y[i] + h * (kA21 * k1[i]);
    eval(pc, P, cyc, t + kC2 * h, yt, k2);
#pragma unroll
    for (int i = 0; i < 9; ++i) yt[i] = y[i] + h * (kA31 * k1[i] + kA32 * k2[i]);
    eval(pc, P, cyc, t + kC3 * h, yt, k3);
#pragma unroll
    for (int i = 0; i < 9; ++i) yt[i] = y[i] + h * (kA41 * k1[i] + kA42 * k2[i] + kA43 * k3[i]);
    eval(pc, P, cyc, t + kC4 * h, yt, k4);
#pragma unroll
    for (int i = 0; i < 9; ++i)
      yt[i] = y[i] + h * (kA51 * k1[i] + kA52 * k2[i] + kA53 * k3[i] + kA54 * k4[i]);
    eval(pc, P, cyc, t + kC5 * h, yt, k5);
#pragma unroll
    for (int i = 0; i < 9; ++i)
      yt[i] = y[i] + h * (kA61 * k1[i] + kA62 * k2[i] + kA63 * k3[i] + kA64 * k4[i] + kA65 * k5[i]);
    eval(pc, P, cyc, t + h, yt, k6);
#pragma unroll
    for (int i = 0; i < 9; ++i)
      yn[i] = y[i] + h * (kB1 * k1[i] + kB3 * k3[i] + kB4 * k4[i] + kB5 * k5[i] + kB6 * k6[i]);
    eval(pc, P, cyc, t + h, yn, k7);
    double err = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double ei =
          h * (kE1 * k1[i] + kE3 * k3[i] + kE4 * k4[i] + kE5 * k5[i] + kE6 * k6[i] + kE7 * k7[i]);
      const double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
      const double r = fabs(ei) / sc;
      if (r > err) err = r;
    }
    steps++;
    if (err <= 1.0) {
      t += h;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        y[i] = yn[i];
        k1[i] = k7[i];
      }
      double fac = err == 0.0 ? 5.0 : 0.9 * ex_core(-0.2 * ln(err, ltab));
      if (fac > 5.0) fac = 5.0;
      h *= fac;
    } else {
      double fac = 0.9 * ex_core(-0.2 * ln(err, ltab));
      if (fac < 0.2) fac = 0.2;
      h *= fac;
    }
    if (h < 1e-14 * (fabs(t) + 1.0)) h = 1e-14 * (fabs(t) + 1.0);
  }
  return steps;
}

// Method 1: the gene block -- E g (0), Var g (3), Cov gu (4), Cov gl (5) --
// in closed form on a smooth piece [a, b], where kon, koff, gamma and the
// labelled fraction lam are constant and alpha(a + s) = al0 + al1 s.  It is a
// lower-triangular linear system with decay rates K, 2K and nu = K + gamma
// (K = kon + koff), the stiff part of the moment ODE.  With g_inf = kon / K
// and D = E g(a) - g_inf:
//   E g(s)    = g_inf + D e^{-Ks}
//   Var g(s)  = Pv + A e^{-Ks} + B e^{-2Ks},  Pv = g_inf koff / K,
//               A = (koff - kon) D / K,  B = Var g(a) - Pv - A
//   Cov gu(s) = Cov gu(a) e^{-nu s} + (1 - lam) J(s),  Cov gl likewise with lam,
//   J(s) = int_0^s e^{-nu (s - r)} (al0 + al1 r) Var g(r) dr
//        = sum over (C, mu) in {(Pv, 0), (A, K), (B, 2K)} of C (al0 E1 + al1 E2),
//   E1 = (e^{-mu s} - e^{-nu s}) / d,  E2 = (e^{-mu s} (d s - 1) + e^{-nu s}) / d^2
//   (as products with 1 / d, formed once per piece),
// d = nu - mu (nu, gamma, gamma - K); for |d s| < 1/2, E1 = e^{-mu s} s phi1(d s)
// and E2 = e^{-mu s} s^2 phi2(d s) by the series phi1(x) = sum (-x)^k / (k+1)!,
// phi2(x) = sum (-x)^k / (k+2)!, k < 16.  E u, E l, Var u, Cov ul and Var l
// (decay gamma and 2 gamma) take the Dormand-Prince steps of method 0 with the
// gene block evaluated exactly at each stage time, and the error norm runs
// over those five, so the step is no longer held below 3.3 / (2K).  The CPU
// restatement performs the same operations in the same order.
struct GBlock {
  double K, gam, nu, ginf, D, Pv, A, B, al0, al1, cu, cl, c4, c5;
  double r0, rK, r2;  // 1 / d for the three sources: 1 / nu, 1 / gamma, 1 / (gamma - K)
};

ABC_DEV void gb_e12(double d, double rd, double s, double emu, double enu, double& E1,
                    double& E2) {
  constexpr double kInvFact[17] = {
      0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
      0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
      0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
      0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45,
      0x1.952c77030ad4ap-49};  // 1/1! .. 1/17!
  const double x = d * s;
  if (fabs(x) < 0.5) {
    const double y = -x;
    double p1 = kInvFact[15], p2 = kInvFact[16];
#pragma unroll
    for (int k = 14; k >= 0; --k) {
      p1 = p1 * y + kInvFact[k];
      p2 = p2 * y + kInvFact[k + 1];
    }
    E1 = emu * (s * p1);
    E2 = emu * ((s * s) * p2);
  } else {
    E1 = (emu - enu) * rd;
    E2 = (emu * (x - 1.0) + enu) * (rd * rd);
  }
}

// p: kon, koff, alpha(a), gamma; al1: d alpha / dt; y: the state at a
ABC_DEV void gb_setup(GBlock& q, const double* p, double al1, double lam, const double* y) {
  q.K = p[0] + p[1];
  q.gam = p[3];
  q.nu = q.K + q.gam;
  q.ginf = p[0] / q.K;
  q.D = y[0] - q.ginf;
  q.Pv = q.ginf * (p[1] / q.K);
  q.A = ((p[1] - p[0]) * q.D) / q.K;
  q.B = (y[3] - q.Pv) - q.A;
  q.al0 = p[2];
  q.al1 = al1;
  q.cu = 1 - lam;
  q.cl = lam;
  q.c4 = y[4];
  q.c5 = y[5];
  q.r0 = 1.0 / q.nu;
  q.rK = 1.0 / q.gam;
  q.r2 = 1.0 / (q.gam - q.K);  // unused (the series) when gamma = K
}

// the gene block at a + s into y[0], y[3], y[4], y[5]
ABC_DEV void gb_eval(const GBlock& q, double s, double* y) {
  const double e1 = ex_core(-(q.K * s));
  const double eg = ex_core(-(q.gam * s));
  const double e2 = e1 * e1;
  const double en = e1 * eg;
  double E10, E20, E1K, E2K, E12, E22;
  gb_e12(q.nu, q.r0, s, 1.0, en, E10, E20);
  gb_e12(q.gam, q.rK, s, e1, en, E1K, E2K);
  gb_e12(q.gam - q.K, q.r2, s, e2, en, E12, E22);
  const double J = q.al0 * ((q.Pv * E10 + q.A * E1K) + q.B * E12) +
                   q.al1 * ((q.Pv * E20 + q.A * E2K) + q.B * E22);
  y[0] = q.ginf + q.D * e1;
  y[3] = (q.Pv + q.A * e1) + q.B * e2;
  y[4] = q.c4 * en + q.cu * J;
  y[5] = q.c5 * en + q.cl * J;
}

// Method 1 on one smooth piece [a, b]; returns the attempted steps.
ABC_DEV int64_t egb_piece(const Piece& pc, const Particle& P, const MomCfgDev& cfg, double a,
                          double b, double* y, const LogEntry* ltab) {
  constexpr int Z[5] = {1, 2, 6, 7, 8};
  const double len = b - a;
  if (len <= 0) return 0;
  const double cyc = cfg.cycle, rtol = cfg.rtol, atol = cfg.atol;
  double p[4];
  rates(pc, P, cyc, a, p);
  const double al1 = P.scaling ? pc.base[2] / cyc : 0.0;
  GBlock q;
  gb_setup(q, p, al1, pc.lam, y);
  double t = a;
  double h = len < 1e-2 ? len : 1e-2;
  double k1[9], k2[9], k3[9], k4[9], k5[9], k6[9], k7[9], yt[9], yn[9];
  int64_t steps = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) yt[i] = y[i];
  gb_eval(q, t - a, yt);
  rhs(p, pc.lam, yt, k1);
  while (t < b) {
    if (t + h > b) h = b - t;
    double ts;
#pragma unroll
    for (int j = 0; j < 5; ++j) { const int i = Z[j]; yt[i] = y[i] + h * (kA21 * k1[i]); }
    ts = t + kC2 * h;
    gb_eval(q, ts - a, yt);
    rates(pc, P, cyc, ts, p);
    rhs(p, pc.lam, yt, k2);
#pragma unroll
    for (int j = 0; j < 5; ++j) { const int i = Z[j]; yt[i] = y[i] + h * (kA31 * k1[i] + kA32 * k2[i]); }
    ts = t + kC3 * h;
    gb_eval(q, ts - a, yt);
    rates(pc, P, cyc, ts, p);
    rhs(p, pc.lam, yt, k3);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int i = Z[j];
      yt[i] = y[i] + h * (kA41 * k1[i] + kA42 * k2[i] + kA43 * k3[i]);
    }
    ts = t + kC4 * h;
    gb_eval(q, ts - a, yt);
    rates(pc, P, cyc, ts, p);
    rhs(p, pc.lam, yt, k4);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int i = Z[j];
      yt[i] = y[i] + h * (kA51 * k1[i] + kA52 * k2[i] + kA53 * k3[i] + kA54 * k4[i]);
    }
    ts = t + kC5 * h;
    gb_eval(q, ts - a, yt);
    rates(pc, P, cyc, ts, p);
    rhs(p, pc.lam, yt, k5);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int i = Z[j];
      yt[i] = y[i] + h * (kA61 * k1[i] + kA62 * k2[i] + kA63 * k3[i] + kA64 * k4[i] + kA65 * k5[i]);
    }
    ts = t + h;
    gb_eval(q, ts - a, yt);
    rates(pc, P, cyc, ts, p);
    rhs(p, pc.lam, yt, k6);
#pragma unroll
    for (int i = 0; i < 9; ++i) yn[i] = yt[i];  // the gene block at t + h
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int i = Z[j];
      yn[i] = y[i] + h * (kB1 * k1[i] + kB3 * k3[i] + kB4 * k4[i] + kB5 * k5[i] + kB6 * k6[i]);
    }
    rhs(p, pc.lam, yn, k7);
    double err = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int i = Z[j];
      const double ei =
          h * (kE1 * k1[i] + kE3 * k3[i] + kE4 * k4[i] + kE5 * k5[i] + kE6 * k6[i] + kE7 * k7[i]);
      const double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
      const double r = fabs(ei) / sc;
      if (r > err) err = r;
    }
    steps++;
    if (err <= 1.0) {
      t += h;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        y[i] = yn[i];
        k1[i] = k7[i];
      }
      double fac = err == 0.0 ? 5.0 : 0.9 * ex_core(-0.2 * ln(err, ltab));
      if (fac > 5.0) fac = 5.0;
      h *= fac;
    } else {
      double fac = 0.9 * ex_core(-0.2 * ln(err, ltab));
      if (fac < 0.2) fac = 0.2;
      h *= fac;
    }
    if (h < 1e-14 * (fabs(t) + 1.0)) h = 1e-14 * (fabs(t) + 1.0);
  }
  return steps;
}

// model(): integrate [tmin, tmax], split at every bin edge k * T / n_steps and
// at the window edges texp, texp + pulse.  The edges are generated in
// ascending order (a sorted merge of the grid with the two window edges);
// coincident edges give empty pieces, which are skipped.
template <int METHOD>
ABC_DEV int64_t moment_solve(const Particle& P, const MomCfgDev& cfg, double texp, double pulse,
                             double tmin, double tmax, double* y, const LogEntry* ltab) {
  const double h = cfg.cycle / cfg.n_steps;
  double k = floor(tmin / h) + 1.0;
  const double we = texp + pulse;
  bool w0 = texp > tmin && texp < tmax;
  bool w1 = we > tmin && we < tmax;
  int64_t steps = 0;
  double a = tmin;
  int grid_left = 4000;  // the same cap on grid edges as the CPU restatement
  for (;;) {
    // next edge: the smallest of the pending grid edge and window edges
    bool have_grid = false;
    double g = 0.0;
    while (grid_left > 0 && k * h < tmax) {
      g = k * h;
      if (g > tmin) { have_grid = true; break; }
      k += 1.0;  // an edge at or before tmin is not an edge (and not counted)
    }
    double b = tmax;
    int src = -1;
    if (have_grid) { b = g; src = 0; }
    if (w0 && (src < 0 || texp < b)) { b = texp; src = 1; }
    if (w1 && (src < 0 || we < b)) { b = we; src = 2; }
    if (src == 0) { k += 1.0; grid_left--; }
    else if (src == 1) w0 = false;
    else if (src == 2) w1 = false;
    if (b > a) {
      const double tb = 0.5 * (a + b);
      const Piece pc = make_piece(P, cfg, texp, pulse, tb);
      steps += METHOD == 1 ? egb_piece(pc, P, cfg, a, b, y, ltab)
                           : dopri_piece(pc, P, cfg, a, b, y, ltab);
      a = b;
    }
    if (src < 0) break;
  }
  return steps;
}

// binomial partitioning at division (model.jl:98-111)
ABC_DEV void periodic_boundary(const double* e, double* v) {
#pragma unroll
  for (int i = 0; i < 9; ++i) v[i] = e[i];
  v[4] = e[4] / 2;
  v[5] = e[5] / 2;
  v[6] = e[6] / 4 + e[1] / 4;
  v[8] = e[8] / 4 + e[2] / 4;
  v[7] = e[7] / 4;
  v[1] = e[1] / 2;
  v[2] = e[2] / 2;
}

template <int METHOD>
ABC_DEV int64_t cycle_map(const Particle& P, const MomCfgDev& cfg, const double* in, double* out,
                          const LogEntry* ltab) {
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = in[i];
  return moment_solve<METHOD>(P, cfg, -1.0, 0.1, 0.0, cfg.cycle, out, ltab);
}

ABC_DEV Particle particle_of(int model, const MomCfgDev& cfg, const double* theta, int64_t ld,
                             int64_t p) {
  Particle P;
  P.th = theta + p;
  P.ld = ld;
  P.v = vary_map(model, cfg.n_steps);
  P.scaling = model != 2;
  return P;
}

}  // namespace

// The cycle map is affine, y -> M y + v (the moment ODE is linear): v is the
// map of the zero state and column i of M the map of unit vector i minus v.
// Column c of particle p (c = 0 the zero state, c = 1 + i unit vector i) goes
// to maps[(p * 10 + c) * 9 + r]; the transient rules below need nothing else,
// so a particle's longest chain is one cycle whatever its iterations.
// Wave w integrates column w % 10 for the 64 particles order[64 (w / 10) ..]
// (an order by predicted cost) and adds each integration's steps to cost[p].
template <int METHOD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void moment_cyclemap_kernel(
    const MomCfgDev cfg, int model, int64_t n, const double* __restrict__ theta, int64_t ld,
    const int32_t* __restrict__ order, double* __restrict__ maps,
    unsigned long long* __restrict__ cost, unsigned long long* __restrict__ steps) {
  __shared__ LogEntry ltab[ABC_LOG_N];
  load_log_table(ltab);
  __syncthreads();
  const int64_t w = blockIdx.x;  // 64 lanes per workgroup
  const int64_t g = w / 10;
  const int col = (int)(w - g * 10);
  const int64_t i = g * 64 + threadIdx.x;
  if (i >= n) return;
  const int64_t p = order[i];
  const Particle P = particle_of(model, cfg, theta, ld, p);
  double in[9], out[9];
  for (int k = 0; k < 9; ++k) in[k] = (col > 0 && k == col - 1) ? 1.0 : 0.0;
  const int64_t st = cycle_map<METHOD>(P, cfg, in, out, ltab);
  double* o = maps + (p * 10 + col) * 9;
  for (int k = 0; k < 9; ++k) o[k] = out[k];
  atomicAdd(cost + p, (unsigned long long)st);
  if (steps) atomicAdd(steps + p, (unsigned long long)st);
}

// Predicted cycle-map cost of particle p, written to cost[p] in units of
// 2^-20, summed over the rate bins of the cycle.  Method 0: the moment ODE's
// fastest rate, 2 max(kon + koff, gamma) (the explicit step is held near a
// fixed multiple of its inverse).  Method 1 (gene block exact): gamma
// alpha^0.3 (a fit of its step counts over the reference prior gave
// steps ~ gamma^0.3 alpha^0.1 and no dependence on kon + koff; the ranking
// is what counts here).
__global__ __launch_bounds__(256) void moment_key_kernel(const MomCfgDev cfg, int model, int64_t n,
                                                         const double* __restrict__ theta,
                                                         int64_t ld,
                                                         unsigned long long* __restrict__ cost) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const Particle P = particle_of(model, cfg, theta, ld, p);
  double key = 0.0;
  for (int b = 0; b < cfg.n_steps; ++b) {
    double lg[4];
    for (int k = 0; k < 4; ++k) lg[k] = theta_at(P, P.v.o[k] + (P.v.cnt[k] > 1 ? b : 0));
    if (cfg.method == 1)
      key += ex10(lg[3] + 0.3 * lg[2]);
    else
      key += 2.0 * fmax(ex10(lg[0]) + ex10(lg[1]), ex10(lg[3]));
  }
  key *= 1048576.0;
  cost[p] = key < 1.8e19 ? (unsigned long long)key : 18000000000000000000ull;
}

// Counting sort of the particles by descending cost, on a log
// scale of 16 buckets per octave (bucket = 16 * floor(log2 c) + the next 4
// bits of c), in one workgroup.  A particle's trajectories integrate the
// same cycles as its neighbours' in this order, at a step size set by the
// same rates, so a wave of 64 of them finishes nearly together.  Order
// within a bucket follows LDS atomics and is not fixed; each trajectory's
// result is independent of it.
__global__ __launch_bounds__(ABC_MOM_ORDER_WG) void moment_order_kernel(
    int64_t n, const unsigned long long* __restrict__ cost, int32_t* __restrict__ order) {
  __shared__ unsigned int hist[ABC_MOM_COST_BUCKETS];
  __shared__ unsigned int scan[ABC_MOM_COST_BUCKETS];
  static_assert(ABC_MOM_COST_BUCKETS == ABC_MOM_ORDER_WG, "one bucket per lane");
  const int tid = threadIdx.x;
  auto bucket = [](unsigned long long c) -> int {
    if (c < 16ull) return (int)c;
    const int e = 63 - __clzll(c);  // >= 4
    return 16 * (e - 3) + (int)((c >> (e - 4)) & 15ull);
  };
  hist[tid] = 0u;
  __syncthreads();
  for (int64_t p = tid; p < n; p += ABC_MOM_ORDER_WG) atomicAdd(&hist[bucket(cost[p])], 1u);
  __syncthreads();
  // descending: lane t holds bucket B - 1 - t; inclusive Hillis-Steele scan
  const unsigned int own = hist[ABC_MOM_COST_BUCKETS - 1 - tid];
  scan[tid] = own;
  __syncthreads();
  for (int d = 1; d < ABC_MOM_ORDER_WG; d <<= 1) {
    const unsigned int add = tid >= d ? scan[tid - d] : 0u;
    __syncthreads();
    scan[tid] += add;
    __syncthreads();
  }
  hist[ABC_MOM_COST_BUCKETS - 1 - tid] = scan[tid] - own;  // first slot of the bucket
  __syncthreads();
  for (int64_t p = tid; p < n; p += ABC_MOM_ORDER_WG) {
    const unsigned int at = atomicAdd(&hist[bucket(cost[p])], 1u);
    order[at] = (int32_t)p;
  }
}

// Periodic state after division, one lane per particle, from (M, v).  Mode 0
// is the reference's rule (model.jl:114-142): iterate the cycle map from iv
// until the gene and unlabelled moments change by at most 1 % (at most 101
// maps).  Mode 1 solves the fixed point exactly: the partition P is linear,
// so the state after division solves (I - P M) x = P v (Gaussian elimination
// with partial pivoting).
__global__ __launch_bounds__(64) void moment_transient_kernel(const MomCfgDev cfg, int64_t n,
                                                               const double* __restrict__ maps,
                                                               double* __restrict__ ss) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* mp = maps + p * 90;
  double M[9][9], v[9], out[9];
  for (int r = 0; r < 9; ++r) v[r] = mp[r];
  for (int i = 0; i < 9; ++i)
    for (int r = 0; r < 9; ++r) M[r][i] = mp[9 * (i + 1) + r] - v[r];
  if (cfg.transient_mode == 1) {
    double Pv[9], A[9][9];
    for (int i = 0; i < 9; ++i) {
      double col[9], pc[9];
      for (int r = 0; r < 9; ++r) col[r] = M[r][i];
      periodic_boundary(col, pc);
      for (int r = 0; r < 9; ++r) A[r][i] = (r == i ? 1.0 : 0.0) - pc[r];
    }
    periodic_boundary(v, Pv);
    for (int col = 0; col < 9; ++col) {
      int piv = col;
      for (int r = col + 1; r < 9; ++r)
        if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
      if (piv != col) {
        for (int k = 0; k < 9; ++k) {
          const double t = A[col][k];
          A[col][k] = A[piv][k];
          A[piv][k] = t;
        }
        const double t = Pv[col];
        Pv[col] = Pv[piv];
        Pv[piv] = t;
      }
      for (int r = col + 1; r < 9; ++r) {
        const double f = A[r][col] / A[col][col];
        for (int k = col; k < 9; ++k) A[r][k] -= f * A[col][k];
        Pv[r] -= f * Pv[col];
      }
    }
    for (int r = 8; r >= 0; --r) {
      double s = Pv[r];
      for (int k = r + 1; k < 9; ++k) s -= A[r][k] * Pv[k];
      Pv[r] = s / A[r][r];
    }
    for (int i = 0; i < 9; ++i) out[i] = Pv[i];
  } else {
    const int check_idx[5] = {0, 1, 3, 4, 6};
    double e1[9], e2[9], iv[9];
    for (int r = 0; r < 9; ++r) {
      double acc = v[r];
      for (int i = 0; i < 9; ++i) acc = acc + M[r][i] * cfg.iv[i];
      e1[r] = acc;
    }
    for (int i = 0; i < 9; ++i) e2[i] = e1[i];
    int k = 0;
    bool conv = false;
    while (!conv && k <= 100) {
      k++;
      periodic_boundary(e1, iv);
      for (int r = 0; r < 9; ++r) {
        double acc = v[r];
        for (int i = 0; i < 9; ++i) acc = acc + M[r][i] * iv[i];
        e2[r] = acc;
      }
      int nz = 0, ok = 0;
      for (int qq = 0; qq < 5; ++qq) {
        const int i = check_idx[qq];
        if (e1[i] > 0.0) {
          nz++;
          if (fabs((e1[i] - e2[i]) / e1[i]) <= 0.01) ok++;
        }
      }
      if (ok == nz) conv = true;
      for (int i = 0; i < 9; ++i) e1[i] = e2[i];
    }
    periodic_boundary(e2, out);
  }
  for (int i = 0; i < 9; ++i) ss[p + (int64_t)i * n] = out[i];
}

// trajectories (model.jl:146-174) for item (particle, condition, age): out
// column (cond * n_ages + age) * 5 + {mean_u, mean_l, var_u, cov_ul, var_l}.
// Wave w takes one (condition, age) item for the 64 particles
// order[64 g .. 64 g + 63], g = w / items, so its lanes share the item's
// cycle structure and have similar step counts.
// A whole cycle that starts on a cycle boundary and does not meet the
// labelling window has the rates of [0, T] shifted by whole cycles and no
// labelling, so it is the particle's affine cycle map, applied from maps.
template <int METHOD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void moment_traj_kernel(const MomCfgDev cfg,
                                                          int model, int64_t n,
                                                          const double* __restrict__ theta,
                                                          int64_t ld,
                                                          const double* __restrict__ ss,
                                                          const double* __restrict__ maps,
                                                          const int32_t* __restrict__ order,
                                                          double* __restrict__ out,
                                                          unsigned long long* __restrict__ steps) {
  __shared__ LogEntry ltab[ABC_LOG_N];
  load_log_table(ltab);
  __syncthreads();
  const int per = cfg.n_conds * cfg.n_ages;
  const int64_t w = blockIdx.x;  // 64 lanes per workgroup
  const int64_t g = w / per;
  const int r = (int)(w - g * per);
  const int64_t i = g * 64 + threadIdx.x;
  if (i >= n) return;
  const int64_t p = order[i];
  const int j = r / cfg.n_ages, ai = r - j * cfg.n_ages;
  const Particle P = particle_of(model, cfg, theta, ld, p);
  const double T = cfg.cycle, t0 = cfg.t0, age = cfg.ages[ai];
  const double pulse = cfg.pulse[j], texp = age - cfg.pulse[j] - cfg.chase[j];
  const int nsols = (age > 0 && age < T) ? (int)floor((age - t0) / T) + 1 : (int)floor((age - t0) / T);
  const double we = texp + pulse;
  const double* mp = maps + p * 90;
  double y[9], tmp[9];
  for (int i = 0; i < 9; ++i) y[i] = ss[p + (int64_t)i * n];
  double tau = t0;
  int64_t st = 0;
  for (int k = 1; k <= nsols; ++k) {
    const double tf = tau < 0 ? tau + T : age;
    if (k > 1) {
      periodic_boundary(y, tmp);
      for (int i = 0; i < 9; ++i) y[i] = tmp[i];
    }
    if (tau < 0 && jmod(tau, T) == 0.0 && (we <= tau || texp >= tf)) {
      // y <- v + M y with M[r][i] = map_i[r] - v[r], in the CPU restatement's order
      for (int r = 0; r < 9; ++r) {
        const double vr = mp[r];
        double acc = vr;
        for (int i = 0; i < 9; ++i) acc = acc + (mp[9 * (i + 1) + r] - vr) * y[i];
        tmp[r] = acc;
      }
      for (int i = 0; i < 9; ++i) y[i] = tmp[i];
    } else {
      st += moment_solve<METHOD>(P, cfg, texp, pulse, tau, tf, y, ltab);
    }
    tau = tf;
  }
  double* o = out + p + (int64_t)r * 5 * n;
  o[0] = y[1];
  o[n] = y[2];
  o[2 * n] = y[6];
  o[3 * n] = y[7];
  o[4 * n] = y[8];
  if (steps) atomicAdd(steps + p, (unsigned long long)st);
}

namespace {

ABC_DEV double wsum(const double* w, const double* s, int col, int nb) {
  double acc = w[0] * s[col];
  for (int b = 1; b < nb; ++b) acc = acc + w[b] * s[5 * b + col];
  return acc;
}

ABC_DEV double wcov_col(const double* s, int cx, int cy, const double* w, int nb) {
  const double mx = wsum(w, s, cx, nb), my = wsum(w, s, cy, nb);
  double acc = w[0] * ((s[cx] - mx) * (s[cy] - my));
  for (int b = 1; b < nb; ++b) acc = acc + w[b] * ((s[5 * b + cx] - mx) * (s[5 * b + cy] - my));
  return acc;
}

}  // namespace

// Per particle: capture downsampling of each condition's moments with the
// capture moments of its (age, pulse/chase) group (model.jl:221-239), then
// the statistics of abc_sim (abc_simulation.jl:23-46): labelled ratio,
// mean correlation and correlation of means per condition, and the mean and
// Fano factor per age of the pulse and chase conditions, in the order
// [pulse mean, pulse FF, chase mean, chase FF, ratios, mean_corr, corr_mean].
// The downsampled moments are written back over the input.
__global__ __launch_bounds__(64) void moment_stats_kernel(const MomCfgDev cfg,
                                                           int64_t n,
                                                           double* __restrict__ mom,
                                                           double* __restrict__ stats) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nb = cfg.n_ages, J = cfg.n_conds;
  double s[ABC_MOM_MAX_AGES * 5];
  for (int j = 0; j < J; ++j) {
    for (int a = 0; a < nb; ++a)
      for (int k = 0; k < 5; ++k) s[5 * a + k] = mom[p + (int64_t)((j * nb + a) * 5 + k) * n];
    if (cfg.downsample) {
      const int g = cfg.is_chase[j];
      for (int i = 0; i < nb; ++i) {
        double* r = s + 5 * i;
        const double mb = cfg.cap[(i * 2 + g) * 3 + 0], m2b = cfg.cap[(i * 2 + g) * 3 + 1],
                     vb = cfg.cap[(i * 2 + g) * 3 + 2];
        const double mu = r[0], ml = r[1], vu = r[2], cv = r[3], vl = r[4];
        r[0] = mu * mb;
        r[1] = ml * mb;
        r[2] = ((mb - m2b) * mu) + (vb * (mu * mu + vu)) + (mb * mb * vu);
        r[4] = ((mb - m2b) * ml) + (vb * (ml * ml + vl)) + (mb * mb * vl);
        r[3] = (vb * (mu * ml + cv)) + (mb * mb * cv);
      }
      for (int a = 0; a < nb; ++a)
        for (int k = 0; k < 5; ++k) mom[p + (int64_t)((j * nb + a) * 5 + k) * n] = s[5 * a + k];
    }
    const double* wj = cfg.weights + nb * j;
    const double su = wsum(wj, s, 0, nb), sl = wsum(wj, s, 1, nb);
    stats[p + (int64_t)(4 * nb + j) * n] = sl / (su + sl);
    const double stds = sqrt(fabs((wsum(wj, s, 2, nb) + wcov_col(s, 0, 0, wj, nb)) *
                                  (wsum(wj, s, 4, nb) + wcov_col(s, 1, 1, wj, nb))));
    stats[p + (int64_t)(4 * nb + J + j) * n] = wsum(wj, s, 3, nb) / stds;
    stats[p + (int64_t)(4 * nb + 2 * J + j) * n] = wcov_col(s, 0, 1, wj, nb) / stds;
    const int slot = j == cfg.pulse_cond ? 0 : (j == cfg.chase_cond ? 1 : -1);
    if (slot >= 0) {
      for (int b = 0; b < nb; ++b) {
        const double* r = s + 5 * b;
        const double m = r[0] + r[1];
        const double eps = m == 0.0 ? 0.0001 : 0.0;
        stats[p + (int64_t)(2 * slot * nb + b) * n] = m;
        stats[p + (int64_t)((2 * slot + 1) * nb + b) * n] = ((r[2] + 2 * r[3]) + r[4]) / (r[0] + r[1] + eps);
      }
    }
  }
}

// one instantiation per integration method (cfg.method), so that each has its
// own register allocation
template __global__ void moment_cyclemap_kernel<0>(MomCfgDev, int, int64_t, const double*, int64_t,
                                                   const int32_t*, double*, unsigned long long*,
                                                   unsigned long long*);
template __global__ void moment_cyclemap_kernel<1>(MomCfgDev, int, int64_t, const double*, int64_t,
                                                   const int32_t*, double*, unsigned long long*,
                                                   unsigned long long*);
template __global__ void moment_traj_kernel<0>(MomCfgDev, int, int64_t, const double*, int64_t,
                                               const double*, const double*, const int32_t*,
                                               double*, unsigned long long*);
template __global__ void moment_traj_kernel<1>(MomCfgDev, int, int64_t, const double*, int64_t,
                                               const double*, const double*, const int32_t*,
                                               double*, unsigned long long*);
