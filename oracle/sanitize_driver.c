/*
 * sanitize_driver.c -- runs every entry point of the CPU oracle on small
 * inputs, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`).  TEST INFRASTRUCTURE ONLY: it checks the
 * checker's memory safety, and prints a few sums so that nothing is
 * optimised away.  Exit status 0 = every call returned 0 and no sanitizer
 * reported anything (the build aborts on the first report).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double REF_PULSE[11] = {0.25, 0.5, 0.75, 1.0, 2.0, 3.0, 22.0, 22.0, 22.0, 22.0, 22.0};
static const double REF_CHASE[11] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0, 6.0};
static const int32_t REF_IS_CHASE[11] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
static const double CFG_PULSE[9] = {0.5, 1.0, 2.0, 3.0, 22.0, 22.0, 22.0, 22.0, 22.0};
static const double CFG_CHASE[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0, 6.0};
static const int32_t CFG_IS_CHASE[9] = {0, 0, 0, 0, 1, 1, 1, 1, 1};

static int fails = 0;
#define CHECK(expr)                                                   \
  do {                                                                \
    int rc_ = (expr);                                                 \
    if (rc_ != 0) {                                                   \
      fprintf(stderr, "%s:%d: %s returned %d\n", __FILE__, __LINE__, #expr, rc_); \
      fails++;                                                        \
    }                                                                 \
  } while (0)

static void simulate_all(const orc_design* d, const char* name) {
  const int nb = d->n_bins, J = d->n_conds, K = 4 * nb + 3 * J, NBJ = nb * J;
  const int64_t n = 3;
  double sum = 0.0;
  for (int m = 1; m <= 5; ++m) {
    const int np = orc_n_params(m);
    double* theta = calloc((size_t)n * np, sizeof(double));
    double* stats = calloc((size_t)n * K, sizeof(double));
    double* mom = calloc((size_t)n * NBJ * 6, sizeof(double));
    double* sup = calloc((size_t)n * NBJ * 2, sizeof(double));
    uint32_t* cu = calloc((size_t)n * d->cells, sizeof(uint32_t));
    uint32_t* cl = calloc((size_t)n * d->cells, sizeof(uint32_t));
    uint8_t* cb = calloc((size_t)n * d->cells, 1);
    int64_t steps[3], react[3];
    orc_sample_prior(d, m, 7, 100, n, theta);
    CHECK(orc_simulate(d, m, 7, 100, n, theta, 1, cu, cl, cb, stats, mom, sup, steps, react));
    for (int64_t q = 0; q < n * K; ++q)
      if (isfinite(stats[q])) sum += stats[q];
    free(theta); free(stats); free(mom); free(sup); free(cu); free(cl); free(cb);
  }
  printf("%s: simulate sum %.6g\n", name, sum);
}

int main(void) {
  /* random numbers and elementary functions */
  {
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
    orc_philox4x32_10(ctr, key, out);
    double s = orc_u01_32(out[0]) + orc_u01_53(out[1], out[2]) + orc_log(3.5) + orc_exp(-2.0) +
               orc_exp10(0.3) + orc_u128_to_f64(1ull << 40, 12345);
    for (uint32_t w = 0; w < 4096; ++w) s += orc_exp1_32(w * 1048573u);
    s += orc_exp1_32(0xFFFFFFFFu) + orc_log(5e-324) + orc_log(1.0 + 1e-15);
    s += (double)orc_exp_scaling_mismatches(-745.2, 709.0, 10000, 9);
    uint32_t po[64];
    orc_poisson_sample(3.5, 1, 64, po);
    orc_poisson_sample(250.0, 2, 64, po);
    orc_poisson_sample(0.0, 3, 64, po);
    printf("rng/math sum %.6g, poisson %u\n", s, po[0]);
  }
  /* stochastic simulation: configured design (Beta capture, split cells),
   * reference design (no capture), reference design with an empirical list */
  double w_cfg[5 * 9], w_ref[5 * 11];
  for (int q = 0; q < 5 * 9; ++q) w_cfg[q] = 0.2;
  for (int q = 0; q < 5 * 11; ++q) w_ref[q] = 0.2;
  orc_design cfg = {21.0, 5, 9, CFG_PULSE, CFG_CHASE, CFG_IS_CHASE, 60, 1, 1, 2, 8, NULL, NULL,
                    w_cfg, 5e-3, 16, 1, 0.0, 200.0, 16};
  simulate_all(&cfg, "config design, split cells");
  orc_design ref = {20.0, 5, 11, REF_PULSE, REF_CHASE, REF_IS_CHASE, 55, 0, 0, 2, 8, NULL, NULL,
                    w_ref, 5e-3, 16, 0, 0.0, 32768.0, 1024};
  simulate_all(&ref, "reference design");
  double betas[40];
  int64_t goff[11];
  for (int q = 0; q < 40; ++q) betas[q] = 0.05 + 0.01 * q;
  for (int q = 0; q <= 10; ++q) goff[q] = 4 * q;
  orc_design emp = ref;
  emp.capture_mode = 2;
  emp.betas = betas;
  emp.group_off = goff;
  simulate_all(&emp, "reference design, empirical capture");
  {
    double theta[6] = {0.5, -0.2, 1.0, -0.5, -0.3, 0.0}, rec[8 * 20];
    CHECK(orc_burn_cells(&cfg, 2, 3, 9, theta, 20, rec));
    double io[2] = {1.0, 2.0};
    orc_burn_step(io, 1, 3.0, 0.5, 1.2, 0.7, 2.0);
    orc_burn_step(io, 0, 3.0, 0.5, 1.2, 0.7, 1e-9);
    uint32_t u[60], l[60];
    uint8_t bin[60];
    double st[4 * 5 + 3 * 9], mo[5 * 9 * 6], su[5 * 9 * 2];
    for (int c = 0; c < 60; ++c) { u[c] = (uint32_t)(c % 7); l[c] = (uint32_t)(c % 3); bin[c] = (uint8_t)(c % 5); }
    orc_cell_stats(&cfg, u, l, bin, st, mo, su);
    printf("burn/cell stats %.6g %.6g\n", rec[2], st[0]);
  }
  /* moment model */
  {
    const double ages[5] = {2.0, 6.0, 10.0, 14.0, 18.0};
    orc_moment_cfg mc = {20.0, 5, 5, ages, 11, REF_PULSE, REF_CHASE, -60.0,
                         {0.0, 0.5, 0, 0, 0, 0, 0, 0, 0}, 0, 1e-8, 1e-10, 1};
    double cap[5 * 2 * 3];
    for (int q = 0; q < 10; ++q) { cap[3 * q] = 0.2; cap[3 * q + 1] = 0.05; cap[3 * q + 2] = 0.01; }
    for (int m = 1; m <= 5; ++m) {
      const int np = m <= 2 ? 5 : 9;
      double theta[2 * 9], stats[2 * 53], moms[2 * 11 * 5 * 5];
      for (int q = 0; q < 2 * np; ++q) theta[q] = -0.5 + 0.1 * q;
      for (int method = 0; method <= 1; ++method) {
        mc.method = method;
        CHECK(orc_moment_sim(m, 2, theta, 2, &mc, REF_IS_CHASE, cap, w_ref, 5, 6, 1, stats, moms));
      }
      double out[11 * 5 * 5], ss[9];
      int64_t ts[55];
      orc_recover(m, theta, &mc, out);
      orc_recover_steps(m, theta, &mc, out, ts);
      if (orc_transient(m, theta, &mc, ss) < 0) fails++;
      mc.transient_mode = 1;
      orc_recover_direct(m, theta, &mc, out);
      mc.transient_mode = 0;
      printf("moments model %d: %.6g\n", m, stats[0]);
    }
  }
  /* errors and acceptance */
  {
    enum { N = 30, G = 7, KS = 53 };
    const int32_t parts[7] = {5, 5, 5, 5, 11, 11, 11};
    double s[N * KS], d[G * KS], se[G * KS], err[N * G];
    for (int q = 0; q < N * KS; ++q) s[q] = 0.1 * (q % 17);
    for (int q = 0; q < G * KS; ++q) { d[q] = 0.13 * (q % 11); se[q] = (q % 5) ? 0.1 : 0.0; }
    orc_errors(N, s, KS, G, d, se, parts, 7, 10.0, err, 1);
    int64_t idx[N];
    double eo[N];
    int64_t a = 0;
    for (int g = 0; g < G; ++g) a += orc_accept_gene(err + (size_t)N * g, N, 4.8, idx, eo);
    printf("errors: %lld accepted\n", (long long)a);
  }
  /* model probabilities and data-side statistics */
  {
    int64_t counts[3 * 2] = {0, 5, 100, 0, 0, 40};
    double prob[6], lb[6], ub[6];
    CHECK(orc_model_probs(3, 2, counts, 11, 50, 0.95, prob, lb, ub, NULL));
    const double v[4] = {0.1, 0.2, 0.4, 0.8};
    const double q = orc_jl_quantile(v, 4, 0.95);
    enum { NC = 120, GN = 3 };
    uint32_t u[NC * GN], l[NC * GN];
    int32_t age[NC], cond[NC], pulse[NC], chase[NC];
    int64_t np = 0, nch = 0;
    for (int c = 0; c < NC; ++c) {
      age[c] = c % 5;
      cond[c] = (c % 12) - 1;
      if (cond[c] >= 0 && cond[c] <= 5) pulse[np++] = c;
      if (cond[c] >= 6) chase[nch++] = c;
    }
    for (int q2 = 0; q2 < NC * GN; ++q2) { u[q2] = (uint32_t)(q2 % 9); l[q2] = (uint32_t)(q2 % 4); }
    double wp[5 * 11];
    for (int q2 = 0; q2 < 55; ++q2) wp[q2] = 0.2;
    double out[GN * (8 * 5 + 6 * 11)], boot[GN * 10 * (4 * 5 + 3 * 11)];
    CHECK(orc_data_stats(GN, NC, u, l, age, cond, 5, 11, pulse, np, chase, nch, wp, 10, 3, out,
                         boot, 1));
    int64_t pos[NC];
    orc_data_resample(1, 2, 3, NC, 3, pos);
    printf("model probs %.6g, quantile %.6g, data stats %.6g\n", prob[1], q, out[0]);
  }
  if (fails) fprintf(stderr, "%d oracle calls failed\n", fails);
  return fails ? 1 : 0;
}
