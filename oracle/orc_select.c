/*
 * orc_select.c -- restatement of the reference's model probabilities (TEST
 * INFRASTRUCTURE ONLY): the CPU replay of the engine's bootstrap.
 *
 *   get_model_probs ........ scripts/model_probs.jl:1-29,
 *                            scripts/constant_model_probs.jl:1-28,
 *                            scripts/non_constant_model_probs.jl:1-30
 *   per-gene branches ...... scripts/model_probs.jl:38-54
 *   quantile ............... Julia Statistics.quantile, default (type 7)
 *
 * Per gene: l_m = number of particles of model group m with error <= eps.
 * No group accepted: probabilities and bounds 0.  One group: 1 for it.
 * Otherwise prob = l / sum(l), and n_boot bootstrap resamples draw sum(l)
 * labels with replacement; the bounds are the 1 - alpha and alpha quantiles of
 * each group's resampled share.  The reference draws from Julia's unseeded
 * global RNG; here draw i of resample b of gene g comes from Philox4x32-10
 * with key (seed_lo, seed_hi ^ 0x5851F42D) and counter (i / 2, b, g, 0):
 * words (0,1) or (2,3) as a 64-bit u, label index floor(u * S / 2^64).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Julia's quantile(v, p) on a sorted vector (alpha = beta = 1) */
static double jl_quantile_sorted(const double* v, int n, double p) {
  if (n == 1) return v[0];
  const double m = 1.0 - p;
  const double aleph = (double)n * p + m;
  int j = (int)aleph;  /* trunc */
  if (j < 1) j = 1;
  if (j > n - 1) j = n - 1;
  double g = aleph - (double)j;
  if (g < 0.0) g = 0.0;
  if (g > 1.0) g = 1.0;
  const double a = v[j - 1], b = v[j];
  return a + g * (b - a);
}

static int cmp_double(const void* x, const void* y) {
  const double a = *(const double*)x, b = *(const double*)y;
  return a < b ? -1 : (a > b ? 1 : 0);
}

int orc_model_probs(int64_t G, int M, const int64_t* counts, uint64_t seed, int n_boot,
                    double alpha, double* prob, double* lb, double* ub, int64_t* boot) {
  if (G < 0 || M < 1 || M > 8 || n_boot < 1 || n_boot > 256) return -1;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ 0x5851F42Du};
  double* fr = (double*)malloc(sizeof(double) * (size_t)n_boot * M);
  for (int64_t g = 0; g < G; ++g) {
    int64_t l[8], S = 0;
    int nz = 0, which = -1;
    for (int m = 0; m < M; ++m) {
      l[m] = counts[g + G * m];
      if (l[m] < 0) { free(fr); return -1; }
      S += l[m];
      if (l[m] > 0) { nz++; which = m; }
    }
    for (int m = 0; m < M; ++m) prob[g + G * m] = lb[g + G * m] = ub[g + G * m] = 0.0;
    if (nz == 0) continue;
    if (nz == 1) {
      prob[g + G * which] = lb[g + G * which] = ub[g + G * which] = 1.0;
      continue;
    }
    for (int m = 0; m < M; ++m) prob[g + G * m] = (double)l[m] / (double)S;
    for (int b = 0; b < n_boot; ++b) {
      int64_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      uint32_t w[4] = {0, 0, 0, 0};
      for (int64_t i = 0; i < S; ++i) {
        if ((i & 1) == 0) {
          const uint32_t c[4] = {(uint32_t)(i >> 1), (uint32_t)b, (uint32_t)g, 0u};
          orc_philox4x32_10(c, key, w);
        }
        const uint64_t u = (i & 1) == 0 ? ((uint64_t)w[1] << 32 | w[0]) : ((uint64_t)w[3] << 32 | w[2]);
        const uint64_t idx = (uint64_t)(((unsigned __int128)u * (uint64_t)S) >> 64);
        int64_t acc = 0;
        int m = 0;
        while (m < M - 1 && idx >= (uint64_t)(acc + l[m])) { acc += l[m]; m++; }
        cnt[m]++;
      }
      for (int m = 0; m < M; ++m) {
        fr[(size_t)m * n_boot + b] = (double)cnt[m] / (double)S;
        if (boot) boot[((size_t)g * n_boot + b) * M + m] = cnt[m];
      }
    }
    for (int m = 0; m < M; ++m) {
      double* col = fr + (size_t)m * n_boot;
      qsort(col, (size_t)n_boot, sizeof(double), cmp_double);
      lb[g + G * m] = jl_quantile_sorted(col, n_boot, 1.0 - alpha);
      ub[g + G * m] = jl_quantile_sorted(col, n_boot, alpha);
    }
  }
  free(fr);
  return 0;
}

/* Test hook: Julia's quantile on a sorted vector. */
double orc_jl_quantile(const double* v, int n, double p) { return jl_quantile_sorted(v, n, p); }
