/*
 * orc_errors.c -- reference-order FP64 errors and fixed-epsilon acceptance
 * (TEST INFRASTRUCTURE ONLY).
 *
 *   nlsqerror_part ........ scripts/compute_errors.jl:30-43
 *   compute_trunc_errors .. scripts/compute_errors.jl:45-70 (sum of parts,
 *                           clamp at 10)
 *   acceptance ............ scripts/accepted_particles.jl:20-22 (err <= eps,
 *                           stable sort by error)
 *
 * Julia evaluates x^2 for Float64 as x*x and sigma^2 with sigma = 0.1 as
 * 0.1*0.1; additions associate left to right.  Compiled with
 * -ffp-contract=off, this reproduces the reference bits.
 */
#include "oracle.h"
#include <stdlib.h>

double orc_error_pair(const double* d, const double* se, const double* s, int K,
                      const int32_t* parts, int n_parts, double clamp) {
  const double sigma = 0.1;
  const double s2 = sigma * sigma;
  double err = 0.0;
  int off = 0;
  for (int p = 0; p < n_parts; ++p) {
    double e = 0.0;
    for (int i = off; i < off + parts[p]; ++i) {
      double eps = (se[i] + d[i] != 0.0) ? 0.0 : 0.0001;
      double diff = d[i] - s[i];
      e += (diff * diff) / ((se[i] * se[i] + s2 * (d[i] * d[i])) + eps);
    }
    err += e / (double)K;
    off += parts[p];
  }
  if (err > clamp) err = clamp;
  return err;
}

void orc_errors(int64_t n, const double* stats, int K, int64_t G, const double* data,
                const double* se, const int32_t* parts, int n_parts, double clamp,
                double* err, int n_threads) {
  (void)n_threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    double s[256], dg[256], sg[256];
    for (int k = 0; k < K; ++k) s[k] = stats[i + n * k];
    for (int64_t g = 0; g < G; ++g) {
      for (int k = 0; k < K; ++k) { dg[k] = data[g + G * k]; sg[k] = se[g + G * k]; }
      err[i + n * g] = orc_error_pair(dg, sg, s, K, parts, n_parts, clamp);
    }
  }
}

typedef struct { double e; int64_t i; } pair_t;

static int cmp_pair(const void* a, const void* b) {
  const pair_t* x = (const pair_t*)a;
  const pair_t* y = (const pair_t*)b;
  if (x->e < y->e) return -1;
  if (x->e > y->e) return 1;
  return x->i < y->i ? -1 : (x->i > y->i);
}

int64_t orc_accept_gene(const double* err_col, int64_t n, double eps, int64_t* idx,
                        double* err_out) {
  pair_t* v = (pair_t*)malloc(sizeof(pair_t) * (n > 0 ? n : 1));
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i)
    if (err_col[i] <= eps) { v[m].e = err_col[i]; v[m].i = i; m++; }
  qsort(v, m, sizeof(pair_t), cmp_pair);
  for (int64_t q = 0; q < m; ++q) {
    idx[q] = v[q].i + 1;
    if (err_out) err_out[q] = v[q].e;
  }
  free(v);
  return m;
}
