/*
 * orc_data.c -- restatement of the data-side summary statistics and their
 * bootstrap standard errors (TEST INFRASTRUCTURE ONLY): the CPU replay of the
 * engine's data_stats_kernel.
 *
 *   get_mean_ccp_data / get_ff_ccp_data ...... scripts/data_summary_statistics.jl:2-32
 *   get_ccp_se ................................ scripts/data_summary_statistics.jl:40-58
 *   get_ratios / get_cond_specific_ratios ...... scripts/data_summary_statistics.jl:61-77
 *   get_ratio_se .............................. scripts/data_summary_statistics.jl:79-97
 *   get_correlations / ..._specific_correlations scripts/data_summary_statistics.jl:100-147
 *   get_correlation_se ........................ scripts/data_summary_statistics.jl:149-177
 *   weighted_cov .............................. scripts/data_summary_statistics.jl:179-181
 *   get_summary_stats ......................... scripts/data_summary_statistics.jl:183-196
 *
 * Per gene, from integer counts u, l of n cells with an age cluster (0..nc-1)
 * and a condition (0..J-1, or -1 for none):
 *   pulse (chase) mean and Fano factor of u + l per cluster over the pulse
 *   (chase) cells; labelled ratio per condition; mean_corr and corr_mean per
 *   condition, weighted by the caller's age x condition shares.
 * Standard errors: n_boot resamples with replacement, four independent ones
 * per gene and resample as in the reference (f = 0 the pulse cells, 1 the
 * chase cells, 2 all cells for the ratios, 3 all cells for the
 * correlations, whose weights are that resample's cluster shares within
 * each condition), then sqrt((1 / (B - 1)) sum_b (x_b - mean x)^2).  Draw r
 * of resample b of family f for gene g is Philox4x32-10 with key (seed_lo,
 * seed_hi ^ 0x2545F491) and counter (r / 2, b, g, f): words (0,1) or (2,3) as
 * a 64-bit u, position floor(u * m / 2^64) in the family's m cells.  The
 * reference draws from Julia's unseeded global RNG, so only the law can match.
 *
 * Sums over cells are exact integers; variances and covariances are
 * (n S_xy - S_x S_y) / (n (n - 1)) from 128-bit integers, rounded once, and
 * are NaN for a single cell (as Julia's var).  Missing clusters give 0, the
 * Fano factor adds 1e-4 to the mean only when it is 0, a condition without
 * cells gives ratio 0, and the floating-point sums run in index order.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KEY_SALT 0x2545F491u

typedef unsigned __int128 u128;

static int64_t resample_pos(uint64_t g, int b, int f, int64_t r, int64_t m, const uint32_t key[2]) {
  const uint32_t c[4] = {(uint32_t)(r >> 1), (uint32_t)b, (uint32_t)g, (uint32_t)f};
  uint32_t w[4];
  orc_philox4x32_10(c, key, w);
  const uint64_t u = (r & 1) == 0 ? ((uint64_t)w[1] << 32 | w[0]) : ((uint64_t)w[3] << 32 | w[2]);
  return (int64_t)(((u128)u * (uint64_t)m) >> 64);
}

void orc_data_resample(int64_t g, int b, int f, int64_t m, uint64_t seed, int64_t* pos) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ KEY_SALT};
  for (int64_t r = 0; r < m; ++r) pos[r] = resample_pos((uint64_t)g, b, f, r, m, key);
}

/* (n Sxy - Sx Sy) / (n (n - 1)) */
static double cov_int(int64_t n, uint64_t Sx, uint64_t Sy, uint64_t Sxy) {
  const u128 a = (u128)(uint64_t)n * Sxy, b = (u128)Sx * Sy;
  const double den = (double)n * (double)(n - 1);
  if (a >= b) return (double)(a - b) / den;
  return -((double)(b - a) / den);
}

/* sum of w .* ((x .- sum(w .* x)) .* (y .- sum(w .* y))) in index order */
static double wcov(const double* x, const double* y, const double* w, int n) {
  double mx = w[0] * x[0], my = w[0] * y[0];
  for (int i = 1; i < n; ++i) mx = mx + w[i] * x[i];
  for (int i = 1; i < n; ++i) my = my + w[i] * y[i];
  double acc = w[0] * ((x[0] - mx) * (y[0] - my));
  for (int i = 1; i < n; ++i) acc = acc + w[i] * ((x[i] - mx) * (y[i] - my));
  return acc;
}

typedef struct {
  int64_t cnt[ORC_DATA_MAX_CLUSTERS];
  uint64_t S[ORC_DATA_MAX_CLUSTERS], SS[ORC_DATA_MAX_CLUSTERS];
} ccp_acc;

typedef struct {
  int64_t cnt[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS];
  uint64_t Su[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS], Sl[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS];
  uint64_t Suu[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS], Sll[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS];
  uint64_t Sul[ORC_DATA_MAX_CONDS][ORC_DATA_MAX_CLUSTERS];
} corr_acc;

/* mean and Fano factor per cluster (get_mean_ccp_data, get_ff_ccp_data) */
static void ccp_stats(const ccp_acc* a, int nc, double* mean, double* ff) {
  for (int c = 0; c < nc; ++c) {
    if (a->cnt[c] == 0) { mean[c] = 0.0; ff[c] = 0.0; continue; }
    const double m = (double)a->S[c] / (double)a->cnt[c];
    const double v = cov_int(a->cnt[c], a->S[c], a->S[c], a->SS[c]);
    const double eps = m == 0.0 ? 0.0001 : 0.0;
    mean[c] = m;
    ff[c] = v / (m + eps);
  }
}

/* get_cond_specific_ratios */
static double ratio_of(int64_t cnt, uint64_t Su, uint64_t Sl) {
  if (cnt == 0) return 0.0;
  const double mu = (double)Su / (double)cnt, ml = (double)Sl / (double)cnt;
  return mu + ml > 0.0 ? ml / (mu + ml) : 0.0;
}

/* get_cond_specific_correlations for condition j with weights w[nc] */
static void corr_stats(const corr_acc* a, int j, int nc, const double* w, double* mean_corr,
                       double* corr_mean) {
  double m1[ORC_DATA_MAX_CLUSTERS] = {0}, m2[ORC_DATA_MAX_CLUSTERS] = {0},
         v1[ORC_DATA_MAX_CLUSTERS] = {0}, v2[ORC_DATA_MAX_CLUSTERS] = {0},
         c12[ORC_DATA_MAX_CLUSTERS] = {0};
  int any_cov = 0;
  for (int c = 0; c < nc; ++c) {
    const int64_t n = a->cnt[j][c];
    if (n == 0) {
      m1[c] = m2[c] = v1[c] = v2[c] = c12[c] = 0.0;
    } else {
      m1[c] = (double)a->Su[j][c] / (double)n;
      m2[c] = (double)a->Sl[j][c] / (double)n;
      v1[c] = cov_int(n, a->Su[j][c], a->Su[j][c], a->Suu[j][c]);
      v2[c] = cov_int(n, a->Sl[j][c], a->Sl[j][c], a->Sll[j][c]);
      c12[c] = cov_int(n, a->Su[j][c], a->Sl[j][c], a->Sul[j][c]);
    }
    if (!(c12[c] == 0.0)) any_cov = 1;
  }
  double s1 = w[0] * v1[0], s2 = w[0] * v2[0], s12 = w[0] * c12[0];
  for (int c = 1; c < nc; ++c) {
    s1 = s1 + w[c] * v1[c];
    s2 = s2 + w[c] * v2[c];
    s12 = s12 + w[c] * c12[c];
  }
  const double tv1 = s1 + wcov(m1, m1, w, nc), tv2 = s2 + wcov(m2, m2, w, nc);
  const int nz = !(tv1 == 0.0) && !(tv2 == 0.0);
  *mean_corr = (any_cov && nz) ? s12 / sqrt(fabs(tv1 * tv2)) : 0.0;
  *corr_mean = nz ? wcov(m1, m2, w, nc) / sqrt(fabs(tv1 * tv2)) : 0.0;
}

/* the statistics of one resample (b >= 0) or of the data themselves (b < 0):
 * st[0..nc) pulse mean, [nc..2nc) pulse FF, [2nc..3nc) chase mean,
 * [3nc..4nc) chase FF, then J ratios, J mean_corr, J corr_mean */
static void one_set(int64_t g, int b, int64_t n, const uint32_t* u, const uint32_t* l,
                    const int32_t* age, const int32_t* cond, int nc, int J, const int32_t* pulse,
                    int64_t np, const int32_t* chase, int64_t nch, const double* w_point,
                    const uint32_t key[2], double* st) {
  for (int f = 0; f < 2; ++f) {
    const int32_t* list = f == 0 ? pulse : chase;
    const int64_t m = f == 0 ? np : nch;
    ccp_acc a;
    memset(&a, 0, sizeof a);
    for (int64_t r = 0; r < m; ++r) {
      const int64_t i = list[b < 0 ? r : resample_pos((uint64_t)g, b, f, r, m, key)];
      const uint64_t x = (uint64_t)u[i] + l[i];
      const int c = age[i];
      a.cnt[c]++;
      a.S[c] += x;
      a.SS[c] += x * x;
    }
    ccp_stats(&a, nc, st + 2 * f * nc, st + (2 * f + 1) * nc);
  }
  {
    int64_t cnt[ORC_DATA_MAX_CONDS] = {0};
    uint64_t Su[ORC_DATA_MAX_CONDS] = {0}, Sl[ORC_DATA_MAX_CONDS] = {0};
    for (int64_t r = 0; r < n; ++r) {
      const int64_t i = b < 0 ? r : resample_pos((uint64_t)g, b, 2, r, n, key);
      const int j = cond[i];
      if (j < 0) continue;
      cnt[j]++;
      Su[j] += u[i];
      Sl[j] += l[i];
    }
    for (int j = 0; j < J; ++j) st[4 * nc + j] = ratio_of(cnt[j], Su[j], Sl[j]);
  }
  {
    corr_acc *a = (corr_acc*)calloc(1, sizeof(corr_acc));
    for (int64_t r = 0; r < n; ++r) {
      const int64_t i = b < 0 ? r : resample_pos((uint64_t)g, b, 3, r, n, key);
      const int j = cond[i];
      if (j < 0) continue;
      const int c = age[i];
      const uint64_t x = u[i], y = l[i];
      a->cnt[j][c]++;
      a->Su[j][c] += x;
      a->Sl[j][c] += y;
      a->Suu[j][c] += x * x;
      a->Sll[j][c] += y * y;
      a->Sul[j][c] += x * y;
    }
    for (int j = 0; j < J; ++j) {
      double w[ORC_DATA_MAX_CLUSTERS];
      if (b < 0) {
        for (int c = 0; c < nc; ++c) w[c] = w_point[c + (int64_t)nc * j];
      } else {
        int64_t tot = 0;
        for (int c = 0; c < nc; ++c) tot += a->cnt[j][c];
        for (int c = 0; c < nc; ++c) w[c] = tot > 0 ? (double)a->cnt[j][c] / (double)tot : 0.0;
      }
      corr_stats(a, j, nc, w, st + 4 * nc + J + j, st + 4 * nc + 2 * J + j);
    }
    free(a);
  }
}

int orc_data_stats(int64_t G, int64_t n, const uint32_t* u, const uint32_t* l, const int32_t* age,
                   const int32_t* cond, int nc, int J, const int32_t* pulse, int64_t np,
                   const int32_t* chase, int64_t nch, const double* w_point, int n_boot,
                   uint64_t seed, double* out, double* boot, int n_threads) {
  if (G < 0 || n < 0 || nc < 1 || nc > ORC_DATA_MAX_CLUSTERS || J < 1 || J > ORC_DATA_MAX_CONDS ||
      n_boot < 2 || np < 0 || nch < 0)
    return -1;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ KEY_SALT};
  const int K1 = 4 * nc + 3 * J;          /* statistics per set */
  (void)n_threads;
#pragma omp parallel for schedule(dynamic) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t g = 0; g < G; ++g) {
    const uint32_t* ug = u + g * n;
    const uint32_t* lg = l + g * n;
    double* pt = (double*)malloc(sizeof(double) * K1);
    double* bs = (double*)malloc(sizeof(double) * K1 * (size_t)n_boot);
    one_set(g, -1, n, ug, lg, age, cond, nc, J, pulse, np, chase, nch, w_point, key, pt);
    for (int b = 0; b < n_boot; ++b)
      one_set(g, b, n, ug, lg, age, cond, nc, J, pulse, np, chase, nch, w_point, key,
              bs + (size_t)b * K1);
    if (boot) memcpy(boot + (size_t)g * K1 * n_boot, bs, sizeof(double) * K1 * (size_t)n_boot);
    double se[4 * ORC_DATA_MAX_CLUSTERS + 3 * ORC_DATA_MAX_CONDS];
    for (int k = 0; k < K1; ++k) {
      double s = bs[k];
      for (int b = 1; b < n_boot; ++b) s = s + bs[(size_t)b * K1 + k];
      const double m = s / (double)n_boot;
      double q = (bs[k] - m) * (bs[k] - m);
      for (int b = 1; b < n_boot; ++b) {
        const double d = bs[(size_t)b * K1 + k] - m;
        q = q + d * d;
      }
      se[k] = sqrt((1.0 / (double)(n_boot - 1)) * q);
    }
    /* columns (G-major): pulse_mean, pulse_ff, pulse_mean_se, pulse_ff_se,
     * chase_mean, chase_ff, chase_mean_se, chase_ff_se (nc each), ratio,
     * ratio_se, mean_corr, mean_corr_se, corr_mean, corr_mean_se (J each) */
    int col = 0;
    for (int f = 0; f < 2; ++f) {
      for (int c = 0; c < nc; ++c) out[g + G * (col + c)] = pt[2 * f * nc + c];
      col += nc;
      for (int c = 0; c < nc; ++c) out[g + G * (col + c)] = pt[(2 * f + 1) * nc + c];
      col += nc;
      for (int c = 0; c < nc; ++c) out[g + G * (col + c)] = se[2 * f * nc + c];
      col += nc;
      for (int c = 0; c < nc; ++c) out[g + G * (col + c)] = se[(2 * f + 1) * nc + c];
      col += nc;
    }
    for (int q = 0; q < 3; ++q) {
      for (int j = 0; j < J; ++j) out[g + G * (col + j)] = pt[4 * nc + q * J + j];
      col += J;
      for (int j = 0; j < J; ++j) out[g + G * (col + j)] = se[4 * nc + q * J + j];
      col += J;
    }
    free(pt);
    free(bs);
  }
  return 0;
}
