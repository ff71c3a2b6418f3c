// abc_data.hip -- the data-side summary statistics and their bootstrap
// standard errors (scripts/data_summary_statistics.jl:1-196: get_ccp_data,
// get_ccp_se, get_ratios, get_ratio_se, get_correlations,
// get_correlation_se, weighted_cov, get_summary_stats).
//
// For each gene, from the integer counts u, l of the n cells (age cluster
// 0..nc-1, condition 0..J-1 or -1), the kernels form the statistics of the
// data themselves, then of n_boot resamples with replacement: four
// independent ones per resample as in the reference -- the pulse cells
// (mean and Fano factor of u + l per cluster), the chase cells (the same),
// all cells (labelled ratio per condition) and all cells again (mean_corr,
// corr_mean per condition, weighted by that resample's cluster shares within
// the condition; the data's own statistics use the caller's shares).  Draw r
// of resample b of family f for gene g is Philox4x32-10 with key (seed_lo,
// seed_hi ^ 0x2545F491) and counter (r / 2, b, g, f): words (0,1) or (2,3)
// as a 64-bit u, position floor(u * m / 2^64) among the family's m cells.
//
// Per set, the cells' sums go to LDS with 64-bit integer atomics (exact and
// independent of their order); then up to 2 nc + 2 J lanes finish the
// statistics, one cluster or condition each, in the CPU restatement's
// floating-point order: variances and covariances (n S_xy - S_x S_y) /
// (n (n - 1)) from 128-bit integers, NaN for one cell; missing clusters 0;
// the Fano factor adds 1e-4 to a zero mean; an empty condition has ratio 0.
// The standard error of each statistic is sqrt((1 / (B - 1)) sum_b (x_b -
// mean)^2), summed in resample order.  The result therefore equals the CPU
// replay bit for bit whatever the launch shape.
//
// data_stats_kernel below does this with one workgroup per gene and the
// per-draw atomics just described; it serves families of more than
// ABC_DATA_SORTED_MAX cells.  Smaller families (the reference data's 5,422
// cells) take data_pack_kernel (each gene's cells packed in group order),
// data_stats_sorted_kernel, which counts the draws per cell first and spreads
// a gene's sets over several workgroups, and data_se_kernel; both paths give
// the same integers and so the same bits.
#include "abc_device.h"
#include "abc_kernels.h"

using namespace abcd;

namespace {

// position among m cells from word pair h of a Philox block
ABC_DEV int64_t pos_of(const u4& w, int h, int64_t m) {
  const uint64_t u = h == 0 ? ((uint64_t)w.y << 32 | w.x) : ((uint64_t)w.w << 32 | w.z);
  return (int64_t)__umul64hi(u, (uint64_t)m);
}

// the same for m < 2^32: with u = hi 2^32 + lo, floor(u m / 2^64) =
// floor((hi m + floor(lo m / 2^32)) / 2^32), since the low word of lo m
// cannot carry the sum past a multiple of 2^64
ABC_DEV uint32_t pos_of32(const u4& w, int h, uint32_t m) {
  const uint32_t lo = h == 0 ? w.x : w.z, hi = h == 0 ? w.y : w.w;
  return (uint32_t)(((uint64_t)hi * m + __umulhi(lo, m)) >> 32);
}

// (n Sxy - Sx Sy) / (n (n - 1)) from exact 128-bit products
ABC_DEV double cov_int(int64_t n, uint64_t Sx, uint64_t Sy, uint64_t Sxy) {
  const uint64_t ahi = __umul64hi((uint64_t)n, Sxy), alo = (uint64_t)n * Sxy;
  const uint64_t bhi = __umul64hi(Sx, Sy), blo = Sx * Sy;
  const double den = (double)n * (double)(n - 1);
  const bool ge = ahi > bhi || (ahi == bhi && alo >= blo);
  if (ge) {
    const uint64_t lo = alo - blo, hi = ahi - bhi - (alo < blo ? 1ull : 0ull);
    return u128_to_f64(hi, lo) / den;
  }
  const uint64_t lo = blo - alo, hi = bhi - ahi - (blo < alo ? 1ull : 0ull);
  return -(u128_to_f64(hi, lo) / den);
}

constexpr int MC = ABC_DATA_MAX_CLUSTERS, MJ = ABC_DATA_MAX_CONDS;

struct DataLds {
  unsigned long long ccp[2][3][MC];      // pulse / chase: cells, S, SS of u + l per cluster
  unsigned long long rat[3][MJ];         // cells, S_u, S_l per condition
  unsigned long long cor[6][MJ][MC];     // cells, S_u, S_l, S_uu, S_ll, S_ul per (condition, cluster)
  double pt[4 * MC + 3 * MJ];            // the data's own statistics
  unsigned int bad;                      // a count of 2^20 or more was met
};

// The statistics of one set from its sums, one lane per cluster (pulse,
// chase) or condition (ratio, correlations); b < 0: the data's own set, whose
// correlations use the caller's weights w_point.
ABC_DEV void finish_set(const DataLds& S, int tid, int b, int nc, int J,
                        const double* __restrict__ w_point, double* st) {
  if (tid < 2 * nc) {
    const int f = tid / nc, c = tid - f * nc;
    const int64_t cnt = (int64_t)S.ccp[f][0][c];
    double mean = 0.0, ff = 0.0;
    if (cnt > 0) {
      const uint64_t s1 = S.ccp[f][1][c], s2 = S.ccp[f][2][c];
      mean = (double)s1 / (double)cnt;
      const double v = cov_int(cnt, s1, s1, s2);
      const double eps = mean == 0.0 ? 0.0001 : 0.0;
      ff = v / (mean + eps);
    }
    st[2 * f * nc + c] = mean;
    st[(2 * f + 1) * nc + c] = ff;
  } else if (tid < 2 * nc + J) {
    const int j = tid - 2 * nc;
    const int64_t cnt = (int64_t)S.rat[0][j];
    double r = 0.0;
    if (cnt > 0) {
      const double mu = (double)S.rat[1][j] / (double)cnt, ml = (double)S.rat[2][j] / (double)cnt;
      r = mu + ml > 0.0 ? ml / (mu + ml) : 0.0;
    }
    st[4 * nc + j] = r;
  } else if (tid < 2 * nc + 2 * J) {
    // one condition: the per-cluster means and weights are formed again from
    // the sums in a second pass instead of being held in arrays, whose
    // registers would bound every kernel that finishes sets
    const int j = tid - 2 * nc - J;
    int64_t tot = 0;
    for (int c = 0; c < nc; ++c) tot += (int64_t)S.cor[0][j][c];
    auto cluster = [&](int c, double& m1, double& m2, double& w) {
      const int64_t cnt = (int64_t)S.cor[0][j][c];
      m1 = cnt == 0 ? 0.0 : (double)S.cor[1][j][c] / (double)cnt;
      m2 = cnt == 0 ? 0.0 : (double)S.cor[2][j][c] / (double)cnt;
      w = b < 0 ? w_point[c + (int64_t)nc * j] : (tot > 0 ? (double)cnt / (double)tot : 0.0);
      return cnt;
    };
    // weighted within-cluster moments and the weighted means of the cluster
    // means (weighted_cov's first two sums)
    bool any_cov = false;
    double s1 = 0.0, s2 = 0.0, s12 = 0.0, mx1 = 0.0, mx2 = 0.0;
    for (int c = 0; c < nc; ++c) {
      double m1, m2, w;
      const int64_t cnt = cluster(c, m1, m2, w);
      double v1 = 0.0, v2 = 0.0, c12 = 0.0;
      if (cnt > 0) {
        const uint64_t su = S.cor[1][j][c], sl = S.cor[2][j][c];
        v1 = cov_int(cnt, su, su, S.cor[3][j][c]);
        v2 = cov_int(cnt, sl, sl, S.cor[4][j][c]);
        c12 = cov_int(cnt, su, sl, S.cor[5][j][c]);
      }
      if (!(c12 == 0.0)) any_cov = true;
      if (c == 0) {
        s1 = w * v1; s2 = w * v2; s12 = w * c12; mx1 = w * m1; mx2 = w * m2;
      } else {
        s1 = s1 + w * v1; s2 = s2 + w * v2; s12 = s12 + w * c12;
        mx1 = mx1 + w * m1; mx2 = mx2 + w * m2;
      }
    }
    // weighted_cov(m1, m1), (m2, m2), (m1, m2): the weighted products of deviations
    double q11 = 0.0, q22 = 0.0, q12 = 0.0;
    for (int c = 0; c < nc; ++c) {
      double m1, m2, w;
      cluster(c, m1, m2, w);
      const double d1 = m1 - mx1, d2 = m2 - mx2;
      if (c == 0) {
        q11 = w * (d1 * d1); q22 = w * (d2 * d2); q12 = w * (d1 * d2);
      } else {
        q11 = q11 + w * (d1 * d1); q22 = q22 + w * (d2 * d2); q12 = q12 + w * (d1 * d2);
      }
    }
    const double tv1 = s1 + q11, tv2 = s2 + q22;
    const bool nz = !(tv1 == 0.0) && !(tv2 == 0.0);
    st[4 * nc + J + j] = (any_cov && nz) ? s12 / sqrt(fabs(tv1 * tv2)) : 0.0;
    st[4 * nc + 2 * J + j] = nz ? q12 / sqrt(fabs(tv1 * tv2)) : 0.0;
  }
}

// Statistic k of gene g: its standard error over the n_boot sets in bg and
// the data's own value, to the output columns in the order of the 14 summary
// files: pulse_mean, pulse_ff, pulse_mean_se, pulse_ff_se, chase_* (nc each),
// ratio, ratio_se, mean_corr, mean_corr_se, corr_mean, corr_mean_se (J each).
ABC_DEV void se_column(int64_t g, int64_t G, int k, int n_boot, int nc, int J, double point,
                       const double* bg, double* __restrict__ out) {
  const int K1 = 4 * nc + 3 * J;
  double s = bg[k];
  for (int b = 1; b < n_boot; ++b) s = s + bg[(size_t)b * K1 + k];
  const double m = s / (double)n_boot;
  double q = (bg[k] - m) * (bg[k] - m);
  for (int b = 1; b < n_boot; ++b) {
    const double d = bg[(size_t)b * K1 + k] - m;
    q = q + d * d;
  }
  const double se = sqrt((1.0 / (double)(n_boot - 1)) * q);
  int colv, cols;
  if (k < 4 * nc) {  // pulse / chase block, mean or FF
    const int blk = k / nc, c = k - blk * nc;  // blk 0 p-mean, 1 p-ff, 2 c-mean, 3 c-ff
    const int f = blk / 2, kind = blk & 1;
    colv = 4 * nc * f + kind * nc + c;
    cols = colv + 2 * nc;
  } else {  // ratio, mean_corr, corr_mean
    const int kk = k - 4 * nc, qd = kk / J, j = kk - qd * J;
    colv = 8 * nc + 2 * J * qd + j;
    cols = colv + J;
  }
  out[g + G * (int64_t)colv] = point;
  out[g + G * (int64_t)cols] = se;
}

ABC_DEV void write_se(const DataLds& S, int tid, int64_t g, int64_t G, int n_boot, int nc, int J,
                      const double* bg, double* __restrict__ out) {
  const int K1 = 4 * nc + 3 * J;
  for (int k = tid; k < K1; k += ABC_DATA_WG) se_column(g, G, k, n_boot, nc, J, S.pt[k], bg, out);
}

}  // namespace

// genes g0 .. g0 + gridDim.x - 1 of G; boot holds this launch's resampled
// statistics, gridDim.x x n_boot x (4 nc + 3 J)
__global__ __launch_bounds__(ABC_DATA_WG) void data_stats_kernel(
    int64_t G, int64_t g0, int64_t n, const uint32_t* __restrict__ u, const uint32_t* __restrict__ l,
    const int32_t* __restrict__ age, const int32_t* __restrict__ cond, int nc, int J,
    const int32_t* __restrict__ pulse, int64_t np, const int32_t* __restrict__ chase, int64_t nch,
    const double* __restrict__ w_point, int n_boot, uint32_t k0, uint32_t k1,
    double* __restrict__ boot, double* __restrict__ out, unsigned int* __restrict__ bad_out) {
  __shared__ DataLds S;
  const int64_t g = g0 + blockIdx.x;
  const int tid = threadIdx.x;
  const uint32_t* ug = u + g * n;
  const uint32_t* lg = l + g * n;
  const int K1 = 4 * nc + 3 * J;
  double* bg = boot + (size_t)blockIdx.x * (size_t)n_boot * (size_t)K1;
  if (tid == 0) S.bad = 0u;
  for (int b = -1; b < n_boot; ++b) {
    unsigned long long* z = &S.ccp[0][0][0];
    for (int i = tid; i < 2 * 3 * MC + 3 * MJ + 6 * MJ * MC; i += ABC_DATA_WG) {
      if (i < 2 * 3 * MC) z[i] = 0ull;
      else if (i < 2 * 3 * MC + 3 * MJ) (&S.rat[0][0])[i - 2 * 3 * MC] = 0ull;
      else (&S.cor[0][0][0])[i - 2 * 3 * MC - 3 * MJ] = 0ull;
    }
    __syncthreads();
    // pulse and chase cells: u + l per age cluster
    for (int f = 0; f < 2; ++f) {
      const int32_t* list = f == 0 ? pulse : chase;
      const int64_t m = f == 0 ? np : nch;
      for (int64_t q = tid; 2 * q < m; q += ABC_DATA_WG) {
        u4 w{0, 0, 0, 0};
        if (b >= 0) w = philox((uint32_t)q, (uint32_t)b, (uint32_t)g, (uint32_t)f, k0, k1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t r = 2 * q + h;
          if (r >= m) break;
          const int64_t pos = b < 0 ? r : pos_of(w, h, m);
          const int32_t i = list[pos];
          if (i < 0 || i >= n) continue;
          const int c = age[i];
          if (c < 0 || c >= nc) continue;
          const uint32_t x0 = ug[i], x1 = lg[i];
          if ((x0 | x1) >= (1u << 20)) { atomicOr(&S.bad, 1u); continue; }
          const unsigned long long x = (unsigned long long)x0 + x1;
          atomicAdd(&S.ccp[f][0][c], 1ull);
          atomicAdd(&S.ccp[f][1][c], x);
          atomicAdd(&S.ccp[f][2][c], x * x);
        }
      }
    }
    // all cells, twice: ratios (f = 2) and correlations (f = 3)
    for (int f = 2; f < 4; ++f) {
      for (int64_t q = tid; 2 * q < n; q += ABC_DATA_WG) {
        u4 w{0, 0, 0, 0};
        if (b >= 0) w = philox((uint32_t)q, (uint32_t)b, (uint32_t)g, (uint32_t)f, k0, k1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t r = 2 * q + h;
          if (r >= n) break;
          const int64_t i = b < 0 ? r : pos_of(w, h, n);
          const int j = cond[i];
          if (j < 0 || j >= J) continue;
          const uint32_t x0 = ug[i], x1 = lg[i];
          if ((x0 | x1) >= (1u << 20)) { atomicOr(&S.bad, 1u); continue; }
          if (f == 2) {
            atomicAdd(&S.rat[0][j], 1ull);
            atomicAdd(&S.rat[1][j], (unsigned long long)x0);
            atomicAdd(&S.rat[2][j], (unsigned long long)x1);
          } else {
            const int c = age[i];
            if (c < 0 || c >= nc) continue;
            const unsigned long long x = x0, y = x1;
            atomicAdd(&S.cor[0][j][c], 1ull);
            atomicAdd(&S.cor[1][j][c], x);
            atomicAdd(&S.cor[2][j][c], y);
            atomicAdd(&S.cor[3][j][c], x * x);
            atomicAdd(&S.cor[4][j][c], y * y);
            atomicAdd(&S.cor[5][j][c], x * y);
          }
        }
      }
    }
    __syncthreads();
    finish_set(S, tid, b, nc, J, w_point, b < 0 ? S.pt : bg + (size_t)b * K1);
    __syncthreads();
  }
  write_se(S, tid, g, G, n_boot, nc, J, bg, out);
  if (tid == 0 && S.bad) atomicOr(bad_out, 1u);
}

// Each family's cells of one gene as one 64-bit word per cell: bits 0-19 u,
// 20-39 l, 40-55 the cell's position in the family (its draw counter), 56-63
// its group + 1 (0: a cell in no group, or padding).  Per gene the families
// follow one another (pulse, chase, all cells by condition, all cells by
// (condition, cluster)), each in ceil(m / 256) * 256 slots, transposed: slot
// j * 256 + t holds sorted position t * ceil(m / 256) + j, so that lane t of
// data_stats_sorted_kernel walks a contiguous run of the sorted cells while
// the 256 lanes read 256 consecutive words.  A count of 2^20 or more in a
// cell with a group sets *bad_out (the same cells the data's own set sums).
// Workgroup L packs block (item % (slots / 256)) of gene item / (slots /
// 256), item = (L % 8) * (gridDim.x / 8) + L / 8, so that one XCD reads
// each gene's counts.
__global__ __launch_bounds__(256) void data_pack_kernel(
    int64_t n, const uint32_t* __restrict__ u, const uint32_t* __restrict__ l, int64_t g0,
    const int32_t* __restrict__ pulse, int64_t np, const int32_t* __restrict__ chase, int64_t nch,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ grp, int64_t slots,
    int64_t genes, unsigned long long* __restrict__ xs, unsigned int* __restrict__ bad_out) {
  const int64_t item = (int64_t)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int64_t nblk = slots / 256;
  if (item >= genes * nblk) return;
  const int64_t gl = item / nblk, g = g0 + gl;
  const int64_t s = (item - gl * nblk) * 256 + threadIdx.x;
  const int64_t c0 = (np + 255) / 256, c1 = (nch + 255) / 256, c2 = (n + 255) / 256;
  int64_t f, at, m, base, cf;
  if (s < 256 * c0) { f = 0; at = s; m = np; base = 0; cf = c0; }
  else if (s < 256 * (c0 + c1)) { f = 1; at = s - 256 * c0; m = nch; base = np; cf = c1; }
  else if (s < 256 * (c0 + c1 + c2)) { f = 2; at = s - 256 * (c0 + c1); m = n; base = np + nch; cf = c2; }
  else { f = 3; at = s - 256 * (c0 + c1 + c2); m = n; base = np + nch + n; cf = c2; }
  const int64_t k = (at & 255) * cf + (at >> 8);
  unsigned long long word = 0ull;
  if (k < m) {
    const int gk = grp[base + k];
    if (gk >= 0) {
      const int32_t pos = perm[base + k];
      const int64_t i = f == 0 ? pulse[pos] : (f == 1 ? chase[pos] : pos);
      const uint32_t x0 = u[g * n + i], x1 = l[g * n + i];
      if ((x0 | x1) >= (1u << 20)) atomicOr(bad_out, 1u);
      else
        word = (unsigned long long)x0 | (unsigned long long)x1 << 20 |
               (unsigned long long)pos << 40 | (unsigned long long)(gk + 1) << 56;
    }
  }
  xs[gl * slots + s] = word;
}

// The same statistics for genes whose families hold at most
// ABC_DATA_SORTED_MAX cells, with the sets spread over workgroups: gene g0 +
// gl, sets sb * spb .. sb * spb + spb - 1 of 1 + n_boot (set 0: the data
// themselves, into pt; set 1 + b: resample b, into boot).  Workgroup L runs
// item (L % 8) * (gridDim.x / 8) + L / 8 of genes x nsb (gl = item / nsb, sb
// = item % nsb): workgroups go round-robin to the eight XCDs, so a gene's nsb
// workgroups share one XCD's L2 for its packed cells.
// Each resample first counts how often it draws each cell of a family
// (16-bit halves of 32-bit LDS words, atomics at random addresses; a count
// never exceeds the family's size), then every lane walks a contiguous run
// of the family's cells in group order (data_pack_kernel's words), sums
// count-weighted values in registers and adds them to the group's sums
// where its run changes group.  The integer sums are those of
// data_stats_kernel, regrouped, so the statistics are identical.
// data_se_kernel then forms the standard errors.  Eight waves per SIMD
// (64 VGPRs): the values spilled are kept per set, outside the draw and cell loops.
__global__ __launch_bounds__(ABC_DATA_WG) __attribute__((amdgpu_waves_per_eu(8))) void
data_stats_sorted_kernel(
    int64_t genes, int nsb, int64_t g0, const unsigned long long* __restrict__ xs, int64_t slots,
    int64_t n, int64_t np, int64_t nch, int nc, int J, const double* __restrict__ w_point,
    int n_boot, int spb, uint32_t k0, uint32_t k1, double* __restrict__ pt,
    double* __restrict__ boot) {
  extern __shared__ uint32_t mult_sh[];  // two 16-bit counters per word
  __shared__ DataLds S;
  const int64_t item = (int64_t)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (item >= genes * nsb) return;  // the whole workgroup: padding to a multiple of 8
  const int64_t gl = item / nsb;  // gene within this launch
  const int64_t g = g0 + gl;
  const int tid = threadIdx.x;
  const unsigned long long* xg = xs + gl * slots;
  const int K1 = 4 * nc + 3 * J;
  double* bg = boot + (size_t)gl * (size_t)n_boot * (size_t)K1;
  const int s_lo = (int)(item - gl * nsb) * spb;
  const int s_hi = s_lo + spb < n_boot + 1 ? s_lo + spb : n_boot + 1;
  for (int set = s_lo; set < s_hi; ++set) {
    const int b = set - 1;
    for (int i = tid; i < 2 * 3 * MC + 3 * MJ + 6 * MJ * MC; i += ABC_DATA_WG) {
      if (i < 2 * 3 * MC) (&S.ccp[0][0][0])[i] = 0ull;
      else if (i < 2 * 3 * MC + 3 * MJ) (&S.rat[0][0])[i - 2 * 3 * MC] = 0ull;
      else (&S.cor[0][0][0])[i - 2 * 3 * MC - 3 * MJ] = 0ull;
    }
    int64_t fb = 0;  // the family's first slot
    // unrolled: each family's cell walk is compiled with its own sums and no
    // branch on the family per cell
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int64_t m = f == 0 ? np : (f == 1 ? nch : n);
      const int64_t cf = (m + ABC_DATA_WG - 1) / ABC_DATA_WG;
      if (b >= 0) {
        for (int64_t i = tid; 2 * i < m; i += ABC_DATA_WG) mult_sh[i] = 0u;
        __syncthreads();
        for (int64_t q = tid; 2 * q < m; q += ABC_DATA_WG) {
          const u4 w = philox((uint32_t)q, (uint32_t)b, (uint32_t)g, (uint32_t)f, k0, k1);
          const uint32_t p0 = pos_of32(w, 0, (uint32_t)m);
          atomicAdd(&mult_sh[p0 >> 1], 1u << (16 * (p0 & 1)));
          if (2 * q + 1 < m) {
            const uint32_t p1 = pos_of32(w, 1, (uint32_t)m);
            atomicAdd(&mult_sh[p1 >> 1], 1u << (16 * (p1 & 1)));
          }
        }
      }
      __syncthreads();
      int cur = -1;
      unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
      auto flush = [&]() {
        if (cur < 0 || a0 == 0ull) return;
        if (f < 2) {
          atomicAdd(&S.ccp[f][0][cur], a0);
          atomicAdd(&S.ccp[f][1][cur], a1);
          atomicAdd(&S.ccp[f][2][cur], a2);
        } else if (f == 2) {
          atomicAdd(&S.rat[0][cur], a0);
          atomicAdd(&S.rat[1][cur], a1);
          atomicAdd(&S.rat[2][cur], a2);
        } else {
          const int j = cur / nc, c = cur - (cur / nc) * nc;
          atomicAdd(&S.cor[0][j][c], a0);
          atomicAdd(&S.cor[1][j][c], a1);
          atomicAdd(&S.cor[2][j][c], a2);
          atomicAdd(&S.cor[3][j][c], a3);
          atomicAdd(&S.cor[4][j][c], a4);
          atomicAdd(&S.cor[5][j][c], a5);
        }
      };
      const unsigned long long* xf = xg + fb + tid;
      for (int64_t j = 0; j < cf; ++j) {
        const unsigned long long v = xf[j * ABC_DATA_WG];
        const int gk = (int)(v >> 56) - 1;
        if (gk != cur) {
          flush();
          cur = gk;
          a0 = a1 = a2 = a3 = a4 = a5 = 0ull;
        }
        if (gk < 0) continue;
        const uint32_t pos = (uint32_t)(v >> 40) & 0xFFFFu;
        const uint32_t mult = b < 0 ? 1u : (mult_sh[pos >> 1] >> (16 * (pos & 1))) & 0xFFFFu;
        if (mult == 0u) continue;
        const unsigned long long x0 = v & 0xFFFFFull, x1 = (v >> 20) & 0xFFFFFull;
        const unsigned long long mm = mult;
        a0 += mm;
        if (f < 2) {
          const unsigned long long x = x0 + x1;
          a1 += mm * x;
          a2 += mm * (x * x);
        } else {
          a1 += mm * x0;
          a2 += mm * x1;
          if (f == 3) {
            a3 += mm * (x0 * x0);
            a4 += mm * (x1 * x1);
            a5 += mm * (x0 * x1);
          }
        }
      }
      flush();
      __syncthreads();
      fb += cf * ABC_DATA_WG;
    }
    finish_set(S, tid, b, nc, J, w_point, b < 0 ? pt + (size_t)gl * K1 : bg + (size_t)b * K1);
    __syncthreads();
  }
}

// Standard errors of data_stats_sorted_kernel's sets: one lane per (gene,
// statistic) of the launch, the same sums in resample order as write_se.
__global__ __launch_bounds__(256) void data_se_kernel(int64_t G, int64_t g0, int64_t genes, int nc,
                                                      int J, int n_boot,
                                                      const double* __restrict__ pt,
                                                      const double* __restrict__ boot,
                                                      double* __restrict__ out) {
  const int K1 = 4 * nc + 3 * J;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= genes * K1) return;
  const int64_t gl = q / K1;
  const int k = (int)(q - gl * K1);
  const double* bg = boot + (size_t)gl * (size_t)n_boot * (size_t)K1;
  se_column(g0 + gl, G, k, n_boot, nc, J, pt[gl * K1 + k], bg, out);
}
