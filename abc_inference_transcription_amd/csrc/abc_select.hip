// abc_select.hip -- model probabilities with bootstrap bounds per gene, the
// step after acceptance (scripts/model_probs.jl:1-63, constant_model_probs.jl,
// non_constant_model_probs.jl: get_model_probs and its per-gene branches).
//
// With l_m accepted particles in model group m and S = sum l: no group
// accepted gives zeros, one group gives 1 for it, and otherwise prob = l / S
// and n_boot resamples of the S labels with replacement.  model_prob_kernel
// spreads a gene's resamples over workgroups (rpb each) and stores every
// resample's share of each group; model_prob_bounds_kernel sorts them and
// takes Julia's type-7 quantiles at 1 - alpha and alpha.
// Draw i of resample b of gene g uses Philox4x32-10 with key (seed_lo,
// seed_hi ^ 0x5851F42D) and counter (i / 2, b, g, 0), words (0,1) or (2,3) as
// a 64-bit u, and takes label floor(u * S / 2^64) (the reference samples from
// Julia's unseeded global RNG, so only the law can match).  Counts are
// integers and every floating-point step is fixed, so the result equals the
// CPU replay bit for bit whatever the launch shape.
#include "abc_device.h"
#include "abc_kernels.h"

using namespace abcd;

namespace {

ABC_DEV double jl_quantile_sorted(const double* v, int n, double p) {
  if (n == 1) return v[0];
  const double m = 1.0 - p;
  const double aleph = (double)n * p + m;
  int j = (int)aleph;
  if (j < 1) j = 1;
  if (j > n - 1) j = n - 1;
  double g = aleph - (double)j;
  if (g < 0.0) g = 0.0;
  if (g > 1.0) g = 1.0;
  const double a = v[j - 1], b = v[j];
  return a + g * (b - a);
}

}  // namespace

// Workgroup (x, y): gene g0 + x, resamples y * rpb .. y * rpb + rpb - 1;
// each resample's share of every group goes to shares[(x M + m) n_boot + b].
// A gene with at most one accepting group is written here by y = 0.
__global__ __launch_bounds__(ABC_SEL_WG) void model_prob_kernel(
    int64_t G, int64_t g0, int M, const int64_t* __restrict__ counts, uint32_t k0, uint32_t k1,
    int n_boot, int rpb, double* __restrict__ shares, double* __restrict__ prob,
    double* __restrict__ lb, double* __restrict__ ub) {
  __shared__ unsigned long long cnt[ABC_SEL_MAX_GROUPS];
  const int64_t gl = blockIdx.x, g = g0 + gl;
  const int tid = threadIdx.x;
  int64_t l[ABC_SEL_MAX_GROUPS];
  int64_t S = 0;
  int nz = 0, which = -1;
#pragma unroll
  for (int m = 0; m < ABC_SEL_MAX_GROUPS; ++m) {
    l[m] = m < M ? counts[g + G * m] : 0;
    S += l[m];
    if (l[m] > 0) {
      nz++;
      which = m;
    }
  }
  if (nz <= 1) {  // uniform over the workgroup
    if (blockIdx.y == 0 && tid < M) {
      const double v = tid == which ? 1.0 : 0.0;
      prob[g + G * tid] = v;
      lb[g + G * tid] = v;
      ub[g + G * tid] = v;
    }
    return;
  }
  // S < 2^32 (the caller checks it), so with u = hi 2^32 + lo a label
  // floor(u S / 2^64) = floor((hi S + floor(lo S / 2^32)) / 2^32): the low
  // word of lo S cannot carry the sum past a multiple of 2^32
  const uint32_t S32 = (uint32_t)S;
  // cumulative counts A[k] = l_0 + ... + l_k: a draw's label is the number of
  // them its index reaches (a prefix, since A is nondecreasing)
  uint32_t A[ABC_SEL_MAX_GROUPS - 1];
  {
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < ABC_SEL_MAX_GROUPS - 1; ++k) {
      acc += (uint32_t)l[k];
      A[k] = acc;
    }
  }
  const int64_t pairs = (S + 1) / 2;
  const int b_lo = (int)blockIdx.y * rpb;
  const int b_hi = b_lo + rpb < n_boot ? b_lo + rpb : n_boot;
  for (int b = b_lo; b < b_hi; ++b) {
    if (tid < M) cnt[tid] = 0ull;
    __syncthreads();
    // this lane's draws, and how many of them reach A[k]: group m has
    // r[m - 1] - r[m] of them (the draws for r[-1], none for r[M - 1])
    uint32_t nd = 0u, r[ABC_SEL_MAX_GROUPS - 1];
#pragma unroll
    for (int k = 0; k < ABC_SEL_MAX_GROUPS - 1; ++k) r[k] = 0u;
    for (int64_t q = tid; q < pairs; q += blockDim.x) {
      const u4 w = philox((uint32_t)q, (uint32_t)b, (uint32_t)g, 0u, k0, k1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (2 * q + h >= S) break;
        const uint32_t lo = h == 0 ? w.x : w.z, hi = h == 0 ? w.y : w.w;
        const uint32_t idx = (uint32_t)(((uint64_t)hi * S32 + __umulhi(lo, S32)) >> 32);
        nd++;
#pragma unroll
        for (int k = 0; k < ABC_SEL_MAX_GROUPS - 1; ++k)
          if (k < M - 1) r[k] += idx >= A[k] ? 1u : 0u;
      }
    }
#pragma unroll
    for (int m = 0; m < ABC_SEL_MAX_GROUPS; ++m) {
      if (m < M) {
        const uint32_t reach = m == 0 ? nd : r[m > 0 ? m - 1 : 0];
        const uint32_t cm = reach - (m < M - 1 ? r[m < ABC_SEL_MAX_GROUPS - 1 ? m : 0] : 0u);
        if (cm) atomicAdd(&cnt[m], (unsigned long long)cm);
      }
    }
    __syncthreads();
    if (tid < M) shares[(gl * M + tid) * n_boot + b] = (double)cnt[tid] / (double)S;
    __syncthreads();
  }
}

// One workgroup per gene g0 + x with two or more accepting groups: each
// group's n_boot shares in ascending order (value i goes to place #{j: x_j <
// x_i} + #{j < i: x_j == x_i}; shares are finite and non-negative, so this is
// the order of any sort), then prob and Julia's type-7 quantiles at 1 - alpha
// and alpha.
__global__ __launch_bounds__(ABC_SEL_WG) void model_prob_bounds_kernel(
    int64_t G, int64_t g0, int M, const int64_t* __restrict__ counts, int n_boot, double alpha,
    const double* __restrict__ shares, double* __restrict__ prob, double* __restrict__ lb,
    double* __restrict__ ub) {
  __shared__ double x[ABC_SEL_MAX_GROUPS * ABC_SEL_MAX_BOOT];
  __shared__ double srt[ABC_SEL_MAX_GROUPS * ABC_SEL_MAX_BOOT];
  const int64_t gl = blockIdx.x, g = g0 + gl;
  const int tid = threadIdx.x;
  int64_t S = 0;
  int nz = 0;
  for (int m = 0; m < M; ++m) {
    const int64_t v = counts[g + G * m];
    S += v;
    nz += v > 0 ? 1 : 0;
  }
  if (nz <= 1) return;  // written by model_prob_kernel
  const double* sh = shares + gl * M * (int64_t)n_boot;
  for (int i = tid; i < M * n_boot; i += ABC_SEL_WG) x[i] = sh[i];
  __syncthreads();
  for (int i = tid; i < M * n_boot; i += ABC_SEL_WG) {
    const int m = i / n_boot, k = i - m * n_boot;
    const double* col = x + m * n_boot;
    const double v = col[k];
    int rank = 0;
    for (int j = 0; j < n_boot; ++j) rank += (col[j] < v || (col[j] == v && j < k)) ? 1 : 0;
    srt[m * n_boot + rank] = v;
  }
  __syncthreads();
  if (tid < M) {
    const double* col = srt + tid * n_boot;
    prob[g + G * tid] = (double)counts[g + G * tid] / (double)S;
    lb[g + G * tid] = jl_quantile_sorted(col, n_boot, 1.0 - alpha);
    ub[g + G * tid] = jl_quantile_sorted(col, n_boot, alpha);
  }
}
