// abc_select.hip -- model probabilities with bootstrap bounds per gene, the
// step after acceptance (scripts/model_probs.jl:1-63, constant_model_probs.jl,
// non_constant_model_probs.jl: get_model_probs and its per-gene branches).
//
// One workgroup per gene.  With l_m accepted particles in model group m and
// S = sum l: no group accepted gives zeros, one group gives 1 for it, and
// otherwise prob = l / S and n_boot resamples of the S labels with
// replacement; each resample's share of every group goes to LDS, and after
// sorting the bounds are Julia's type-7 quantiles at 1 - alpha and alpha.
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

__global__ __launch_bounds__(ABC_SEL_WG) void model_prob_kernel(
    int64_t G, int M, const int64_t* __restrict__ counts, uint32_t k0, uint32_t k1, int n_boot,
    double alpha, double* __restrict__ prob, double* __restrict__ lb, double* __restrict__ ub) {
  __shared__ double fr[ABC_SEL_MAX_GROUPS * ABC_SEL_MAX_BOOT];
  __shared__ unsigned long long cnt[ABC_SEL_MAX_GROUPS];
  const int64_t g = blockIdx.x;
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
    if (tid < M) {
      const double v = tid == which ? 1.0 : 0.0;
      prob[g + G * tid] = v;
      lb[g + G * tid] = v;
      ub[g + G * tid] = v;
    }
    return;
  }
  const uint64_t uS = (uint64_t)S;
  const int64_t pairs = (S + 1) / 2;
  for (int b = 0; b < n_boot; ++b) {
    if (tid < M) cnt[tid] = 0ull;
    __syncthreads();
    uint32_t c[ABC_SEL_MAX_GROUPS];
#pragma unroll
    for (int m = 0; m < ABC_SEL_MAX_GROUPS; ++m) c[m] = 0u;
    for (int64_t q = tid; q < pairs; q += blockDim.x) {
      const u4 w = philox((uint32_t)q, (uint32_t)b, (uint32_t)g, 0u, k0, k1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (2 * q + h >= S) break;
        const uint64_t u = h == 0 ? ((uint64_t)w.y << 32 | w.x) : ((uint64_t)w.w << 32 | w.z);
        const uint64_t idx = __umul64hi(u, uS);
        // label group: the number of cumulative counts (over the first M - 1
        // groups) that idx reaches
        uint64_t acc = 0;
        int m = 0;
#pragma unroll
        for (int k = 0; k < ABC_SEL_MAX_GROUPS - 1; ++k) {
          if (k < M - 1) {
            acc += (uint64_t)l[k];
            if (idx >= acc) m = k + 1;
          }
        }
#pragma unroll
        for (int k = 0; k < ABC_SEL_MAX_GROUPS; ++k) c[k] += (k == m);
      }
    }
#pragma unroll
    for (int m = 0; m < ABC_SEL_MAX_GROUPS; ++m)
      if (m < M && c[m]) atomicAdd(&cnt[m], (unsigned long long)c[m]);
    __syncthreads();
    if (tid < M) fr[tid * n_boot + b] = (double)cnt[tid] / (double)S;
    __syncthreads();
  }
  if (tid < M) {
    // insertion sort of this group's resampled shares
    double* col = fr + tid * n_boot;
    for (int i = 1; i < n_boot; ++i) {
      const double x = col[i];
      int j = i;
      while (j > 0 && col[j - 1] > x) {
        col[j] = col[j - 1];
        --j;
      }
      col[j] = x;
    }
    prob[g + G * tid] = (double)counts[g + G * tid] / (double)S;
    lb[g + G * tid] = jl_quantile_sorted(col, n_boot, 1.0 - alpha);
    ub[g + G * tid] = jl_quantile_sorted(col, n_boot, alpha);
  }
}
