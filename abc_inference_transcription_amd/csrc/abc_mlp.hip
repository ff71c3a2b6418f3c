// abc_mlp.hip -- the surrogate: an MLP from (model, log10 parameters) to the
// mean and log-variance of every summary statistic, trained on the simulator's
// own output, and the screen that lets the exact simulation run only on
// particles that can still be accepted.  No reference counterpart (the
// reference simulates every prior particle, scripts/abc_simulation.jl:88-97).
//
// Network: x (n_in, zero-padded to 64) -> [width, SiLU] x depth -> n_out
// (padded to a multiple of 64).  Every product is a bf16 "NT" GEMM
// C[m][n] = sum_k A[m][k] * B[n][k] on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation: forward Z = H W^T, backward dH = dZ W (W kept transposed as a
// second bf16 copy), dW = dZ^T H (epilogues store the transposed activations
// these need).  Master weights, gradients and Adam moments are FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "../../include/abc.h"
#include "abc_kernels.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

enum { EPI_F32 = 0, EPI_SILU = 1, EPI_DSILU = 2 };

// K per pipeline stage (two 32-deep MFMA steps) and the LDS row length
// (+8 bf16 of padding against bank conflicts)
constexpr int BK = 64, LDSK = BK + 8;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// 256 threads, workgroup tile BM x BN (BM 128 or 64), wave tile BM/2 x BN/2
// (2 x 2 waves).  blockIdx.z selects a K chunk [z*K, (z+1)*K) of rows with
// leading dimension ld (split-K: EPI_F32 then writes its partial tile to
// Cf + z*M*N).  K advances in stages of BK = 64 through two LDS buffers: while
// the MFMAs read one stage, the next is loaded from global memory into
// registers and stored to the other buffer, one barrier per stage.  Every
// output sums its K in order, 32 at a time, whatever BM.
template <int BM, int BN, int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(const bf16* __restrict__ A,
                                                      const bf16* __restrict__ B, int M, int N,
                                                      int K, int ld, const float* __restrict__ bias,
                                                      float* __restrict__ Cf,
                                                      bf16* __restrict__ Cb,
                                                      bf16* __restrict__ CbT,
                                                      const float* __restrict__ Zprev,
                                                      float* __restrict__ Zout) {
  constexpr int WN = BN / 2, NCB = WN / 16, NBQ = BN / 32;
  constexpr int WM = BM / 2, NRB = WM / 16, NAQ = BM / 32;
  __shared__ bf16 As[2][BM][LDSK];
  __shared__ bf16 Bs[2][BN][LDSK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * K, kend = kbeg + K;
  if (EPI == EPI_F32) Cf += (size_t)blockIdx.z * M * N;
  f32x4 acc[NRB][NCB];
#pragma unroll
  for (int a = 0; a < NRB; ++a)
#pragma unroll
    for (int b = 0; b < NCB; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // this thread's slice of a stage: rows c >> 3, 8 columns at (c & 7) * 8
  bf16x8 ra[NAQ], rb[NBQ];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NAQ; ++q) {
      const int c = tid + q * 256;
      ra[q] = *reinterpret_cast<const bf16x8*>(A + (size_t)(m0 + (c >> 3)) * ld + k0 + (c & 7) * 8);
    }
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int c = tid + q * 256;
      rb[q] = *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + (c >> 3)) * ld + k0 + (c & 7) * 8);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NAQ; ++q) {
      const int c = tid + q * 256;
      *reinterpret_cast<bf16x8*>(&As[buf][c >> 3][(c & 7) * 8]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int c = tid + q * 256;
      *reinterpret_cast<bf16x8*>(&Bs[buf][c >> 3][(c & 7) * 8]) = rb[q];
    }
  };
  load(kbeg);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) load(k0 + BK);  // in flight while the MFMAs below run
#pragma unroll
    for (int ks = 0; ks < BK / 32; ++ks) {
      bf16x8 af[NRB], bfr[NCB];
#pragma unroll
      for (int r4 = 0; r4 < NRB; ++r4)
        af[r4] = *reinterpret_cast<const bf16x8*>(
            &As[buf][wm * WM + r4 * 16 + (lane & 15)][ks * 32 + (lane >> 4) * 8]);
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
        bfr[cb] = *reinterpret_cast<const bf16x8*>(
            &Bs[buf][wn * WN + cb * 16 + (lane & 15)][ks * 32 + (lane >> 4) * 8]);
#pragma unroll
      for (int r4 = 0; r4 < NRB; ++r4)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
          acc[r4][cb] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[r4], bfr[cb], acc[r4][cb], 0, 0, 0);
    }
    if (more) stash(buf ^ 1);  // the other buffer: its last readers passed the barrier below
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int r4 = 0; r4 < NRB; ++r4)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      // this lane's 4 outputs are rows m..m+3 of column n: one 8-byte store
      // of the transposed copy (rows consecutive in C^T)
      const int mb = m0 + wm * WM + r4 * 16 + (lane >> 4) * 4;
      const int n = n0 + wn * WN + cb * 16 + (lane & 15);
      bf16x4 tv;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = mb + q;
        float v = acc[r4][cb][q];
        const size_t o = (size_t)m * N + n;
        if (EPI == EPI_F32) {
          if (bias) v += bias[n];
          Cf[o] = v;
        } else if (EPI == EPI_SILU) {
          const float z = v + bias[n];
          Zout[o] = z;
          const bf16 h = (bf16)(z * sigm(z));
          Cb[o] = h;
          tv[q] = h;
        } else {
          const float z = Zprev[o];
          const float s = sigm(z);
          const bf16 d = (bf16)(v * (s * (1.0f + z * (1.0f - s))));
          Cb[o] = d;
          tv[q] = d;
        }
      }
      if (EPI != EPI_F32) *reinterpret_cast<bf16x4*>(CbT + (size_t)n * M + mb) = tv;
    }
}

// C[i] = sum over the S split-K partials P[s][i], in the order s = 0, 1, ...
__global__ void splitk_reduce_kernel(const float* __restrict__ P, int S, int64_t MN,
                                     float* __restrict__ C) {
  // four consecutive outputs per thread (MN is a multiple of 4)
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i >= MN) return;
  float4 v = *reinterpret_cast<const float4*>(P + i);
  for (int s = 1; s < S; ++s) {
    const float4 w = *reinterpret_cast<const float4*>(P + (size_t)s * MN + i);
    v.x += w.x;
    v.y += w.y;
    v.z += w.z;
    v.w += w.w;
  }
  *reinterpret_cast<float4*>(C + i) = v;
}

// x (B x n_in FP32, row-major) -> X (B x in_pad bf16) and X^T (in_pad x B).
__global__ void cast_input_kernel(const float* __restrict__ x, int Bn, int Bpad, int n_in,
                                  int in_pad, bf16* __restrict__ X, bf16* __restrict__ XT) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)Bpad * in_pad) return;
  const int b = (int)(q / in_pad), k = (int)(q % in_pad);
  const float v = (b < Bn && k < n_in) ? x[(size_t)b * n_in + k] : 0.0f;
  X[q] = (bf16)v;
  XT[(size_t)k * Bpad + b] = (bf16)v;
}

// Gaussian negative log-likelihood per element, mean over B*ns.
// O row: [mu (ns) | logvar (ns) | pad]; dO and dO^T in bf16.  The first n_cnt
// statistics are counts: their mu slot is a log-rate o with the Poisson
// negative log-likelihood exp(o) - t o (log t! dropped) on [-30, 8], continued
// linearly outside (the gradient there is the bound's, exp(bound) - t, so an
// output that strays is pulled back and exp never overflows), and their
// logvar slot is unused.
// One workgroup per 64 x 64 tile of (sample b, output k): the gradient is
// written row-major to dO with k fastest, and through an LDS transpose to dO^T
// with b fastest, so both stores are coalesced; one loss atomic per tile.
__global__ __launch_bounds__(256) void nll_kernel(const float* __restrict__ O,
                                                  const float* __restrict__ T, int Bn, int Bpad,
                                                  int ns, int n_cnt, int out_pad,
                                                  bf16* __restrict__ dO, bf16* __restrict__ dOT,
                                                  float* __restrict__ loss) {
  __shared__ bf16 tile[64][64 + 2];
  __shared__ float red[256];
  const int b0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 columns x 4 rows per pass
  const float scale = 1.0f / ((float)Bn * (float)ns);
  float l = 0.0f;
  for (int rr = ty; rr < 64; rr += 4) {
    const int b = b0 + rr, k = k0 + tx;
    float g = 0.0f;
    if (b < Bn && k < 2 * ns) {
      const int s = k < ns ? k : k - ns;
      const float mu = O[(size_t)b * out_pad + s];
      const float lv_raw = O[(size_t)b * out_pad + ns + s];
      const bool clipped = lv_raw < -12.0f || lv_raw > 12.0f;
      const float lv = fminf(fmaxf(lv_raw, -12.0f), 12.0f);
      const float y = T[(size_t)b * ns + s];
      const float r = y - mu;
      const float inv = __expf(-lv);
      if (s < n_cnt) {
        if (k < ns) {
          const float o = fminf(fmaxf(mu, -30.0f), 8.0f);
          const float lam = __expf(o);
          g = (lam - y) * scale;
          l += ((lam - y * o) + (lam - y) * (mu - o)) * scale;
        }
      } else if (k < ns) {
        g = -r * inv * scale;
        l += 0.5f * (lv + r * r * inv) * scale;
      } else {
        g = clipped ? 0.0f : 0.5f * (1.0f - r * r * inv) * scale;
      }
    }
    const bf16 gb = (bf16)g;
    dO[(size_t)b * out_pad + k] = gb;
    tile[rr][tx] = gb;
  }
  __syncthreads();
  for (int kk = ty; kk < 64; kk += 4) dOT[(size_t)(k0 + kk) * Bpad + b0 + tx] = tile[tx][kk];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// bias gradient: row sums of dZ^T (rows x Bpad bf16), one workgroup per row:
// 16-byte loads (8 values), then a wave reduction and one across the 4 waves
__global__ __launch_bounds__(256) void rowsum_kernel(const bf16* __restrict__ dZT, int Bpad,
                                                     float* __restrict__ db) {
  const int r = blockIdx.x;
  const bf16x8* row = reinterpret_cast<const bf16x8*>(dZT + (size_t)r * Bpad);
  float s = 0.0f;
  for (int v = threadIdx.x; v < Bpad / 8; v += blockDim.x) {
    const bf16x8 x = row[v];
#pragma unroll
    for (int e = 0; e < 8; ++e) s += (float)x[e];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) db[r] = (red[0] + red[1]) + (red[2] + red[3]);
}

struct LayerDesc {
  int in, out;      // padded dims
  int64_t w_off;    // offset of W (out x in) in the flat parameter array
  int64_t b_off;    // offset of b (out)
};

// Adam on the flat FP32 parameters; refreshes the bf16 copies W (out x in)
// and W^T (in x out) of every weight.
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int64_t n, float lr, float b1, float b2,
                            float eps, float wd, float c1, float c2, const LayerDesc* __restrict__ L,
                            int nl, const int64_t* __restrict__ bf_off, bf16* __restrict__ Wb,
                            bf16* __restrict__ WbT) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float gg = g[p];
  const float mm = b1 * m[p] + (1.0f - b1) * gg;
  const float vv = b2 * v[p] + (1.0f - b2) * gg * gg;
  m[p] = mm;
  v[p] = vv;
  float ww = w[p];
  ww -= lr * ((mm / c1) / (sqrtf(vv / c2) + eps) + wd * ww);
  w[p] = ww;
  for (int l = 0; l < nl; ++l) {
    const int64_t sz = (int64_t)L[l].in * L[l].out;
    if (p >= L[l].w_off && p < L[l].w_off + sz) {
      const int64_t e = p - L[l].w_off;
      const int o = (int)(e / L[l].in), i = (int)(e % L[l].in);
      Wb[bf_off[l] + e] = (bf16)ww;
      WbT[bf_off[l] + (int64_t)i * L[l].out + o] = (bf16)ww;
    }
  }
}

__global__ void refresh_bf16_kernel(const float* __restrict__ w, const LayerDesc* __restrict__ L,
                                    int nl, const int64_t* __restrict__ bf_off,
                                    bf16* __restrict__ Wb, bf16* __restrict__ WbT, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int l = 0; l < nl; ++l) {
    const int64_t sz = (int64_t)L[l].in * L[l].out;
    if (p >= L[l].w_off && p < L[l].w_off + sz) {
      const int64_t e = p - L[l].w_off;
      const int o = (int)(e / L[l].in), i = (int)(e % L[l].in);
      Wb[bf_off[l] + e] = (bf16)w[p];
      WbT[bf_off[l] + (int64_t)i * L[l].out + o] = (bf16)w[p];
    }
  }
}

__global__ void zero_f32_kernel(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// copy the (mu, logvar) block of O (Bpad x out_pad) into y (B x 2 ns)
__global__ void gather_out_kernel(const float* __restrict__ O, int Bn, int ns, int out_pad,
                                  float* __restrict__ y) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)Bn * 2 * ns) return;
  const int b = (int)(q / (2 * ns)), k = (int)(q % (2 * ns));
  y[q] = O[(size_t)b * out_pad + k];
}

int g1(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- surrogate I/O
// Canonical features of a particle: one-hot(model) and its rates per bin,
// [kon(nb), koff, alpha(nb), gamma(nb), lambda, splice] in log10 / 3, with
// constant rates repeated over the bins (layout of scripts/model.jl:30-43).
__global__ void features_kernel(int model, int64_t n, const double* __restrict__ theta, int nb,
                                float* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nk = model == 3 ? nb : 1, na = model == 4 ? nb : 1, ng = model == 5 ? nb : 1;
  const int ok = 0, ooff = nk, oa = nk + 1, og = oa + na, ol = og + ng, os = ol + 1;
  const int nin = 5 + 3 * nb + 3;
  float* r = x + i * nin;
  for (int q = 0; q < 5; ++q) r[q] = (q == model - 1) ? 1.0f : 0.0f;
  int c = 5;
  for (int j = 0; j < nb; ++j) r[c++] = (float)(theta[i + n * (ok + (nk > 1 ? j : 0))] / 3.0);
  r[c++] = (float)(theta[i + n * ooff] / 3.0);
  for (int j = 0; j < nb; ++j) r[c++] = (float)(theta[i + n * (oa + (na > 1 ? j : 0))] / 3.0);
  for (int j = 0; j < nb; ++j) r[c++] = (float)(theta[i + n * (og + (ng > 1 ? j : 0))] / 3.0);
  r[c++] = (float)(theta[i + n * ol] / 3.0);
  r[c++] = (float)(theta[i + n * os] / 3.0);
}

// Targets: standardised asinh of every statistic (stats n x K column-major).
__global__ void targets_kernel(int64_t n, const double* __restrict__ stats, int K,
                               const float* __restrict__ mean, const float* __restrict__ sd,
                               float* __restrict__ t) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * K) return;
  const int64_t i = q / K;
  const int k = (int)(q % K);
  const double s = stats[i + n * k];
  float v = 0.0f;
  if (s == s && fabs(s) < 1e300) v = (float)((asinh(s) - (double)mean[k]) / (double)sd[k]);
  t[q] = v;
}

// Screen: the surrogate gives, per statistic, y ~ N(mu, sigma^2) in the
// standardised asinh space, s = sinh(mean_k + sd_k y).  Its central moments
// (E s, c2, c3, c4) come from 7-point Gauss-Hermite quadrature.  The
// statistics of different (bin, condition) groups come from disjoint cells,
// so the error err_g = sum_k W_gk (d_gk - s_k)^2 concentrates around
//   E err = sum W (delta^2 + c2),  Var err = sum W^2 (4 delta^2 c2 - 4 delta c3 + c4 - c2^2)
// with delta = d - E s.  A particle passes when E err - z sd err can reach a
// gene's threshold (or when its prediction is not finite).  Genes whose data
// are not finite have NaN errors and accept nobody.
__global__ __launch_bounds__(256) void screen_kernel(const float* __restrict__ y, int64_t n, int K,
                                                     const float* __restrict__ mean,
                                                     const float* __restrict__ sd, float z,
                                                     const float* __restrict__ Wf,
                                                     const double* __restrict__ data, int64_t G,
                                                     const double* __restrict__ thr, float margin,
                                                     double clamp, uint8_t* __restrict__ pass,
                                                     float* __restrict__ score) {
  constexpr int PT = 16;
  __shared__ float e1[PT][ABC_MAX_STATS / 2], c2[PT][ABC_MAX_STATS / 2];
  __shared__ float c3[PT][ABC_MAX_STATS / 2], c4[PT][ABC_MAX_STATS / 2];
  __shared__ int hit[PT], bad[PT];
  __shared__ float s_sum[4][PT], s_max[4][PT];  // per wave: expected pairs, largest margin
  const double gx[7] = {-3.7504397177257425, -2.366759410734541, -1.1544053947399682, 0.0,
                        1.1544053947399682, 2.366759410734541, 3.7504397177257425};
  const double gw[7] = {0.000548268855972217, 0.03075712396758652, 0.2401231786050127,
                        0.45714285714285724, 0.2401231786050127, 0.03075712396758652,
                        0.000548268855972217};
  const int64_t i0 = (int64_t)blockIdx.x * PT;
  if (threadIdx.x < PT) {
    hit[threadIdx.x] = (i0 + threadIdx.x < n) ? 0 : 1;
    bad[threadIdx.x] = 0;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < PT * K; q += blockDim.x) {
    const int p = q / K, k = q % K;
    const int64_t i = i0 + p;
    double m1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0;
    if (i < n) {
      const double mu = y[i * 2 * K + k];
      const double lv = fmin(fmax((double)y[i * 2 * K + K + k], -12.0), 12.0);
      const double s = exp(0.5 * lv);
      double sv[7];
      for (int r = 0; r < 7; ++r) {
        sv[r] = sinh((double)mean[k] + (double)sd[k] * (mu + s * gx[r]));
        m1 += gw[r] * sv[r];
      }
      for (int r = 0; r < 7; ++r) {
        const double e = sv[r] - m1, e2 = e * e;
        v2 += gw[r] * e2;
        v3 += gw[r] * e2 * e;
        v4 += gw[r] * e2 * e2;
      }
      if (!(fabs(m1) < 1e30) || !(v4 < 1e300)) hit[p] = bad[p] = 1;
    }
    e1[p][k] = (float)m1;
    c2[p][k] = (float)v2;
    c3[p][k] = (float)v3;
    c4[p][k] = (float)v4;
  }
  __syncthreads();
  // per particle over this thread's genes: sum of P(err <= t) under the
  // Gaussian approximation, and the largest standardised margin (t - E) / sd
  float psum[PT], pmax[PT];
#pragma unroll
  for (int p = 0; p < PT; ++p) { psum[p] = 0.f; pmax[p] = -HUGE_VALF; }
  for (int64_t g = threadIdx.x; g < G; g += blockDim.x) {
    const double tg = thr[g];
    if (tg >= clamp) {  // every clamped error ties at the threshold: nobody may be skipped
      for (int p = 0; p < PT; ++p) {
        hit[p] = 1;
        psum[p] += 1.f;
        pmax[p] = HUGE_VALF;
      }
      continue;
    }
    const float t = (float)(tg * (1.0 + (double)margin));
    float em[PT], ev[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) { em[p] = 0.f; ev[p] = 0.f; }
    for (int k = 0; k < K; ++k) {
      const float d = (float)data[g + G * k];
      const float w = Wf[g + G * k];
      const float w2 = w * w;
#pragma unroll
      for (int p = 0; p < PT; ++p) {
        const float dl = d - e1[p][k];
        const float a = c2[p][k];
        em[p] += w * (dl * dl + a);
        ev[p] += w2 * fmaxf(4.f * dl * dl * a - 4.f * dl * c3[p][k] + c4[p][k] - a * a, 0.f);
      }
    }
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const float sd = sqrtf(ev[p]);
      if (em[p] - z * sd <= t) hit[p] = 1;  // benign race: every writer stores 1
      if (score) {
        const float d = t - em[p];
        const float u = sd > 0.f ? d / sd : (d >= 0.f ? HUGE_VALF : -HUGE_VALF);
        psum[p] += normcdff(u);
        pmax[p] = fmaxf(pmax[p], u);
      }
    }
  }
  if (score) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      float a = psum[p], b = pmax[p];
      for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b = fmaxf(b, __shfl_xor(b, o));
      }
      if (lane == 0) { s_sum[wave][p] = a; s_max[wave][p] = b; }
    }
  }
  __syncthreads();
  if (threadIdx.x < PT && i0 + threadIdx.x < n) {
    const int p = threadIdx.x;
    pass[i0 + p] = (uint8_t)hit[p];
    if (score) {
      // a prediction that is not finite passes whatever the threshold
      float a = 0.f, b = -HUGE_VALF;
      for (int w = 0; w < 4; ++w) { a += s_sum[w][p]; b = fmaxf(b, s_max[w][p]); }
      score[2 * (i0 + p)] = bad[p] ? HUGE_VALF : a;
      score[2 * (i0 + p) + 1] = bad[p] ? HUGE_VALF : b;
    }
  }
}

struct abc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int n_in = 0, in_pad = 0, width = 0, depth = 0, n_out = 0, out_pad = 0, ns = 0, max_batch = 0;
  int n_cnt = 0;  // leading count statistics (Poisson loss on a log-rate output)
  int tile_rows = 0;  // GEMM workgroup rows: 0 = by shape, else 64 or 128
  int tile_cols = 0;  // GEMM workgroup columns: 0 = by shape, else 64 or 128
  int64_t n_par = 0, n_bf = 0;
  std::vector<LayerDesc> layers;
  std::vector<int64_t> bf_off;
  LayerDesc* d_layers = nullptr;
  int64_t* d_bf_off = nullptr;
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  bf16 *Wb = nullptr, *WbT = nullptr;
  // activations (max_batch rows, padded to 128)
  bf16 *X = nullptr, *XT = nullptr;
  std::vector<float*> Z;
  std::vector<bf16*> H, HT;
  float* O = nullptr;
  bf16 *dA = nullptr, *dAT = nullptr, *dB = nullptr, *dBT = nullptr;
  float* loss = nullptr;
  float* splitk_ws = nullptr;  // split-K partial tiles of the weight gradients
  size_t splitk_cap = 0;       // floats
  int64_t step = 0;
  std::string msg;
};

namespace {

int mfail(abc_mlp* p, int code, const std::string& s) {
  if (p) p->msg = s;
  return code;
}

#define MCHK(p, expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return mfail(p, ABC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Workgroups of a gemm grid; splitting K fills the machine when M x N is small
// (the weight gradients: M, N = layer widths, K = batch).
int splitk_for(int M, int N, int K) {
  // counted in 64-row tiles, so the split does not depend on the tile rows
  // chosen below (results do not depend on them either)
  const int tiles = (M / 64) * (N / (N % 128 == 0 ? 128 : 64));
  int S = 1;
  while (tiles * S < 256 && K % (2 * S * BK) == 0 && K / (2 * S) >= 4 * BK) S *= 2;
  return S;
}

// Workgroup tile rows: 64 when 128-row tiles would leave fewer than two
// workgroups per CU (one wave per SIMD cannot hide the load and LDS latency).
int tile_rows(int M, int N, int S) {
  const int tiles = (M / 128) * (N / (N % 128 == 0 ? 128 : 64)) * S;
  return (tiles < 512 && M % 64 == 0) ? 64 : 128;
}

// Workgroup tile columns (N % 128 == 0): 64 when 128-column tiles would leave
// fewer than four workgroups per CU.
int tile_cols(int M, int N, int S, int bm) {
  return (M / bm) * (N / 128) * S < 1024 ? 64 : 128;
}

template <int EPI>
void gemm(abc_mlp* p, const bf16* A, const bf16* B, int M, int N, int K, const float* bias,
          float* Cf, bf16* Cb, bf16* CbT, const float* Zprev, float* Zout) {
  int S = 1;
  float* out = Cf;
  if (EPI == EPI_F32 && !bias && p->splitk_ws) {
    S = splitk_for(M, N, K);
    if ((size_t)S * M * N > p->splitk_cap) S = 1;
    if (S > 1) out = p->splitk_ws;
  }
  const int bm = p->tile_rows ? p->tile_rows : tile_rows(M, N, S);
  const int bn = N % 128 != 0 ? 64 : (p->tile_cols ? p->tile_cols : tile_cols(M, N, S, bm));
  if (bn == 128 && bm == 128) {
    hipLaunchKernelGGL((gemm_nt_kernel<128, 128, EPI>), dim3(M / 128, N / 128, S), dim3(256), 0,
                       p->stream, A, B, M, N, K / S, K, bias, out, Cb, CbT, Zprev, Zout);
  } else if (bn == 128) {
    hipLaunchKernelGGL((gemm_nt_kernel<64, 128, EPI>), dim3(M / 64, N / 128, S), dim3(256), 0,
                       p->stream, A, B, M, N, K / S, K, bias, out, Cb, CbT, Zprev, Zout);
  } else if (bm == 128) {
    hipLaunchKernelGGL((gemm_nt_kernel<128, 64, EPI>), dim3(M / 128, N / 64, S), dim3(256), 0,
                       p->stream, A, B, M, N, K / S, K, bias, out, Cb, CbT, Zprev, Zout);
  } else {
    hipLaunchKernelGGL((gemm_nt_kernel<64, 64, EPI>), dim3(M / 64, N / 64, S), dim3(256), 0,
                       p->stream, A, B, M, N, K / S, K, bias, out, Cb, CbT, Zprev, Zout);
  }
  if (S > 1)
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(g1((int64_t)M * N / 4, 256)), dim3(256), 0,
                       p->stream, p->splitk_ws, S, (int64_t)M * N, Cf);
}

int forward(abc_mlp* p, int64_t Bn, const float* x_dev) {
  const int Bpad = (int)((Bn + 127) / 128 * 128);
  hipLaunchKernelGGL(cast_input_kernel, dim3(g1((int64_t)Bpad * p->in_pad, 256)), dim3(256), 0,
                     p->stream, x_dev, (int)Bn, Bpad, p->n_in, p->in_pad, p->X, p->XT);
  const bf16* in = p->X;
  for (int l = 0; l < p->depth; ++l) {
    const LayerDesc& L = p->layers[l];
    gemm<EPI_SILU>(p, in, p->Wb + p->bf_off[l], Bpad, L.out, L.in, p->w + L.b_off, nullptr,
                   p->H[l], p->HT[l], nullptr, p->Z[l]);
    in = p->H[l];
  }
  const LayerDesc& L = p->layers[p->depth];
  gemm<EPI_F32>(p, in, p->Wb + p->bf_off[p->depth], Bpad, L.out, L.in, p->w + L.b_off, p->O,
                nullptr, nullptr, nullptr, nullptr);
  MCHK(p, hipGetLastError());
  return ABC_OK;
}

// fwd + loss + bwd into p->g; loss (mean NLL) accumulated in p->loss
int grad(abc_mlp* p, int64_t Bn, const float* x_dev, const float* t_dev) {
  int rc = forward(p, Bn, x_dev);
  if (rc) return rc;
  const int Bpad = (int)((Bn + 127) / 128 * 128);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, p->stream, p->loss, (int64_t)1);
  hipLaunchKernelGGL(nll_kernel, dim3(Bpad / 64, p->out_pad / 64), dim3(256), 0, p->stream,
                     p->O, t_dev, (int)Bn, Bpad, p->ns, p->n_cnt, p->out_pad, p->dA, p->dAT,
                     p->loss);
  // dA / dAT hold dZ of the current layer (output layer first)
  bf16 *cur = p->dA, *curT = p->dAT, *nxt = p->dB, *nxtT = p->dBT;
  for (int l = p->depth; l >= 0; --l) {
    const LayerDesc& L = p->layers[l];
    const bf16* Hin_T = l == 0 ? p->XT : p->HT[l - 1];
    // dW_l = dZ_l^T H_{l-1}: (out x Bpad) . (in x Bpad)^T
    gemm<EPI_F32>(p, curT, Hin_T, L.out, L.in, Bpad, nullptr, p->g + L.w_off, nullptr, nullptr,
                  nullptr, nullptr);
    hipLaunchKernelGGL(rowsum_kernel, dim3(L.out), dim3(256), 0, p->stream, curT, Bpad,
                       p->g + L.b_off);
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}); W_l^T copy is (in x out)
      gemm<EPI_DSILU>(p, cur, p->WbT + p->bf_off[l], Bpad, L.in, L.out, nullptr, nullptr, nxt,
                      nxtT, p->Z[l - 1], nullptr);
      std::swap(cur, nxt);
      std::swap(curT, nxtT);
    }
  }
  MCHK(p, hipGetLastError());
  return ABC_OK;
}

}  // namespace

extern "C" {

int32_t abc_mlp_create(abc_mlp** out, int32_t device, int32_t n_in, int32_t width, int32_t depth,
                       int32_t n_stats, int32_t max_batch, uint64_t seed) {
  if (!out || n_in < 1 || n_in > 256 || width < 64 || width % 128 || depth < 1 || depth > 16 ||
      n_stats < 1 || n_stats > 1024 || max_batch < 1)
    return ABC_E_ARG;
  *out = nullptr;
  abc_mlp* p = new abc_mlp();
  p->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete p; return ABC_E_HIP; }
  p->n_in = n_in;
  p->in_pad = (n_in + 63) / 64 * 64;
  p->width = width;
  p->depth = depth;
  p->ns = n_stats;
  p->n_out = 2 * n_stats;
  p->out_pad = (p->n_out + 127) / 128 * 128;  // rows of dW_out are tiled by 128
  p->max_batch = (max_batch + 127) / 128 * 128;
  int64_t off = 0, boff = 0;
  for (int l = 0; l <= depth; ++l) {
    LayerDesc L;
    L.in = l == 0 ? p->in_pad : width;
    L.out = l == depth ? p->out_pad : width;
    L.w_off = off;
    off += (int64_t)L.in * L.out;
    L.b_off = off;
    off += L.out;
    p->layers.push_back(L);
    p->bf_off.push_back(boff);
    boff += (int64_t)L.in * L.out;
  }
  p->n_par = off;
  p->n_bf = boff;
  // He-uniform initial weights on the real (unpadded) block, zeros elsewhere
  std::vector<float> w0(off, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l <= depth; ++l) {
    const LayerDesc& L = p->layers[l];
    const int rin = l == 0 ? n_in : width, rout = l == depth ? p->n_out : width;
    const float a = std::sqrt(3.0f / (float)rin);
    std::uniform_real_distribution<float> U(-a, a);
    for (int o = 0; o < rout; ++o)
      for (int i = 0; i < rin; ++i) w0[L.w_off + (int64_t)o * L.in + i] = U(rng);
  }
  const int64_t Bp = p->max_batch;
  bool ok = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess;
  auto A = [&](auto** ptr, size_t bytes) {
    if (ok && hipMalloc((void**)ptr, bytes) != hipSuccess) ok = false;
  };
  A(&p->w, 4 * off); A(&p->m, 4 * off); A(&p->v, 4 * off); A(&p->g, 4 * off);
  A(&p->Wb, 2 * boff); A(&p->WbT, 2 * boff);
  A(&p->d_layers, sizeof(LayerDesc) * p->layers.size());
  A(&p->d_bf_off, sizeof(int64_t) * p->bf_off.size());
  A(&p->X, 2 * Bp * p->in_pad); A(&p->XT, 2 * Bp * p->in_pad);
  p->Z.assign(depth, nullptr); p->H.assign(depth, nullptr); p->HT.assign(depth, nullptr);
  for (int l = 0; l < depth; ++l) {
    A(&p->Z[l], 4 * Bp * width); A(&p->H[l], 2 * Bp * width); A(&p->HT[l], 2 * Bp * width);
  }
  const int64_t dmax = std::max<int64_t>(width, p->out_pad);
  A(&p->O, 4 * Bp * p->out_pad);
  A(&p->dA, 2 * Bp * dmax); A(&p->dAT, 2 * Bp * dmax);
  A(&p->dB, 2 * Bp * dmax); A(&p->dBT, 2 * Bp * dmax);
  A(&p->loss, 64);
  // split-K partials: at most 256 tiles of 128 x 128 floats in flight
  p->splitk_cap = (size_t)256 * 128 * 128;
  A(&p->splitk_ws, 4 * p->splitk_cap);
  if (!ok) { abc_mlp_destroy(p); return ABC_E_OOM; }
  std::vector<float> zeros(off, 0.0f);
  if (hipMemcpy(p->w, w0.data(), 4 * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->m, zeros.data(), 4 * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->v, zeros.data(), 4 * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_layers, p->layers.data(), sizeof(LayerDesc) * p->layers.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_bf_off, p->bf_off.data(), sizeof(int64_t) * p->bf_off.size(), hipMemcpyHostToDevice) != hipSuccess) {
    abc_mlp_destroy(p);
    return ABC_E_HIP;
  }
  hipLaunchKernelGGL(refresh_bf16_kernel, dim3(g1(off, 256)), dim3(256), 0, p->stream, p->w,
                     p->d_layers, (int)p->layers.size(), p->d_bf_off, p->Wb, p->WbT, off);
  if (hipStreamSynchronize(p->stream) != hipSuccess) { abc_mlp_destroy(p); return ABC_E_HIP; }
  *out = p;
  return ABC_OK;
}

int32_t abc_mlp_destroy(abc_mlp* p) {
  if (!p) return ABC_OK;
  if (p->stream) hipStreamSynchronize(p->stream);
  void* bufs[] = {p->w, p->m, p->v, p->g, p->Wb, p->WbT, p->d_layers, p->d_bf_off, p->X, p->XT,
                  p->O, p->dA, p->dAT, p->dB, p->dBT, p->loss, p->splitk_ws};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (auto* z : p->Z) if (z) hipFree(z);
  for (auto* h : p->H) if (h) hipFree(h);
  for (auto* h : p->HT) if (h) hipFree(h);
  if (p->stream && p->own_stream) hipStreamDestroy(p->stream);
  delete p;
  return ABC_OK;
}

const char* abc_mlp_last_error(const abc_mlp* p) { return p ? p->msg.c_str() : "null model"; }

int64_t abc_mlp_n_params(const abc_mlp* p) { return p ? p->n_par : ABC_E_ARG; }

int32_t abc_mlp_set_tile_rows(abc_mlp* p, int32_t rows) {
  if (!p || (rows != 0 && rows != 64 && rows != 128)) return ABC_E_ARG;
  p->tile_rows = rows;
  return ABC_OK;
}

int32_t abc_mlp_set_tile_cols(abc_mlp* p, int32_t cols) {
  if (!p || (cols != 0 && cols != 64 && cols != 128)) return ABC_E_ARG;
  p->tile_cols = cols;
  return ABC_OK;
}

int32_t abc_mlp_set_count_outputs(abc_mlp* p, int32_t n_counts) {
  if (!p || n_counts < 0 || n_counts > p->ns) return ABC_E_ARG;
  p->n_cnt = n_counts;
  return ABC_OK;
}

int32_t abc_mlp_set_stream(abc_mlp* p, void* s) {
  if (!p) return ABC_E_ARG;
  hipStreamSynchronize(p->stream);
  if (p->own_stream) hipStreamDestroy(p->stream);
  if (s) {
    p->stream = (hipStream_t)s;
    p->own_stream = false;
  } else {
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return ABC_E_HIP;
    p->own_stream = true;
  }
  return ABC_OK;
}

int32_t abc_mlp_get_params(abc_mlp* p, float* host) {
  if (!p || !host) return ABC_E_ARG;
  MCHK(p, hipStreamSynchronize(p->stream));
  MCHK(p, hipMemcpy(host, p->w, 4 * p->n_par, hipMemcpyDeviceToHost));
  return ABC_OK;
}

int32_t abc_mlp_set_params(abc_mlp* p, const float* host) {
  if (!p || !host) return ABC_E_ARG;
  MCHK(p, hipStreamSynchronize(p->stream));
  MCHK(p, hipMemcpy(p->w, host, 4 * p->n_par, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(refresh_bf16_kernel, dim3(g1(p->n_par, 256)), dim3(256), 0, p->stream, p->w,
                     p->d_layers, (int)p->layers.size(), p->d_bf_off, p->Wb, p->WbT, p->n_par);
  MCHK(p, hipStreamSynchronize(p->stream));
  return ABC_OK;
}

int32_t abc_mlp_forward_dev(abc_mlp* p, int64_t n, const float* x_dev, float* y_dev) {
  if (!p || n < 0 || (n > 0 && (!x_dev || !y_dev))) return ABC_E_ARG;
  for (int64_t o = 0; o < n; o += p->max_batch) {
    const int64_t b = std::min<int64_t>(p->max_batch, n - o);
    int rc = forward(p, b, x_dev + o * p->n_in);
    if (rc) return rc;
    hipLaunchKernelGGL(gather_out_kernel, dim3(g1(b * p->n_out, 256)), dim3(256), 0, p->stream,
                       p->O, (int)b, p->ns, p->out_pad, y_dev + o * p->n_out);
  }
  MCHK(p, hipGetLastError());
  return ABC_OK;
}

int32_t abc_mlp_grad_dev(abc_mlp* p, int64_t n, const float* x_dev, const float* t_dev,
                         float* grad_host, float* loss_host) {
  if (!p || n < 1 || n > p->max_batch || !x_dev || !t_dev) return ABC_E_ARG;
  int rc = grad(p, n, x_dev, t_dev);
  if (rc) return rc;
  MCHK(p, hipStreamSynchronize(p->stream));
  if (grad_host) MCHK(p, hipMemcpy(grad_host, p->g, 4 * p->n_par, hipMemcpyDeviceToHost));
  if (loss_host) MCHK(p, hipMemcpy(loss_host, p->loss, 4, hipMemcpyDeviceToHost));
  return ABC_OK;
}

int32_t abc_mlp_train_step_dev(abc_mlp* p, int64_t n, const float* x_dev, const float* t_dev,
                               float lr, float beta1, float beta2, float eps, float weight_decay,
                               float* loss_dev_or_null) {
  if (!p || n < 1 || n > p->max_batch || !x_dev || !t_dev) return ABC_E_ARG;
  int rc = grad(p, n, x_dev, t_dev);
  if (rc) return rc;
  p->step++;
  const float c1 = 1.0f - std::pow(beta1, (float)p->step);
  const float c2 = 1.0f - std::pow(beta2, (float)p->step);
  hipLaunchKernelGGL(adam_kernel, dim3(g1(p->n_par, 256)), dim3(256), 0, p->stream, p->w, p->m,
                     p->v, p->g, p->n_par, lr, beta1, beta2, eps, weight_decay, c1, c2,
                     p->d_layers, (int)p->layers.size(), p->d_bf_off, p->Wb, p->WbT);
  if (loss_dev_or_null)
    MCHK(p, hipMemcpyAsync(loss_dev_or_null, p->loss, 4, hipMemcpyDeviceToDevice, p->stream));
  MCHK(p, hipGetLastError());
  return ABC_OK;
}

int32_t abc_mlp_synchronize(abc_mlp* p) {
  if (!p) return ABC_E_ARG;
  MCHK(p, hipStreamSynchronize(p->stream));
  return ABC_OK;
}

int32_t abc_mlp_n_stats(const abc_mlp* p) { return p ? p->ns : ABC_E_ARG; }

int32_t abc_features_dev(int32_t model, int64_t n, const double* theta_dev, int32_t n_bins,
                         float* x_dev, void* stream) {
  if (model < 1 || model > 5 || n < 0 || n_bins < 1 || n_bins > ABC_MAX_BINS ||
      (n > 0 && (!theta_dev || !x_dev)))
    return ABC_E_ARG;
  if (n == 0) return ABC_OK;
  hipLaunchKernelGGL(features_kernel, dim3(g1(n, 256)), dim3(256), 0, (hipStream_t)stream, model,
                     n, theta_dev, n_bins, x_dev);
  return hipGetLastError() == hipSuccess ? ABC_OK : ABC_E_HIP;
}

int32_t abc_targets_dev(int64_t n, const double* stats_dev, int32_t K, const float* mean_dev,
                        const float* sd_dev, float* t_dev, void* stream) {
  if (n < 0 || K < 1 || (n > 0 && (!stats_dev || !mean_dev || !sd_dev || !t_dev))) return ABC_E_ARG;
  if (n == 0) return ABC_OK;
  hipLaunchKernelGGL(targets_kernel, dim3(g1(n * K, 256)), dim3(256), 0, (hipStream_t)stream, n,
                     stats_dev, K, mean_dev, sd_dev, t_dev);
  return hipGetLastError() == hipSuccess ? ABC_OK : ABC_E_HIP;
}

}  // extern "C"
