// abc_gemm.hip -- particle x gene error screening on MFMA (split-BF16) with
// exact FP64 rescoring, the acceptance path of scripts/compute_errors.jl:30-70
// and scripts/accepted_particles.jl:12-32.
//
// err(i,g) = sum_k W_gk (d_gk - s_ik)^2 = c_g + sum_k s_ik (-2 W_gk d_gk) + sum_k s_ik^2 W_gk
// with W_gk = 1 / (K (se^2 + 0.01 d^2 + eps)) and c_g = sum_k W_gk d_gk^2.
// The two sums are one GEMM of A = [S | S^2] (n x 2K) with B = [-2WD ; W]
// (2K x G).  Every operand x is split x = hi + lo (two bf16) and the product
// uses hi*hi + hi*lo + lo*hi with FP32 accumulation on the matrix cores.
//
// Error bound (per element): |a*b - (ah*bh + ah*bl + al*bh)| <= 3.02*2^-16 |a||b|
// and FP32 accumulation of 3*Kp products adds <= 3*Kp*2^-24 sum|terms|; with
// Kp <= 128 both stay below DELTA = 2^-12 of sum_k |a_k||b_k|.  By
// Cauchy-Schwarz over W >= 0:  sum_k |a_k||b_k| <= 2 sqrt(y c) + y, where
// y = sum_k W s^2 is the second accumulator.  A pair is a candidate whenever
// its lower bound can reach the gene's threshold; candidates are rescored in
// FP64 in the reference order, so the accepted sets equal the FP64 ones.
#include <hip/hip_runtime.h>

#include "abc_kernels.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define SCREEN_DELTA 2.44140625e-4  // 2^-12
#define SCREEN_LIMIT 1e15           // operands beyond this go straight to FP64

__device__ __forceinline__ void split_bf16(double x, bf16* hi, bf16* lo) {
  const bf16 h = (bf16)(float)x;
  const double r = x - (double)(float)h;
  *hi = h;
  *lo = (bf16)(float)r;
}

// One thread per gene: B rows (transposed: Bt[g][k], 2*Kp per gene), c_g and
// a flag for genes whose weights leave the screened range.
__global__ __launch_bounds__(256) void gene_prep_kernel(const double* __restrict__ data,
                                                        const double* __restrict__ se, int64_t G,
                                                        int K, int Kp, int64_t Gpad,
                                                        bf16* __restrict__ Bh,
                                                        bf16* __restrict__ Bl,
                                                        double* __restrict__ cg,
                                                        uint8_t* __restrict__ gflag,
                                                        float* __restrict__ Wf) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= Gpad) return;
  bf16* bh = Bh + g * 2 * Kp;
  bf16* bl = Bl + g * 2 * Kp;
  const bf16 z = (bf16)0.0f;
  if (g >= G) {
    for (int k = 0; k < 2 * Kp; ++k) { bh[k] = z; bl[k] = z; }
    return;
  }
  const double s2 = 0.1 * 0.1;
  double c = 0.0;
  bool bad = false;
  for (int k = 0; k < Kp; ++k) {
    if (k < K) {
      const double d = data[g + G * k];
      const double sg = se[g + G * k];
      const double eps = (sg + d != 0.0) ? 0.0 : 0.0001;
      const double den = (sg * sg + s2 * (d * d)) + eps;
      const double W = 1.0 / ((double)K * den);
      const double WD = W * d;
      c += WD * d;
      if (!(fabs(W) <= SCREEN_LIMIT) || !(fabs(WD) <= SCREEN_LIMIT)) bad = true;
      Wf[g + G * k] = (float)W;
      split_bf16(-2.0 * WD, bh + k, bl + k);
      split_bf16(W, bh + Kp + k, bl + Kp + k);
    } else {
      bh[k] = z; bl[k] = z; bh[Kp + k] = z; bl[Kp + k] = z;
    }
  }
  if (!(c <= SCREEN_LIMIT)) bad = true;
  cg[g] = c;
  gflag[g] = bad ? 1 : 0;
}

// One thread per particle: A rows [s | s^2] split into bf16 hi/lo.  Row i of
// the n rows is stats[i + ld * k] (a block of rows of a larger batch).
__global__ __launch_bounds__(256) void particle_prep_kernel(const double* __restrict__ stats,
                                                            int64_t n, int64_t ld, int K, int Kp,
                                                            int64_t npad,
                                                            bf16* __restrict__ Ah,
                                                            bf16* __restrict__ Al,
                                                            uint8_t* __restrict__ pflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  bf16* ah = Ah + i * 2 * Kp;
  bf16* al = Al + i * 2 * Kp;
  const bf16 z = (bf16)0.0f;
  bool bad = false;
  for (int k = 0; k < Kp; ++k) {
    if (i < n && k < K) {
      const double s = stats[i + ld * k];
      if (!(fabs(s) <= SCREEN_LIMIT)) bad = true;
      const double sc = bad ? 0.0 : s;
      split_bf16(sc, ah + k, al + k);
      split_bf16(sc * sc, ah + Kp + k, al + Kp + k);
    } else {
      ah[k] = z; al[k] = z; ah[Kp + k] = z; al[Kp + k] = z;
    }
  }
  if (i < n) pflag[i] = bad ? 1 : 0;
}

// 256 threads = 4 waves; the workgroup covers 64 particles x 64 genes, each
// wave a 32 x 32 block as 2 x 2 MFMA 16x16x32 tiles.  Lane layout of
// v_mfma_f32_16x16x32_bf16: A lane l holds row l%16, k = (l/16)*8 .. +7;
// B lane l holds column l%16, the same k; D lane l holds column l%16,
// rows (l/16)*4 .. +3.
__global__ __launch_bounds__(256) void gemm_screen_kernel(
    const bf16* __restrict__ Ah, const bf16* __restrict__ Al, int64_t n, int Kp,
    const bf16* __restrict__ Bh, const bf16* __restrict__ Bl, int64_t G,
    const double* __restrict__ cg, const uint8_t* __restrict__ gflag,
    const uint8_t* __restrict__ pflag, const double* __restrict__ thr, double margin_rel,
    double clamp, Candidate* __restrict__ out, unsigned long long* __restrict__ count,
    unsigned long long cap, float* __restrict__ approx_out, float* __restrict__ bound_out,
    PoolEntry* __restrict__ pool, int64_t capg, uint32_t* __restrict__ pool_n) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t i0 = (int64_t)blockIdx.x * 64 + (wave >> 1) * 32;
  const int64_t g0 = (int64_t)blockIdx.y * 64 + (wave & 1) * 32;
  const int row = lane & 15, kq = (lane >> 4) * 8;
  const int KK = 2 * Kp;
  f32x4 accC[2][2], accS[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accC[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      accS[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  const int nchunk = Kp / 32;
  // cross half (k < Kp) into accC, square half (k >= Kp) into accS
  auto chunk = [&](f32x4 (&acc)[2][2], int k) {
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const int64_t r = i0 + rb * 16 + row;
      ah[rb] = *reinterpret_cast<const bf16x8*>(Ah + r * KK + k);
      al[rb] = *reinterpret_cast<const bf16x8*>(Al + r * KK + k);
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int64_t cc = g0 + cb * 16 + row;
      bh[cb] = *reinterpret_cast<const bf16x8*>(Bh + cc * KK + k);
      bl[cb] = *reinterpret_cast<const bf16x8*>(Bl + cc * KK + k);
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
      }
  };
  for (int ch = 0; ch < nchunk; ++ch) chunk(accC, ch * 32 + kq);
  for (int ch = 0; ch < nchunk; ++ch) chunk(accS, Kp + ch * 32 + kq);
  // epilogue: bounds, thresholds, candidate emission.  A pair is a candidate
  // unless its lower bound provably exceeds the gene's threshold; a
  // non-finite approximation or bound (FP32 overflow) makes it one.
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int64_t g = g0 + cb * 16 + (lane & 15);
    const bool gok = g < G;
    const double c = gok ? cg[g] : 0.0;
    const double t = gok ? thr[g] : 0.0;
    const bool gbad = gok && gflag[g] != 0;
    unsigned cmask = 0;  // bit rb * 4 + q: row i0 + rb * 16 + (lane >> 4) * 4 + q
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t i = i0 + rb * 16 + (lane >> 4) * 4 + q;
        if (!gok || i >= n) continue;
        const double cross = (double)accC[rb][cb][q];
        const double sq = (double)accS[rb][cb][q];
        const double approx = (c + cross) + sq;
        const double y = fmax(sq, 0.0) * (1.0 + 2.0 * SCREEN_DELTA) + 1e-30;
        const double bound = SCREEN_DELTA * (2.0 * sqrt(y * fmax(c, 0.0)) + y) +
                             1e-12 * fabs(approx) + 1e-30;
        if (approx_out) {
          approx_out[i + n * g] = (float)approx;
          bound_out[i + n * g] = (float)bound;
        }
        double lo = approx - bound;
        if (t < 1e300) lo -= margin_rel * fabs(t);
        if (lo > clamp) lo = clamp;
        const bool must = pflag[i] != 0 || gbad;
        if (must || !(lo > t)) cmask |= 1u << (rb * 4 + q);
      }
    if (pool) {
      // lanes l, l + 16, l + 32, l + 48 hold the same gene: one atomic per
      // gene and wave reserves their pool slots
      const int col = lane & 15, rgrp = lane >> 4;
      const unsigned cnt = (unsigned)__popc(cmask);
      const unsigned c0 = __shfl(cnt, col), c1 = __shfl(cnt, col + 16),
                     c2 = __shfl(cnt, col + 32), c3 = __shfl(cnt, col + 48);
      const unsigned tot = c0 + c1 + c2 + c3;
      const unsigned pre = (rgrp > 0 ? c0 : 0u) + (rgrp > 1 ? c1 : 0u) + (rgrp > 2 ? c2 : 0u);
      unsigned base = 0;
      if (rgrp == 0 && tot > 0) base = atomicAdd(pool_n + g, tot);
      base = __shfl(base, col);
      int64_t slot = (int64_t)base + pre;
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        if (!(cmask & (1u << bit))) continue;
        const int64_t i = i0 + (bit >> 2) * 16 + (lane >> 4) * 4 + (bit & 3);
        if (slot < capg) pool[g * capg + slot] = PoolEntry{0.0, i};
        slot++;
      }
    } else if (out) {
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        if (!(cmask & (1u << bit))) continue;
        const int64_t i = i0 + (bit >> 2) * 16 + (lane >> 4) * 4 + (bit & 3);
        const double cross = (double)accC[bit >> 2][cb][bit & 3];
        const double sq = (double)accS[bit >> 2][cb][bit & 3];
        const unsigned long long s = atomicAdd(count, 1ull);
        if (s < cap) out[s] = Candidate{(c + cross) + sq, i, (int32_t)g, 0};
      }
    }
  }
}

// FP64 rescoring of candidates in the reference operation order
// (compute_errors.jl:30-64).  Candidate.particle holds the batch-local row.
__global__ __launch_bounds__(256) void rescore_kernel(const double* __restrict__ stats, int64_t n,
                                                      const double* __restrict__ data,
                                                      const double* __restrict__ se, int64_t G,
                                                      ErrSpec E, Candidate* __restrict__ cand,
                                                      const unsigned long long* __restrict__ count,
                                                      unsigned long long cap) {
  const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long m = *count;
  if (m > cap) m = cap;
  if (q >= m) return;
  Candidate c = cand[q];
  const int64_t i = c.particle, g = c.gene;
  const double s2 = 0.1 * 0.1;
  double err = 0.0;
  int k = 0;
  for (int pp = 0; pp < E.n_parts; ++pp) {
    double e = 0.0;
    const int end = k + E.parts[pp];
    for (; k < end; ++k) {
      const double d = data[g + G * k];
      const double sg = se[g + G * k];
      const double s = stats[i + n * k];
      const double eps = (sg + d != 0.0) ? 0.0 : 0.0001;
      const double diff = d - s;
      e += (diff * diff) / ((sg * sg + s2 * (d * d)) + eps);
    }
    err += e / (double)E.K;
  }
  if (err > E.clamp) err = E.clamp;
  c.err = err;
  cand[q] = c;
}
