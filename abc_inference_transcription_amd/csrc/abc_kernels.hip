// abc_kernels.hip -- gfx950 kernels of the ABC simulation hot path.
//
//   prior_kernel ...... fix_params (scripts/abc_simulation.jl:3-11) and the
//                       config priors; one thread per particle
//   rates_kernel ...... log10 parameters -> per-particle rates, thinning
//                       bounds, periodic gene law, burn-in length, cost key
//                       (scripts/model.jl:1-43)
//   bucket_* / *scan .. heaviest-first queue order and work-item offsets
//   burn_kernel ....... burn-in of every cell (gene path + Poisson means of the
//                       unlabelled molecules) up to its labelling window; the
//                       stochastic counterpart of scripts/model.jl:114-142
//   ssa_kernel ........ persistent Gillespie simulation of the labelling
//                       window, one (cell, sub-lane) item per lane, lanes refill
//                       from a global item queue the moment their item ends
//                       (lane-level work stealing); replaces the moment ODE
//                       solves of scripts/model.jl:146-206
//   stats_kernel ...... per-particle statistics from observed cell counts with
//                       exact integer accumulation in LDS
//                       (scripts/abc_simulation.jl:23-46 semantics on cells,
//                       estimators of scripts/data_summary_statistics.jl)
//   err_fp64_kernel ... reference-order FP64 error matrix
//                       (scripts/compute_errors.jl:30-70)
//   accept_fp64_kernel  fixed-epsilon candidates from the same FP64 errors
//                       (scripts/accepted_particles.jl:20-22)
#include "abc_kernels.h"

using namespace abcd;

// ------------------------------------------------------------------ layout
struct Layout {
  int n_kon, n_al, n_ga, o_kon, o_koff, o_al, o_ga, o_lam, o_spl, n_par;
};

__host__ __device__ inline Layout layout_of(int model, int nb) {
  Layout L;
  L.n_kon = model == 3 ? nb : 1;
  L.n_al = model == 4 ? nb : 1;
  L.n_ga = model == 5 ? nb : 1;
  L.o_kon = 0;
  L.o_koff = L.n_kon;
  L.o_al = L.o_koff + 1;
  L.o_ga = L.o_al + L.n_al;
  L.o_lam = L.o_ga + L.n_ga;
  L.o_spl = L.o_lam + 1;
  L.n_par = L.o_spl + 1;
  return L;
}

int abc_layout_n_params(int model, int nb) { return layout_of(model, nb).n_par; }

// ------------------------------------------------------------------- prior
struct PriorStream {
  uint32_t k0, k1, tag, pid;  // tag: 0xFF000000 | model, disjoint from every cell word
  int q;
  u4 w;
  ABC_DEV double next() {
    const int call = q >> 1, half = q & 1;
    if (half == 0) w = philox((uint32_t)call, 0u, tag, pid, k0, k1);
    q++;
    return half == 0 ? u01_53(w.x, w.y) : u01_53(w.z, w.w);
  }
  ABC_DEV double unif(double lo, double hi) {
    const double u = next();
    return lo + (hi - lo) * u;
  }
};

#define FC_LO (-0.698970004336018804786)
#define FC_HI (0.698970004336018804786)

__global__ __launch_bounds__(256) void prior_kernel(const DesignDev* __restrict__ D, int model,
                                                    uint32_t k0, uint32_t k1, uint32_t first,
                                                    int64_t n, double* __restrict__ theta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nb = D->nb;
  const Layout L = layout_of(model, nb);
  PriorStream s{k0, k1, 0xFF000000u | (uint32_t)model, first + (uint32_t)i, 0, u4{0, 0, 0, 0}};
  double* th = theta + i;  // column c at th[n * c]
  if (D->prior_mode == 0) {
    for (int c = 0; c < L.n_kon; ++c) th[n * (L.o_kon + c)] = s.unif(-3.0, 3.0);
    th[n * L.o_koff] = s.unif(-3.0, 3.0);
    for (int c = 0; c < L.n_al; ++c) th[n * (L.o_al + c)] = s.unif(-3.0, 3.0);
    for (int c = 0; c < L.n_ga; ++c) th[n * (L.o_ga + c)] = s.unif(-3.0, 2.0);
    th[n * L.o_lam] = s.unif(-0.7, 0.0);
    th[n * L.o_spl] = D->splice_log10;
  } else {
    const double kb = s.unif(-2.0, 2.0);
    if (L.n_kon > 1) {
      for (int c = 0; c < nb; ++c) th[n * (L.o_kon + c)] = kb + s.unif(FC_LO, FC_HI);
    } else {
      th[n * L.o_kon] = kb;
    }
    const double koff = s.unif(-3.0, 3.0);
    th[n * L.o_koff] = koff;
    const double bb = s.unif(-1.0, 3.0);
    if (L.n_al > 1) {
      for (int c = 0; c < nb; ++c) th[n * (L.o_al + c)] = (bb + s.unif(FC_LO, FC_HI)) + koff;
    } else {
      th[n * L.o_al] = bb + koff;
    }
    const double gb = s.unif(-2.0, 1.0);
    if (L.n_ga > 1) {
      for (int c = 0; c < nb; ++c) th[n * (L.o_ga + c)] = gb + s.unif(FC_LO, FC_HI);
    } else {
      th[n * L.o_ga] = gb;
    }
    th[n * L.o_lam] = s.unif(-0.7, 0.0);
    th[n * L.o_spl] = s.unif(-1.0, 2.0);
  }
}

// ------------------------------------------------------------------- rates
__global__ __launch_bounds__(256) void rates_kernel(const DesignDev* __restrict__ D, int model,
                                                    int64_t n, const double* __restrict__ theta,
                                                    PRates* __restrict__ R,
                                                    float* __restrict__ cost) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nb = D->nb;
  const Layout L = layout_of(model, nb);
  const double T = D->T;
  const double h = T / (double)nb;
  PRates r;
  r.scaling = model != 2;
  r.step = model == 2 ? nb : 1;
  for (int j = 0; j < nb; ++j) {
    r.kon[j] = ex10(theta[i + n * (L.o_kon + (L.n_kon > 1 ? j : 0))]);
    r.koff[j] = ex10(theta[i + n * L.o_koff]);
    const double al = ex10(theta[i + n * (L.o_al + (L.n_al > 1 ? j : 0))]);
    r.amax[j] = r.scaling ? (al * (double)(nb + j + 1)) / (double)nb : al;
    r.gam[j] = ex10(theta[i + n * (L.o_ga + (L.n_ga > 1 ? j : 0))]);
  }
  r.bs = ex10(theta[i + n * L.o_spl]);
  for (int j = 0; j < nb; ++j) {
    r.ikon[j] = 1.0 / r.kon[j];
    r.ikoff[j] = 1.0 / r.koff[j];
    r.iku[j] = 1.0 / (r.gam[j] + r.bs);
    r.ikt[j] = 1.0 / r.gam[j];
  }
  const double lam = ex10(theta[i + n * L.o_lam]);
  r.thr_lam = (uint64_t)(lam * 4294967296.0);
  double pi;
  if (L.n_kon > 1) {
    double p = 0.0, E = 1.0;
    for (int j = 0; j < nb; ++j) {
      const double rr = r.kon[j] + r.koff[j];
      const double pj = r.kon[j] / rr;
      const double e = ex(-(rr * h));
      p = pj + (p - pj) * e;
      E = E * e;
    }
    pi = p / (1.0 - E);
  } else {
    pi = r.kon[0] / (r.kon[0] + r.koff[0]);
  }
  if (pi > 1.0) pi = 1.0;
  if (pi < 0.0) pi = 0.0;
  r.thr_pi = (uint64_t)(pi * 4294967296.0);
  double gsum = r.gam[0];
  for (int j = 1; j < nb; ++j) gsum = gsum + r.gam[j];
  const double s = 0.5 * ex(-(gsum * h));
  double S = 1.0;
  for (int q = 0; q < D->bmin; ++q) S = S * s;
  int B = D->bmin;
  while (S > D->burn_eps && B < D->bmax) { S = S * s; B++; }
  r.B = B;
  r.gene_varies = L.n_kon > 1;
  // Mean birth rate with the gene's activity (before capture): the per-cell
  // split below uses it with the operations of the CPU replay, since it
  // selects random streams.  The typical sub-lanes M of an average cell
  // (longest window, mean capture) only orders the work queue.
  double births = 0.0;
  for (int j = 0; j < nb; ++j) births = births + r.amax[j] * (r.kon[j] / (r.kon[j] + r.koff[j]));
  r.brate = births / (double)nb;
  int M = 1;
  if (D->split_target > 0.0) {
    const double est = ((r.brate * D->split_scale) * D->wmax) * 3.0;
    while (M < D->split_max && est > D->split_target * (double)M) M *= 2;
  }
  r.M = M;
  R[i] = r;
  // Expected Philox draws per work item (a scheduling key only, never a
  // result): gene switches over the whole path + molecule events (births
  // times 1 + splicing + loss, a 1/M share) over about half the longest window.
  double sw = 0.0, obs = 0.0, loss = 0.0;
  for (int j = 0; j < nb; ++j) {
    const double pon = r.kon[j] / (r.kon[j] + r.koff[j]);
    sw += 2.0 * r.kon[j] * (1.0 - pon);
    const double b = r.amax[j] * pon * D->capture_mean;
    obs += b;
    loss += b * (r.bs / (r.bs + r.gam[j]) + 1.0);
  }
  cost[i] = (float)((sw / (double)nb) * T * ((double)B + 0.5) +
                    ((obs + loss) / ((double)nb * (double)M)) * (0.5 * D->wmax) + 4.0);
}

// -------------------------------------------------------------- queue order
// Particles enter the work queue heaviest first (longest-processing-time
// order) through a bucket sort of the expected cost on its float exponent and
// top two mantissa bits (quarter-octave buckets).  Order inside a bucket is
// whatever the atomics give: every random stream is keyed by particle, cell
// and sub-lane, never by queue position, so the order changes the schedule
// and nothing else.  These are plain kernels (no library sort, no memset
// node), so a simulate call can be captured in a HIP graph.
__device__ __forceinline__ int cost_bucket(float c) {
  const int k = (int)((__float_as_uint(c) >> 21) & 0x3FFu) - 127 * 4;  // 4*exponent + 2 bits
  return k < 0 ? 0 : (k > ABC_NBUCKET - 1 ? ABC_NBUCKET - 1 : k);
}

__global__ __launch_bounds__(256) void bucket_hist_kernel(const float* __restrict__ cost, int64_t n,
                                                          unsigned* __restrict__ hist) {
  __shared__ unsigned h[ABC_NBUCKET];
  for (int b = threadIdx.x; b < ABC_NBUCKET; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[cost_bucket(cost[i])], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < ABC_NBUCKET; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// cursor[b] = number of particles in heavier buckets (launch: 1 x ABC_NBUCKET)
__global__ __launch_bounds__(ABC_NBUCKET) void bucket_scan_kernel(const unsigned* __restrict__ hist,
                                                                  unsigned* __restrict__ cursor) {
  __shared__ unsigned s[ABC_NBUCKET];
  const int t = threadIdx.x;
  const unsigned x = hist[ABC_NBUCKET - 1 - t];  // heaviest bucket first
  s[t] = x;
  __syncthreads();
  for (int d = 1; d < ABC_NBUCKET; d <<= 1) {
    const unsigned v = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  cursor[ABC_NBUCKET - 1 - t] = s[t] - x;
}

__global__ __launch_bounds__(256) void bucket_scatter_kernel(const float* __restrict__ cost,
                                                             int64_t n, unsigned* __restrict__ cursor,
                                                             int32_t* __restrict__ order) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned pos = atomicAdd(&cursor[cost_bucket(cost[i])], 1u);
  order[pos] = (int32_t)i;
}

// Identity of the particles of one model in a multi-model batch: global
// particle index src[i] (first + i without src) and the model's tag
// (model << 24) for the cell word of its counters.
__global__ __launch_bounds__(256) void fill_ids_kernel(int64_t n, int64_t first,
                                                       const int64_t* __restrict__ src,
                                                       uint32_t mtag, int64_t* __restrict__ pidx,
                                                       uint32_t* __restrict__ mtags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    pidx[i] = src ? src[i] : first + i;
    mtags[i] = mtag;
  }
}

__global__ __launch_bounds__(256) void copy_u64_kernel(const unsigned long long* __restrict__ src,
                                                       unsigned long long* __restrict__ dst,
                                                       int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Sub-lane offsets of every particle's cells after the burn-in: one
// workgroup per particle scans its cells' M = 2^msh into coff[p * (C + 1) + c]
// (exclusive; coff[.. + C] is the total) and pitems[p].
__global__ __launch_bounds__(256) void cell_items_kernel(int64_t n, int C,
                                                         const CellStart* __restrict__ rec,
                                                         int32_t* __restrict__ coff,
                                                         int64_t* __restrict__ pitems) {
  __shared__ int32_t s[256];
  const int t = threadIdx.x;
  const int per = (C + 255) / 256;
  const int c0 = t * per, c1 = min(C, c0 + per);
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
    const CellStart* rp = rec + (size_t)p * (size_t)C;
    int32_t acc = 0;
    for (int c = c0; c < c1; ++c) acc += 1 << rp[c].msh;
    s[t] = acc;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int32_t x = t >= d ? s[t - d] : 0;
      __syncthreads();
      s[t] += x;
      __syncthreads();
  }
  int32_t run = t ? s[t - 1] : 0;
  int32_t* co = coff + (size_t)p * (size_t)(C + 1);
  for (int c = c0; c < c1; ++c) {
    co[c] = run;
    run += 1 << rp[c].msh;
  }
  if (t == 255) {
    co[C] = s[255];
    pitems[p] = s[255];
  }
  __syncthreads();  // s[] is reused by the next particle
  }
}

// Work items of particle rank r = the sub-lanes of all its cells (pitems).
// off[r + 1] = inclusive sum over ranks <= r, computed per block of
// ABC_SCAN_BLOCK ranks here (block totals to bsum), then offset by
// bsum_scan_kernel / add_bsum_kernel.
__global__ __launch_bounds__(256) void items_scan_kernel(const int64_t* __restrict__ pitems,
                                                         const int32_t* __restrict__ order,
                                                         int64_t n, int64_t* __restrict__ off,
                                                         int64_t* __restrict__ bsum) {
  constexpr int PER = ABC_SCAN_BLOCK / 256;
  __shared__ int64_t s[256];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * ABC_SCAN_BLOCK + (int64_t)t * PER;
  int64_t v[PER];
  int64_t acc = 0;
  for (int q = 0; q < PER; ++q) {
    const int64_t r = base + q;
    acc += r < n ? pitems[order[r]] : 0;
    v[q] = acc;
  }
  s[t] = acc;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const int64_t x = t >= d ? s[t - d] : 0;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  const int64_t pre = t ? s[t - 1] : 0;
  for (int q = 0; q < PER; ++q)
    if (base + q < n) off[base + q + 1] = pre + v[q];
  if (t == 255) bsum[blockIdx.x] = s[255];
  if (blockIdx.x == 0 && t == 0) off[0] = 0;
}

// Exclusive scan of nblk block totals in place (launch: 1 x 256).
__global__ __launch_bounds__(256) void bsum_scan_kernel(int64_t* __restrict__ bsum, int64_t nblk) {
  __shared__ int64_t s[256];
  const int t = threadIdx.x;
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += 256) {
    const int64_t b = b0 + t;
    const int64_t x = b < nblk ? bsum[b] : 0;
    s[t] = x;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int64_t y = t >= d ? s[t - d] : 0;
      __syncthreads();
      s[t] += y;
      __syncthreads();
    }
    if (b < nblk) bsum[b] = carry + s[t] - x;
    carry += s[255];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void add_bsum_kernel(int64_t* __restrict__ off, int64_t n,
                                                       const int64_t* __restrict__ bsum) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  off[r + 1] += bsum[r / ABC_SCAN_BLOCK];
}

// ------------------------------------------------------- shared helpers
// Random streams of a cell: counter (draw_lo, draw_hi[15:0] | stream << 16,
// cell | model << 24, particle), key (seed_lo, seed_hi); stream 0 holds the
// set-up draws and the gene clock, stream 1 + sub the molecule events and
// division coins of sub-lane sub.
// time of grid edge k (edges every h = T / nb hours; edge nbB is t = 0)
__device__ __forceinline__ double grid_time(int64_t k, int64_t nbB, double h) {
  return (double)(k - nbB) * h;
}

__device__ __forceinline__ u4 draw(uint64_t ctr, uint32_t stream, uint32_t cell, uint32_t pid,
                                   uint32_t k0, uint32_t k1) {
  return philox((uint32_t)ctr, ((uint32_t)(ctr >> 32) & 0xFFFFu) | (stream << 16), cell, pid, k0, k1);
}

// ----- burn-in: before the labelling window molecules are never labelled and
// never act on the gene or on each other, so given the gene path the observed
// unspliced and spliced molecules alive at the window start are independent
// Poisson counts (births a Poisson process of rate alpha(t) beta / M while the
// gene is on; decay, splicing and halving at division thin every molecule
// independently; Jahnke & Huisinga, J. Math. Biol. 54, 2007).  The burn-in
// simulates the gene path only and carries the two Poisson means lu
// (unspliced: leave at gamma + splice) and lt (all: leave at gamma).

// E1 = int_0^D e^{-k(D-s)} ds, E2 = int_0^D s e^{-k(D-s)} ds, e = e^{-kD}, ik = 1/k
__device__ __forceinline__ void seg_int(double k, double ik, double D, double e, double& E1,
                                        double& E2) {
  const double x = k * D;
  if (x < 1e-3) {
    E1 = D * (1.0 - x * (0.5 - x * (1.0 / 6.0 - x * (1.0 / 24.0))));
    E2 = (D * D) * (0.5 - x * (1.0 / 6.0 - x * (1.0 / 24.0 - x * (1.0 / 120.0))));
  } else {
    E1 = (1.0 - e) * ik;
    E2 = (D - E1) * ik;
  }
}

// advance the Poisson means over [t, t + D]: birth rate c0 + c1 s while on;
// ku, kt are the exit rates (with reciprocals iku, ikt)
__device__ __forceinline__ void burn_accum(double& lu, double& lt, bool g, double c0, double c1,
                                           double ku, double iku, double kt, double ikt, double D) {
  if (!(D > 0.0)) return;
  // arguments <= 0 (D > 0, rates > 0): no NaN or overflow test
  const double eu = ex_core(-(ku * D)), et = ex_core(-(kt * D));
  lu = lu * eu;
  lt = lt * et;
  if (g) {
    double E1u, E2u, E1t, E2t;
    seg_int(ku, iku, D, eu, E1u, E2u);
    seg_int(kt, ikt, D, et, E1t, E2t);
    lu = lu + (c0 * E1u + c1 * E2u);
    lt = lt + (c0 * E1t + c1 * E2t);
  }
}

// consecutive words of a sub-lane's molecule stream
struct WordSrc {
  u4 w;
  int i;  // next word; 4 = empty
  uint64_t& ctr;  // Philox blocks drawn so far on this stream
  uint32_t stream, cell, pid, k0, k1;
  __device__ uint32_t next() {
    if (i == 4) {
      w = draw(ctr++, stream, cell, pid, k0, k1);
      i = 0;
    }
    const uint32_t v = i == 0 ? w.x : (i == 1 ? w.y : (i == 2 ? w.z : w.w));
    i++;
    return v;
  }
};

// ln(k!) for integer-valued k >= 0: table below 10, Stirling series above
__device__ __forceinline__ double logfact(double kd, const LogEntry* lt) {
  if (kd < 10.0) return abc_logfact[(int)kd];
  const double x = kd + 1.0, ix = 1.0 / x, ix2 = ix * ix;
  return ((x - 0.5) * ln(x, lt) - x) +
         (0x1.d67f1c864beb5p-1 +
          ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0 - ix2 * (1.0 / 1680.0)))));
}

// Poisson(mu): inversion below 10, else the transformed rejection method with
// squeeze (Hormann, Insurance: Math. Econ. 12, 1993), at most 1000 tries.
__device__ uint32_t poisson(WordSrc& src, double mu, const LogEntry* lt, const double* rk) {
  if (!(mu > 0.0)) return 0u;
  if (mu < 10.0) {
    // inversion: p_k = p_{k-1} * mu * RN(1/k) (rk: the table in LDS)
    const uint32_t w0 = src.next(), w1 = src.next();
    const double u = u01_53(w0, w1);
    double p = ex(-mu), F = p;
    uint32_t k = 0;
    while (u > F && k < 200u) {
      k++;
      p = (p * mu) * rk[k];
      F = F + p;
    }
    return k;
  }
  const double smu = ex(0.5 * ln(mu, lt));
  const double bb = 0.931 + 2.53 * smu, aa = -0.059 + 0.02483 * bb;
  const double vr = 0.9277 - 3.6224 / (bb - 2.0);
  for (int tries = 0; tries < 1000; ++tries) {
    const double U = u01_32(src.next()) - 0.5;
    const double V = u01_32(src.next());
    const double us = 0.5 - fabs(U);
    const double kd = floor(((2.0 * aa) / us + bb) * U + mu + 0.43);
    if (kd < 0.0) continue;
    if (us >= 0.07 && V <= vr) return (uint32_t)kd;
    if (us < 0.013 && V > us) continue;
    // rare: the envelope constant and ln(mu) are recomputed here (same bits)
    const double inva = 1.1239 + 1.1328 / (bb - 3.4);
    if ((ln(V, lt) + ln(inva, lt)) - ln(aa / (us * us) + bb, lt) <=
        (kd * ln(mu, lt) - mu) - logfact(kd, lt))
      return (uint32_t)kd;
  }
  return (uint32_t)floor(mu + 0.5);
}

// Set-up of cell c of particle p on its stream 0 (counter 0 ...): age and bin,
// the gene's state from its periodic stationary law, the capture efficiency.
struct CellSetup {
  double age, bc;
  int bin;
  bool g;
};

__device__ __forceinline__ CellSetup cell_setup(const DesignDev* __restrict__ D, const PRates* rp,
                                                int c, uint32_t cw, uint32_t pid, uint32_t k0,
                                                uint32_t k1,
                                                const double* __restrict__ betas,
                                                const int64_t* __restrict__ goff, uint64_t& gctr) {
  const double T = D->T;
  const int nb = D->nb, J = D->J;
  CellSetup o;
  const u4 w = draw(gctr++, 0u, cw, pid, k0, k1);
  if (D->age_mode == 1) {
    const double th = u01_53(w.x, w.y);
    o.age = th * T;
    o.bin = (int)(th * (double)nb);
    if (o.bin >= nb) o.bin = nb - 1;
  } else {
    o.bin = (c / J) % nb;
    o.age = ((double)(2 * o.bin + 1) / (double)(2 * nb)) * T;
  }
  o.g = (uint64_t)w.z < rp->thr_pi;
  const int cond = c % J;
  if (D->cap_mode == 2) {
    const int grp = o.bin * 2 + D->is_chase[cond];
    const int64_t off0 = goff[grp];
    const uint32_t cnt = (uint32_t)(goff[grp + 1] - off0);
    o.bc = betas[off0 + (int64_t)lemire(w.w, cnt)];
  } else if (D->cap_mode == 1) {
    // Beta(a, b) = a-th smallest of a+b-1 uniforms
    const int a = D->beta_a, nu = D->beta_a + D->beta_b - 1;
    uint32_t s0 = 0xFFFFFFFFu, s1 = 0xFFFFFFFFu, s2 = 0xFFFFFFFFu, s3 = 0xFFFFFFFFu;
    u4 v{0, 0, 0, 0};
    for (int qq = 0; qq < nu; ++qq) {
      if ((qq & 3) == 0) v = draw(gctr++, 0u, cw, pid, k0, k1);
      const int sel = qq & 3;
      const uint32_t x = sel == 0 ? v.x : (sel == 1 ? v.y : (sel == 2 ? v.z : v.w));
      if (x < s3) {
        s3 = x;
        if (s3 < s2) { const uint32_t tt = s2; s2 = s3; s3 = tt; }
        if (s2 < s1) { const uint32_t tt = s1; s1 = s2; s2 = tt; }
        if (s1 < s0) { const uint32_t tt = s0; s0 = s1; s1 = tt; }
      }
    }
    o.bc = u01_32(a == 1 ? s0 : (a == 2 ? s1 : (a == 3 ? s2 : s3)));
  } else {
    o.bc = 1.0;
  }
  return o;
}

// Burn-in, one lane per cell (persistent, cells of the heaviest particles
// first): the gene path from the start of the burn-in to the labelling window
// start ws, carrying the Poisson means of the cell's observed unspliced (lu)
// and all (lt) molecules; the cell's state at ws goes to rec[p * C + c].
__global__ __launch_bounds__(SSA_WG) void burn_kernel(const DesignDev* __restrict__ D,
                                                      const PRates* __restrict__ R,
                                                      const int32_t* __restrict__ order, int64_t n,
                                                      uint32_t first, uint32_t k0, uint32_t k1,
                                                      uint32_t mtag_all,
                                                      const uint32_t* __restrict__ mtags,
                                                      const double* __restrict__ betas,
                                                      const int64_t* __restrict__ goff,
                                                      unsigned long long* __restrict__ queue,
                                                      CellStart* __restrict__ rec,
                                                      unsigned long long* __restrict__ steps_out,
                                                      unsigned long long* __restrict__ react_out,
                                                      const int64_t* __restrict__ pidx,
                                                      unsigned long long* __restrict__ burn_tally) {
  const double T = D->T;
  const int nb = D->nb, J = D->J, C = D->C;
  const double h = T / (double)nb;  // hours between rate-bin edges
  const unsigned long long total = (unsigned long long)n * (unsigned long long)C;
  unsigned long long lane_steps = 0;  // this lane's burn-in draws, added to *burn_tally at exit
  __shared__ LogEntry32 ltab32[ABC_LOG_N];
  load_log32_table(ltab32);
  __syncthreads();
  for (;;) {
    const unsigned long long q = atomicAdd(queue, 1ull);
    if (q >= total) break;
    const int p = order[q / (unsigned long long)C];
    const int c = (int)(q % (unsigned long long)C);
    const PRates* rp = R + p;
    const uint32_t pid = pidx ? (uint32_t)pidx[p] : first + (uint32_t)p;
    // counter word: the cell and the particle's model (the key is the seed alone)
    const uint32_t cw = (uint32_t)c | (mtags ? mtags[p] : mtag_all);
    uint64_t gctr = 0;
    const CellSetup cs0 = cell_setup(D, rp, c, cw, pid, k0, k1, betas, goff, gctr);
    const int cond = c % J;
    const double age = cs0.age, bc = cs0.bc;
    const double ws = age - D->pulse[cond] - D->chase[cond];
    const int step = rp->step;
    const bool scaling = rp->scaling, gvar = rp->gene_varies;
    const double bs = rp->bs;
    const int64_t nbB = (int64_t)rp->B * nb;
    int64_t k = 0, knext = step;
    int j = 0;
    double t = grid_time(0, nbB, h), cs = t, gk = grid_time(knext, nbB, h), tg = 0.0;
    double lu = 0.0, lt = 0.0;
    bool g = cs0.g, gneed = true;
    double ikon = rp->ikon[0], ikoff = rp->ikoff[0];
    double atx = rp->amax[0] * bc, gam = rp->gam[0], bound = D->bound[0];
    double iku = rp->iku[0], ikt = rp->ikt[0], abnd = atx / bound;
    unsigned long long st = 0, re = 0;
    for (;;) {
      double tb = age;
      if (gk < tb) tb = gk;
      if (ws < tb) tb = ws;
      if (gneed) {
        const u4 v = draw(gctr++, 0u, cw, pid, k0, k1);
        st++;
        tg = t + (double)exp1_32(v.x, ltab32) * (g ? ikoff : ikon);
        gneed = false;
      }
      // births at rate atx (T + a) / bound_j (size scaling) or atx, a = t - cycle start
      {
        const double t1 = tg < tb ? tg : tb;
        double c0 = atx, c1 = 0.0;
        if (scaling) {
          c1 = abnd;
          c0 = c1 * (T + (t - cs));
        }
        burn_accum(lu, lt, g, c0, c1, gam + bs, iku, gam, ikt, t1 - t);
      }
      if (tg < tb) {
        t = tg;
        g = !g;
        re++;
        gneed = true;
        continue;
      }
      t = tb;
      if (tb == gk) {
        k = knext;
        knext += step;
        j += step;  // k mod nb (step is 1 or nb)
        if (j >= nb) j -= nb;
        if (j == 0) {
          cs = tb;
          lu = lu * 0.5;  // division halves the Poisson means
          lt = lt * 0.5;
        }
        if (step == 1) {
          ikon = rp->ikon[j];
          ikoff = rp->ikoff[j];
          atx = rp->amax[j] * bc;
          gam = rp->gam[j];
          bound = D->bound[j];
          iku = rp->iku[j];
          ikt = rp->ikt[j];
          abnd = atx / bound;
        }
        if (gvar) gneed = true;  // the gene's rate changed: restart its clock
        gk = grid_time(knext, nbB, h);
      }
      if (tb == ws) break;
    }
    CellStart r;
    r.age = age;
    r.bc = bc;
    r.tg = tg;
    r.lu = lu;
    r.lt = lt;
    r.k = k;
    r.gctr = gctr;
    r.bin = (uint8_t)cs0.bin;
    r.g = (uint8_t)g;
    r.gneed = (uint8_t)gneed;
    // sub-lanes of this cell: smallest power of two M with the cell's expected
    // molecule events in its window <= split_target * M (M selects random
    // streams: the CPU replay uses the same operations).  Events: the
    // unlabelled molecules at the window start (at most 2 each) and births
    // at the cell's own capture over its own window (3 events each).
    {
      int M = 1, msh = 0;
      if (D->split_target > 0.0) {
        const double w = D->pulse[cond] + D->chase[cond];
        const double est = (lu + lt) + ((rp->brate * bc) * w) * 3.0;
        while (M < D->split_max && est > D->split_target * (double)M) {
          M *= 2;
          msh++;
        }
      }
      r.msh = (uint8_t)msh;
    }
    rec[(size_t)p * (size_t)C + (size_t)c] = r;
    atomicAdd(steps_out + p, st);
    atomicAdd(react_out + p, re);
    lane_steps += st;
  }
  atomicAdd(burn_tally, lane_steps);
}

// ---------------------------------------------------------------- the SSA
// Labelling window of a work item = (particle, cell, sub-lane).  Sub-lane sub
// draws its unlabelled molecules at the window start, Poisson(lu / M)
// unspliced and Poisson((lt - lu) / M) spliced, from its own stream 1 + sub,
// then simulates the four species to the sampling age: it replays the cell's
// gene path from the shared stream 0 (continued from the burn-in) and carries
// its own 1/M share of the births and those molecules' fates (direct method).
// Molecules never act on the gene, so the sum over sub-lanes has the law of
// the whole cell.  Lanes refill from a global item queue the moment their item
// ends, so heavy and light particles share the machine without idle lanes.
// Held to 128 VGPRs, four waves per SIMD: the compiler's own allocation (140
// VGPRs, three waves) leaves more of the FP64 and Philox latency exposed.
// The four-wave build measured 7 % less SSA time per step; with the Philox
// key the same in every lane it fits 128 VGPRs without spilling.
__global__ __launch_bounds__(SSA_WG) __attribute__((amdgpu_waves_per_eu(4))) void ssa_kernel(const DesignDev* __restrict__ D,
                                                     const PRates* __restrict__ R,
                                                     const int32_t* __restrict__ order,
                                                     const int64_t* __restrict__ off, int64_t n,
                                                     uint32_t first, uint32_t k0, uint32_t k1,
                                                     uint32_t mtag_all,
                                                     const uint32_t* __restrict__ mtags,
                                                     const CellStart* __restrict__ rec,
                                                     const int32_t* __restrict__ coff,
                                                     unsigned long long* __restrict__ queue,
                                                     uint32_t* __restrict__ out_u,
                                                     uint32_t* __restrict__ out_l,
                                                     uint8_t* __restrict__ out_bin,
                                                     unsigned long long* __restrict__ steps_out,
                                                     const int64_t* __restrict__ pidx,
                                                     uint32_t claim) {
  const double T = D->T;
  const int nb = D->nb, J = D->J, C = D->C;
  const double h = T / (double)nb;  // hours between rate-bin edges
  const unsigned long long total = (unsigned long long)off[n];
  __shared__ LogEntry ltab[ABC_LOG_N];
  __shared__ LogEntry32 ltab32[ABC_LOG_N];
  __shared__ double rk[ABC_RECIP_N];
  load_log_table(ltab);
  load_log32_table(ltab32);
  for (int q = threadIdx.x; q < ABC_RECIP_N; q += blockDim.x) rk[q] = abc_recip_k[q];
  __syncthreads();

  uint32_t cw = 0;  // counter word of the current item: cell | model << 24
  // per-lane item state
  bool have = false;
  int p = 0, c = 0, j = 0, step = 1, scaling = 0, gvar = 0;
  uint32_t pid = 0, sstream = 1, n0 = 0, n1 = 0, n2 = 0, n3 = 0;
  // Philox blocks drawn on the gene stream (continued from the burn-in, which
  // had drawn gctr0) and on the item's molecule stream: every step draws one
  // block on exactly one of them, so the item's steps are mctr + gctr - gctr0
  uint64_t gctr = 0, gctr0 = 0, mctr = 0;
  // labelled when pw <= thr_m1 = floor(lambda 2^32) - 1 (lab_on: lambda > 0)
  uint32_t thr_m1 = 0, shi = 0;  // shi: the molecule stream's counter word, stream << 16
  bool lab_on = false;
  int64_t k = 0, knext = 0, nbB = 0;
  // tcs = T - (start of the current cycle): the size-scaling test compares
  // with t + tcs = T + (time since the cycle start)
  double t = 0.0, tb = 0.0, gk = 0.0, tg = 0.0, age = 0.0, we = 0.0, tcs = 0.0;
  double ikon = 0.0, ikoff = 0.0, atx = 0.0, gam = 0.0, bs = 0.0, bc = 1.0, bound = 0.0;
  double ishare = 1.0;  // 1 / (the cell's sub-lanes), a power of two
  bool g = false, in_win = false, div = false;
  int div_s = 0;
  uint32_t div_rem = 0, div_keep = 0;
  // gene clock to (re)draw; pending molecule candidate (its residual Exp(1)
  // hazard and unused words)
  bool gneed = false, pend = false;
  double erem = 0.0;
  uint32_t py = 0, pz = 0, pw = 0;
  // next item of the lane's claim: claim (a power of two) consecutive items
  // per queue atomic, a launch shape only (items are independent).  The
  // queue advances by whole claims, so a claim is used up exactly when qnext
  // reaches a multiple of it.
  unsigned long long qnext = 0;
  // where the lane's last item lies, so the next item of the same claim (the
  // next queue position) is found by one step instead of two searches: queue
  // rank of its particle, item within the particle, the particle's items,
  // the end of the cell's sub-lanes (item offsets within the particle) and
  // the sub-lane
  int32_t wr = 0, wrem = 0, witems = 0, wcend = 0, wsub = 0;

  for (;;) {
    if (!have) {
      const bool fresh = (qnext & (unsigned long long)(claim - 1u)) == 0ull;
      if (fresh) qnext = atomicAdd(queue, (unsigned long long)claim);
      const unsigned long long q = qnext++;
      if (q >= total) break;
      if (fresh) {
        // particle rank: largest r with off[r] <= q
        int64_t lo = 0, hi = n;
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if ((unsigned long long)off[mid] <= q) lo = mid; else hi = mid;
        }
        wr = (int32_t)lo;
        p = order[lo];
        wrem = (int32_t)(q - (unsigned long long)off[lo]);
        witems = (int32_t)(off[lo + 1] - off[lo]);
        // cell: largest c with coff[c] <= wrem among the particle's C cells
        const int32_t* co = coff + (size_t)p * (size_t)(C + 1);
        int clo = 0, chi = C;
        while (chi - clo > 1) {
          const int mid = (clo + chi) >> 1;
          if (co[mid] <= wrem) clo = mid; else chi = mid;
        }
        c = clo;
        wsub = wrem - co[c];
        wcend = co[c + 1];
      } else if (++wrem < witems) {
        // the next item of the same particle: the next sub-lane, or the next
        // cell's first (every cell has at least one sub-lane)
        if (wrem < wcend) {
          wsub++;
        } else {
          c++;
          wsub = 0;
          wcend = coff[(size_t)p * (size_t)(C + 1) + (size_t)(c + 1)];
        }
      } else {
        // the first item of the next particle in queue order
        wr++;
        p = order[wr];
        wrem = 0;
        witems = (int32_t)(off[wr + 1] - off[wr]);
        c = 0;
        wsub = 0;
        wcend = coff[(size_t)p * (size_t)(C + 1) + 1];
      }
      const PRates* rp = R + p;
      const int sub = wsub;
      pid = pidx ? (uint32_t)pidx[p] : first + (uint32_t)p;
      cw = (uint32_t)c | (mtags ? mtags[p] : mtag_all);
      sstream = 1u + (uint32_t)sub;
      shi = sstream << 16;
      const CellStart* rc = rec + ((size_t)p * (size_t)C + (size_t)c);
      ishare = ldexp(1.0, -(int)rc->msh);  // x / M == x * ishare exactly
      mctr = 0;
      // every loop-carried value is reset here, so none stays live across the
      // sampler below
      div_s = 0;
      div_rem = 0;
      div_keep = 0;
      erem = 0.0;
      py = pz = pw = 0;
      {
        // the sub-lane's unlabelled molecules at the window start
        WordSrc src{u4{0, 0, 0, 0}, 4, mctr, sstream, cw, pid, k0, k1};
        const double lu0 = rc->lu, ls = rc->lt - lu0;
        n0 = poisson(src, lu0 * ishare, ltab, rk);
        n1 = poisson(src, ls > 0.0 ? ls * ishare : 0.0, ltab, rk);
      }
      const CellStart r0 = *rc;
      n2 = n3 = 0;
      const int cond = c % J;
      age = r0.age;
      bc = r0.bc;
      we = age - D->chase[cond];
      if (sub == 0) out_bin[(size_t)p * (size_t)C + (size_t)c] = r0.bin;
      step = rp->step;
      scaling = rp->scaling;
      gvar = rp->gene_varies;
      lab_on = rp->thr_lam != 0ull;
      thr_m1 = (uint32_t)(rp->thr_lam - 1ull);
      bs = rp->bs;
      nbB = (int64_t)rp->B * nb;
      k = r0.k;
      knext = k + step;
      j = (int)((uint64_t)k % (uint64_t)nb);
      tcs = T - grid_time(k - j, nbB, h);
      gk = grid_time(knext, nbB, h);
      t = age - D->pulse[cond] - D->chase[cond];
      gctr = r0.gctr;
      gctr0 = gctr;
      g = r0.g != 0;
      tg = r0.tg;
      gneed = r0.gneed != 0;
      ikon = rp->ikon[j];  // gene clock: waiting time = E * (1 / rate)
      ikoff = rp->ikoff[j];
      atx = (rp->amax[j] * bc) * ishare;
      gam = rp->gam[j];
      bound = D->bound[j];
      in_win = true;
      div = false;
      pend = false;
      tb = age;
      if (gk < tb) tb = gk;
      if (we < tb) tb = we;
      have = true;
    }

    // propensities of this sub-lane's molecule channels (direct method)
    const double a_tx = g ? atx : 0.0;
    const uint32_t nun = n0 + n2, ntot = n0 + n1 + n2 + n3;
    const double a_spl = bs * (double)nun;
    const double a_dec = gam * (double)ntot;
    const double c1 = a_tx, c2 = c1 + a_spl;
    const double a0 = c2 + a_dec;

    // One Philox block per lane and iteration, in uniform control flow:
    // division coins, the gene clock (shared stream 0) or a new molecule
    // candidate (stream 1 + sub).  A candidate pre-empted by the gene clock or
    // a boundary stays pending: its words are unused and its Exp(1) hazard,
    // less what the interval used up, is again Exp(1) (memoryless).
    // (division coins, else the gene clock, else a new candidate, else no draw)
    const bool kgen = !div && gneed;
    const bool kmol = !div && !gneed && !pend && a0 > 0.0;
    const bool mol = div || kmol;
    const uint64_t ctr = mol ? mctr : gctr;
    const u4 w = philox((uint32_t)ctr, ((uint32_t)(ctr >> 32) & 0xFFFFu) | (mol ? shi : 0u), cw,
                        pid, k0, k1);
    mctr += mol ? 1ull : 0ull;
    gctr += kgen ? 1ull : 0ull;
    if (div) {
      // binomial(count, 1/2) partition of species div_s, 128 coin flips per block
      const uint32_t m = div_rem < 128u ? div_rem : 128u;
      uint32_t cnt = __popc(m >= 32u ? w.x : (w.x & ((1u << m) - 1u)));
      if (m > 32u) cnt += __popc(m >= 64u ? w.y : (w.y & ((1u << (m - 32u)) - 1u)));
      if (m > 64u) cnt += __popc(m >= 96u ? w.z : (w.z & ((1u << (m - 64u)) - 1u)));
      if (m > 96u) cnt += __popc(m >= 128u ? w.w : (w.w & ((1u << (m - 96u)) - 1u)));
      div_keep += cnt;
      div_rem -= m;
      if (div_rem == 0u) {
        if (div_s == 0) n0 = div_keep;
        else if (div_s == 1) n1 = div_keep;
        else if (div_s == 2) n2 = div_keep;
        else n3 = div_keep;
        div_s++;
        while (div_s < 4 && (div_s == 1 ? n1 : (div_s == 2 ? n2 : n3)) == 0u) div_s++;
        if (div_s < 4) {
          div_rem = div_s == 1 ? n1 : (div_s == 2 ? n2 : n3);
          div_keep = 0;
        } else {
          div = false;
        }
      }
      continue;
    }
    const double E = (double)exp1_32(w.x, ltab32);
    if (kgen) {
      tg = t + E * (g ? ikoff : ikon);
      gneed = false;
      if (!pend && a0 > 0.0) continue;  // a molecule candidate comes first
    } else if (kmol) {
      erem = E;
      py = w.y;
      pz = w.z;
      pw = w.w;
      pend = true;
    }

    // next event: the molecule candidate, the gene switch or a boundary
    const double tm = (pend && a0 > 0.0) ? t + erem / a0 : __longlong_as_double(0x7ff0000000000000ll);
    if (tm < tg && tm < tb) {
      t = tm;
      pend = false;
      // The channel, then its molecule, as predicated arithmetic: lanes of a
      // wave that fire different channels run one instruction stream.
      //   transcription (rr < c1): kept by the size-scaling thinning test on
      //     pz, labelled when pw < thr_lam inside the window;
      //   splicing (c1 <= rr < c2): molecule lemire(pz, n0 + n2) of the
      //     unspliced, unlabelled first;
      //   decay (rr >= c2): molecule lemire(pz, ntot) in species order.
      // u01_32(py) * a0 with the power-of-two scale on a0: the same single
      // rounding of the same exact product, one instruction less
      const double rr = ((double)py + 0.5) * (a0 * 0x1p-32);
      const bool tx = rr < c1;
      const bool spl = !tx && rr < c2;
      const bool dec = !tx && !spl;
      const bool acc = tx && (!scaling || u01_32(pz) * bound < t + tcs);
      const bool lab = in_win && lab_on && pw <= thr_m1;
      const uint32_t x = lemire(pz, spl ? nun : ntot);
      const bool ge0 = x >= n0, ge1 = x >= n0 + n1, ge2 = x >= n0 + n1 + n2;
      // each species gains one, loses one or stays (the conditions exclude
      // each other): one select chain and one add per species
      n0 += (acc && !lab) ? 1u : ((!tx && !ge0) ? 0xFFFFFFFFu : 0u);
      n1 += (spl && !ge0) ? 1u : ((dec && ge0 && !ge1) ? 0xFFFFFFFFu : 0u);
      n2 += (acc && lab) ? 1u : (((spl && ge0) || (dec && ge1 && !ge2)) ? 0xFFFFFFFFu : 0u);
      n3 += (spl && ge0) ? 1u : ((dec && ge2) ? 0xFFFFFFFFu : 0u);
      continue;
    }
    if (pend) {
      // no molecule event before min(tg, tb): a0 times that interval is used up
      const double t1 = tg < tb ? tg : tb;
      erem = erem - a0 * (t1 - t);
      if (erem < 0.0) erem = 0.0;
    }
    if (tg < tb) {
      t = tg;
      g = !g;
      gneed = true;
      continue;
    }
    // boundary: sampling age, rate-bin / cycle edge, or the window end
    t = tb;
    if (tb == age) {
      const size_t o = (size_t)p * (size_t)C + (size_t)c;
      atomicAdd(out_u + o, n0 + n1);
      atomicAdd(out_l + o, n2 + n3);
      atomicAdd(steps_out + p, (unsigned long long)(mctr + (gctr - gctr0)));
      have = false;
      continue;
    }
    if (tb == gk) {
      k = knext;
      knext += step;
      j += step;  // k mod nb (step is 1 or nb)
      if (j >= nb) j -= nb;
      if (j == 0) {
        tcs = T - tb;
        div_s = 0;
        while (div_s < 4 && (div_s == 0 ? n0 : (div_s == 1 ? n1 : (div_s == 2 ? n2 : n3))) == 0u) div_s++;
        if (div_s < 4) {
          div = true;
          div_rem = div_s == 0 ? n0 : (div_s == 1 ? n1 : (div_s == 2 ? n2 : n3));
          div_keep = 0;
        }
      }
      if (step == 1) {
        const PRates* rp = R + p;
        ikon = rp->ikon[j];
        ikoff = rp->ikoff[j];
        atx = (rp->amax[j] * bc) * ishare;
        gam = rp->gam[j];
        bound = D->bound[j];
      }
      if (gvar) gneed = true;  // the gene's rate changed: restart its clock (memoryless)
      gk = grid_time(knext, nbB, h);
    }
    if (in_win && tb == we) in_win = false;
    tb = age;
    if (gk < tb) tb = gk;
    if (in_win && we < tb) tb = we;
  }
}

// -------------------------------------------------------------- statistics
__device__ __forceinline__ void add128(unsigned long long* lo, unsigned long long* hi,
                                       unsigned long long x) {
  const unsigned long long old = atomicAdd(lo, x);
  if (old + x < old) atomicAdd(hi, 1ull);
}

struct U128 { uint64_t hi, lo; };

__device__ __forceinline__ U128 mul64(uint64_t a, uint64_t b) {
  return U128{__umul64hi(a, b), a * b};
}
__device__ __forceinline__ U128 add(U128 a, U128 b) {
  U128 r{a.hi + b.hi, a.lo + b.lo};
  if (r.lo < a.lo) r.hi++;
  return r;
}
__device__ __forceinline__ U128 sub(U128 a, U128 b) {  // requires a >= b
  U128 r{a.hi - b.hi, a.lo - b.lo};
  if (a.lo < b.lo) r.hi--;
  return r;
}
__device__ __forceinline__ bool ge(U128 a, U128 b) {
  return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo);
}
// n * (hi:lo) with n < 2^32 and the product < 2^128
__device__ __forceinline__ U128 mul_small(uint64_t n, U128 v) {
  U128 lo = mul64(n, v.lo);
  lo.hi += n * v.hi;
  return lo;
}

__device__ __forceinline__ double var_of(uint64_t n, uint64_t S, U128 SS) {
  const U128 num = sub(mul_small(n, SS), mul64(S, S));
  return u128_to_f64(num.hi, num.lo) / ((double)n * (double)(n - 1));
}
__device__ __forceinline__ double cov_of(uint64_t n, uint64_t Sa, uint64_t Sb, U128 Sab) {
  const U128 a = mul_small(n, Sab), b = mul64(Sa, Sb);
  const double den = (double)n * (double)(n - 1);
  if (ge(a, b)) { const U128 d = sub(a, b); return u128_to_f64(d.hi, d.lo) / den; }
  const U128 d = sub(b, a);
  return -(u128_to_f64(d.hi, d.lo) / den);
}

__device__ __forceinline__ double wcov(const double* x, const double* y, const double* w, int nb) {
  double mx = w[0] * x[0], my = w[0] * y[0];
  for (int b = 1; b < nb; ++b) mx = mx + w[b] * x[b];
  for (int b = 1; b < nb; ++b) my = my + w[b] * y[b];
  double acc = w[0] * ((x[0] - mx) * (y[0] - my));
  for (int b = 1; b < nb; ++b) acc = acc + w[b] * ((x[b] - mx) * (y[b] - my));
  return acc;
}

__global__ __launch_bounds__(256) void stats_kernel(const DesignDev* __restrict__ D,
                                                    const uint32_t* __restrict__ cu,
                                                    const uint32_t* __restrict__ cl,
                                                    const uint8_t* __restrict__ cb, int64_t n,
                                                    double* __restrict__ stats_ref,
                                                    double* __restrict__ moments,
                                                    double* __restrict__ superset) {
  const int nb = D->nb, J = D->J, C = D->C, NBJ = nb * J;
  __shared__ unsigned int s_n[ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ unsigned long long s_su[ABC_MAX_BINS * ABC_MAX_CONDS], s_sl[ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ unsigned long long s_uu[2][ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ unsigned long long s_ll[2][ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ unsigned long long s_ul[2][ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ double s_m1[ABC_MAX_BINS * ABC_MAX_CONDS], s_m2[ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ double s_v1[ABC_MAX_BINS * ABC_MAX_CONDS], s_v2[ABC_MAX_BINS * ABC_MAX_CONDS];
  __shared__ double s_c12[ABC_MAX_BINS * ABC_MAX_CONDS];
  // one particle per workgroup at a time (grid-stride over particles)
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    for (int q = threadIdx.x; q < NBJ; q += blockDim.x) {
      s_n[q] = 0; s_su[q] = 0; s_sl[q] = 0;
      s_uu[0][q] = s_uu[1][q] = 0; s_ll[0][q] = s_ll[1][q] = 0; s_ul[0][q] = s_ul[1][q] = 0;
    }
    __syncthreads();
    const uint32_t* u = cu + (size_t)i * C;
    const uint32_t* l = cl + (size_t)i * C;
    const uint8_t* bn = cb + (size_t)i * C;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
      if (bn[c] >= nb) continue;  // never produced by the SSA; keeps LDS indexing in range
      const int idx = bn[c] + nb * (c % J);
      const unsigned long long a = u[c], b = l[c];
      atomicAdd(&s_n[idx], 1u);
      atomicAdd(&s_su[idx], a);
      atomicAdd(&s_sl[idx], b);
      add128(&s_uu[0][idx], &s_uu[1][idx], a * a);
      add128(&s_ll[0][idx], &s_ll[1][idx], b * b);
      add128(&s_ul[0][idx], &s_ul[1][idx], a * b);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < NBJ; idx += blockDim.x) {
      const uint64_t nn = s_n[idx];
      const U128 uu{s_uu[1][idx], s_uu[0][idx]}, ll{s_ll[1][idx], s_ll[0][idx]},
          ul{s_ul[1][idx], s_ul[0][idx]};
      double mt = 0.0, fft = 0.0, m1 = 0.0, m2 = 0.0, v1 = 0.0, v2 = 0.0, c12 = 0.0;
      if (nn > 0) {
        const double q = (double)nn;
        m1 = (double)s_su[idx] / q;
        m2 = (double)s_sl[idx] / q;
        v1 = var_of(nn, s_su[idx], uu);
        v2 = var_of(nn, s_sl[idx], ll);
        c12 = cov_of(nn, s_su[idx], s_sl[idx], ul);
        const uint64_t St = s_su[idx] + s_sl[idx];
        const U128 Stt = add(add(uu, add(ul, ul)), ll);
        mt = (double)St / q;
        const double vt = var_of(nn, St, Stt);
        fft = vt / (mt + (mt == 0.0 ? 0.0001 : 0.0));
      }
      s_m1[idx] = m1; s_m2[idx] = m2; s_v1[idx] = v1; s_v2[idx] = v2; s_c12[idx] = c12;
      if (moments) {
        double* mo = moments + i;
        const int64_t base = 6 * (int64_t)idx;
        mo[n * (base + 0)] = (double)nn;
        mo[n * (base + 1)] = m1;
        mo[n * (base + 2)] = m2;
        mo[n * (base + 3)] = v1;
        mo[n * (base + 4)] = c12;
        mo[n * (base + 5)] = v2;
        if (nn == 0) {
          for (int qq = 0; qq < 6; ++qq) mo[n * (base + qq)] = 0.0;
        }
      }
      if (superset) {
        superset[i + n * idx] = mt;
        superset[i + n * (NBJ + idx)] = fft;
      }
    }
    __syncthreads();
    const int tid = threadIdx.x;
    if (stats_ref && tid < 2 * nb) {
      const int grp = tid / nb, b = tid % nb;
      uint64_t nn = 0, St = 0;
      U128 Stt{0, 0};
      for (int jj = 0; jj < J; ++jj) {
        if (D->is_chase[jj] != grp) continue;
        const int idx = b + nb * jj;
        nn += s_n[idx];
        St += s_su[idx] + s_sl[idx];
        const U128 uu{s_uu[1][idx], s_uu[0][idx]}, ll{s_ll[1][idx], s_ll[0][idx]},
            ul{s_ul[1][idx], s_ul[0][idx]};
        Stt = add(Stt, add(add(uu, add(ul, ul)), ll));
      }
      double mean = 0.0, ff = 0.0;
      if (nn > 0) {
        mean = (double)St / (double)nn;
        const double v = var_of(nn, St, Stt);
        ff = v / (mean + (mean == 0.0 ? 0.0001 : 0.0));
      }
      stats_ref[i + n * ((2 * grp) * nb + b)] = mean;
      stats_ref[i + n * ((2 * grp + 1) * nb + b)] = ff;
    } else if (stats_ref && tid >= 64 && tid < 64 + J) {
      const int jj = tid - 64;
      uint64_t nn = 0, su = 0, sl = 0;
      bool anynz = false;
      for (int b = 0; b < nb; ++b) {
        const int idx = b + nb * jj;
        nn += s_n[idx]; su += s_su[idx]; sl += s_sl[idx];
        if (s_c12[idx] != 0.0) anynz = true;
      }
      double ratio = 0.0;
      if (nn > 0) {
        const double a1 = (double)su / (double)nn, a2 = (double)sl / (double)nn;
        if (a1 + a2 > 0.0) ratio = a2 / (a1 + a2);
      }
      const double* w = D->weights + nb * jj;
      const double* m1 = s_m1 + nb * jj;
      const double* m2 = s_m2 + nb * jj;
      const double* v1 = s_v1 + nb * jj;
      const double* v2 = s_v2 + nb * jj;
      const double* c12 = s_c12 + nb * jj;
      double s1 = w[0] * v1[0], s2 = w[0] * v2[0], s12 = w[0] * c12[0];
      for (int b = 1; b < nb; ++b) s1 = s1 + w[b] * v1[b];
      for (int b = 1; b < nb; ++b) s2 = s2 + w[b] * v2[b];
      for (int b = 1; b < nb; ++b) s12 = s12 + w[b] * c12[b];
      const double tv1 = s1 + wcov(m1, m1, w, nb);
      const double tv2 = s2 + wcov(m2, m2, w, nb);
      double mc = 0.0, cm = 0.0;
      if (anynz && tv1 != 0.0 && tv2 != 0.0) mc = s12 / sqrt(fabs(tv1 * tv2));
      if (tv1 != 0.0 && tv2 != 0.0) cm = wcov(m1, m2, w, nb) / sqrt(fabs(tv1 * tv2));
      stats_ref[i + n * (4 * nb + jj)] = ratio;
      stats_ref[i + n * (4 * nb + J + jj)] = mc;
      stats_ref[i + n * (4 * nb + 2 * J + jj)] = cm;
    }
    __syncthreads();  // the LDS sums are cleared for the next particle
  }
}

// ------------------------------------------------------------------ errors
// One FP64 error in the order of compute_errors.jl:30-64.
__device__ __forceinline__ double err_pair(const double* __restrict__ data,
                                           const double* __restrict__ se, int64_t G, int64_t g,
                                           const double* __restrict__ stats, int64_t n,
                                           int64_t i, const ErrSpec& E) {
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
  return err;
}

// err holds the columns of genes g0 .. g0 + gridDim.y - 1
__global__ __launch_bounds__(256) void err_fp64_kernel(const double* __restrict__ stats,
                                                       int64_t n, const double* __restrict__ data,
                                                       const double* __restrict__ se, int64_t G,
                                                       int64_t g0, ErrSpec E,
                                                       double* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = g0 + blockIdx.y;
  if (i >= n) return;
  err[i + n * (int64_t)blockIdx.y] = err_pair(data, se, G, g, stats, n, i, E);
}

__global__ __launch_bounds__(256) void accept_fp64_kernel(const double* __restrict__ stats,
                                                          int64_t n,
                                                          const double* __restrict__ data,
                                                          const double* __restrict__ se,
                                                          int64_t G, int64_t g0, ErrSpec E,
                                                          double eps, int64_t first,
                                                          Candidate* __restrict__ cand,
                                                          unsigned long long* __restrict__ count,
                                                          unsigned long long cap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = g0 + blockIdx.y;
  if (i >= n) return;
  const double e = err_pair(data, se, G, g, stats, n, i, E);
  if (e <= eps) {
    const unsigned long long slot = atomicAdd(count, 1ull);
    if (slot < cap) cand[slot] = Candidate{e, first + i, (int32_t)g, 0};
  }
}

// The device's elementary functions on n inputs, for validation: kind 0 =
// ex (every kernel except the loops below), 1 = ex_core (the burn-in and the
// moment model's loops; arguments <= 709.78), 2 = ln (LDS table)
__global__ __launch_bounds__(256) void math_probe_kernel(int kind, int64_t n,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y) {
  __shared__ LogEntry ltab[ABC_LOG_N];
  load_log_table(ltab);
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    y[i] = kind == 0 ? ex(v) : (kind == 1 ? ex_core(v) : ln(v, ltab));
  }
}

__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
