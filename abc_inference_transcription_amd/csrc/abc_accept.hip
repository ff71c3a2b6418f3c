// abc_accept.hip -- per-gene top-k acceptance kept in device memory: the
// selection of scripts/accepted_particles.jl:12-32 (per gene, particles
// ordered by error with ties by index, a stable sortperm) over errors of
// scripts/compute_errors.jl:30-70, for the top-k rule with k accepted
// particles per gene.
//
// Every gene owns a pool of capg = k + n_sub entries (error, particle index)
// in HBM.  Between batches a pool holds at most k entries: the gene's best k
// so far, and its threshold thr[g] is the k-th error.  One batch of at most
// n_sub particles then proceeds without a host round trip:
//   gemm_screen_kernel (abc_gemm.hip) appends each (particle, gene) pair whose
//     proven lower bound can reach thr[g] to the gene's pool (at most n_sub
//     new entries per gene, so a pool never overflows);
//   pool_rescore_kernel replaces each new entry by its FP64 error in the
//     reference operation order and its global particle index;
//   pool_select_kernel keeps each gene's k smallest (error, index) keys
//     (radix select over the 128-bit key, then a stable in-place compaction)
//     and lowers thr[g] to the k-th error.
// pool_sort_kernel orders each gene's final entries by (error, index).
#include <hip/hip_runtime.h>

#include "abc_kernels.h"

namespace {

constexpr int kWG = 256;

// Order-preserving map of an FP64 value onto uint64 (x <= y iff key(x) <=
// key(y) for non-NaN values; errors are >= +0, but the map is total).
__device__ __forceinline__ uint64_t err_key(double e) {
  const uint64_t b = (uint64_t)__double_as_longlong(e);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_err(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ bool key_le(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
  return ah < bh || (ah == bh && al <= bl);
}

// Exclusive prefix of a per-thread flag over the 256 threads of the
// workgroup, in thread order; *total gets the sum.  Uses scratch[4].
__device__ __forceinline__ unsigned block_scan_flag(bool f, unsigned* scratch, unsigned* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(f);
  const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
  if (lane == 0) scratch[wave] = (unsigned)__popcll(m);
  __syncthreads();
  unsigned pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) {
    const unsigned s = scratch[w];
    if (w < wave) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + below;
}

}  // namespace

// Workgroups (blockIdx.x over genes, grid-stride; blockIdx.y of gridDim.y
// over each gene's entries): the FP64 error of every entry appended since the
// last selection, in the order of compute_errors.jl:30-64 (no contraction),
// and its global particle index.  One lane per entry.  The batch's statistics
// are read in their column-major layout (row i, stat k at scol[i + ld k]):
// a gene's new entries come from the screen in nearly increasing row order,
// so the lanes' loads of one statistic are mostly contiguous.  An entry's
// index field holds its batch row on entry.
__global__ __launch_bounds__(256) void pool_rescore_kernel(
    const double* __restrict__ scol, int64_t ld, const double* __restrict__ data,
    const double* __restrict__ se, int64_t G, ErrSpec E, PoolEntry* __restrict__ pool,
    int64_t capg, const uint32_t* __restrict__ pool_n, const uint32_t* __restrict__ pool_base,
    int64_t first_index, const int64_t* __restrict__ pidx) {
  __shared__ double s_d[ABC_MAX_STATS], s_den[ABC_MAX_STATS];
  const int K = E.K;
  const double s2 = 0.1 * 0.1;
  const int64_t stride = (int64_t)kWG * gridDim.y;
  for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
    const int64_t cnt = min((int64_t)pool_n[g], capg);
    const int64_t b = pool_base[g] + (int64_t)blockIdx.y * kWG;
    if (cnt <= b) continue;  // uniform over the workgroup
    for (int k = threadIdx.x; k < K; k += kWG) {
      const double d = data[g + G * k];
      const double sg = se[g + G * k];
      const double eps = (sg + d != 0.0) ? 0.0 : 0.0001;
      s_d[k] = d;
      s_den[k] = (sg * sg + s2 * (d * d)) + eps;
    }
    __syncthreads();
    PoolEntry* pg = pool + g * capg;
    for (int64_t e = b + threadIdx.x; e < cnt; e += stride) {
      const int64_t i = pg[e].idx;
      const double* s = scol + i;
      double err = 0.0;
      int k = 0;
      for (int pp = 0; pp < E.n_parts; ++pp) {
        double acc = 0.0;
        const int end = k + E.parts[pp];
        for (; k < end; ++k) {
          const double diff = s_d[k] - s[ld * k];
          acc += (diff * diff) / s_den[k];
        }
        err += acc / (double)K;
      }
      if (err > E.clamp) err = E.clamp;
      pg[e] = PoolEntry{err, pidx ? pidx[i] : first_index + i};
    }
    __syncthreads();
  }
}

// One workgroup per gene (grid-stride): keep the k smallest (error, index)
// keys among the pool's entries with a non-NaN error (NaN errors are never
// accepted), compact them to the front of the pool in their current order,
// and lower thr[g] to the k-th error once the pool holds k entries.  Genes
// without new entries are skipped.  n_rescored accumulates the FP64
// rescorings (entries appended since the last selection).
__global__ __launch_bounds__(256) void pool_select_kernel(
    PoolEntry* __restrict__ pool, int64_t capg, uint32_t* __restrict__ pool_n,
    uint32_t* __restrict__ pool_base, int64_t G, int64_t k, double* __restrict__ thr,
    unsigned long long* __restrict__ n_rescored) {
  __shared__ unsigned hist[256];
  __shared__ unsigned scratch[4];
  __shared__ unsigned s_cnt[4];
  __shared__ double s_mx[4];
  __shared__ uint64_t s_ph, s_pl, s_mh, s_ml;  // key prefix found so far and its mask
  __shared__ unsigned s_rank, s_bincnt, s_running;
  __shared__ uint64_t s_kh, s_kl;              // the k-th key
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
    const int64_t cnt = min((int64_t)pool_n[g], capg);
    const int64_t b = pool_base[g];
    if (cnt == b) continue;  // uniform
    PoolEntry* pg = pool + g * capg;
    if (tid == 0) {
      atomicAdd(n_rescored, (unsigned long long)(cnt - b));
      s_running = 0;
    }
    // valid entries and their largest error
    unsigned nv = 0;
    double mx = -HUGE_VAL;
    for (int64_t e = tid; e < cnt; e += kWG) {
      const double x = pg[e].err;
      if (x == x) { nv++; mx = fmax(mx, x); }
    }
    for (int o = 32; o > 0; o >>= 1) {
      nv += __shfl_xor(nv, o);
      mx = fmax(mx, __shfl_xor(mx, o));
    }
    if (lane == 0) { s_cnt[wave] = nv; s_mx[wave] = mx; }
    __syncthreads();
    const unsigned tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    const double vmax = fmax(fmax(s_mx[0], s_mx[1]), fmax(s_mx[2], s_mx[3]));
    const bool keep_all = (int64_t)tot <= k;
    if (!keep_all && k > 0) {
      // radix select of the key of rank k-1 among the valid entries, 8 bits a
      // pass from the top of the 128-bit key (error bits, then index)
      if (tid == 0) { s_ph = s_pl = s_mh = s_ml = 0; s_rank = (unsigned)(k - 1); }
      for (int pass = 0; pass < 16; ++pass) {
        hist[tid] = 0;
        __syncthreads();
        const uint64_t ph = s_ph, pl = s_pl, mh = s_mh, ml = s_ml;
        const bool hi = pass < 8;
        const int sh = 56 - 8 * (pass & 7);
        for (int64_t e = tid; e < cnt; e += kWG) {
          const PoolEntry x = pg[e];
          if (!(x.err == x.err)) continue;
          const uint64_t kh = err_key(x.err), kl = (uint64_t)x.idx;
          if ((kh & mh) != ph || (kl & ml) != pl) continue;
          atomicAdd(&hist[(unsigned)(((hi ? kh : kl) >> sh) & 0xFF)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
          const unsigned r = s_rank;
          unsigned c = 0;
          int d = 0;
          for (; d < 255; ++d) {
            if (c + hist[d] > r) break;
            c += hist[d];
          }
          s_rank = r - c;
          s_bincnt = hist[d];
          if (hi) { s_ph |= (uint64_t)d << sh; s_mh |= 0xFFull << sh; }
          else { s_pl |= (uint64_t)d << sh; s_ml |= 0xFFull << sh; }
        }
        __syncthreads();
        if (s_bincnt == 1) break;  // uniform: the prefix now names one entry
      }
      // the full k-th key: the entry matching the prefix (after all 16 passes
      // every match carries the same key)
      const uint64_t ph = s_ph, pl = s_pl, mh = s_mh, ml = s_ml;
      for (int64_t e = tid; e < cnt; e += kWG) {
        const PoolEntry x = pg[e];
        if (!(x.err == x.err)) continue;
        const uint64_t kh = err_key(x.err), kl = (uint64_t)x.idx;
        if ((kh & mh) == ph && (kl & ml) == pl) { s_kh = kh; s_kl = kl; }
      }
      __syncthreads();
    }
    const uint64_t KH = keep_all ? ~0ull : s_kh, KL = keep_all ? ~0ull : s_kl;
    const bool none = !keep_all && k <= 0;
    // stable in-place compaction of the kept entries: the positions a chunk
    // writes never exceed those it reads, and all its reads precede its writes
    for (int64_t base = 0; base < cnt; base += kWG) {
      const int64_t e = base + tid;
      PoolEntry x{};
      bool f = false;
      if (e < cnt && !none) {
        x = pg[e];
        f = (x.err == x.err) && key_le(err_key(x.err), (uint64_t)x.idx, KH, KL);
      }
      const unsigned run = s_running;
      unsigned total;
      const unsigned pos = block_scan_flag(f, scratch, &total);
      if (f) pg[run + pos] = x;
      if (tid == 0) s_running = run + total;
      __syncthreads();
    }
    if (tid == 0) {
      const unsigned kept = s_running;
      pool_n[g] = kept;
      pool_base[g] = kept;
      if (k > 0 && (int64_t)kept >= k) {
        const double kth = keep_all ? vmax : key_err(KH);
        if (kth < thr[g]) thr[g] = kth;
      } else if (k <= 0) {
        thr[g] = -HUGE_VAL;
      }
    }
    __syncthreads();
  }
}

// One workgroup per gene (grid-stride): the pool's entries (at most
// ABC_POOL_SORT_MAX) sorted in place by (error, index), a bitonic sort in LDS.
__global__ __launch_bounds__(256) void pool_sort_kernel(PoolEntry* __restrict__ pool,
                                                        int64_t capg,
                                                        const uint32_t* __restrict__ pool_n,
                                                        int64_t G) {
  __shared__ uint64_t kh[ABC_POOL_SORT_MAX], kl[ABC_POOL_SORT_MAX];
  const int tid = threadIdx.x;
  for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
    const int cnt = (int)min((int64_t)pool_n[g], (int64_t)ABC_POOL_SORT_MAX);
    if (cnt <= 1) continue;  // uniform
    int N = 2;
    while (N < cnt) N <<= 1;
    PoolEntry* pg = pool + g * capg;
    for (int e = tid; e < N; e += kWG) {
      if (e < cnt) {
        kh[e] = err_key(pg[e].err);
        kl[e] = (uint64_t)pg[e].idx;
      } else {
        kh[e] = ~0ull;
        kl[e] = ~0ull;
      }
    }
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = tid; t < N / 2; t += kWG) {
          const int i = 2 * t - (t & (stride - 1));
          const int j = i + stride;
          const bool up = (i & size) == 0;
          const bool gt = !key_le(kh[i], kl[i], kh[j], kl[j]);
          if (gt == up) {
            const uint64_t a = kh[i], c = kl[i];
            kh[i] = kh[j]; kl[i] = kl[j];
            kh[j] = a; kl[j] = c;
          }
        }
        __syncthreads();
      }
    }
    for (int e = tid; e < cnt; e += kWG) pg[e] = PoolEntry{key_err(kh[e]), (int64_t)kl[e]};
    __syncthreads();
  }
}

// thr[g] = min(thr[g], lower[g]) for the G genes (known upper bounds of the
// k-th error supplied by the caller).
__global__ __launch_bounds__(256) void thr_min_kernel(double* __restrict__ thr,
                                                      const double* __restrict__ lower,
                                                      int64_t G) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G && lower[g] < thr[g]) thr[g] = lower[g];
}
