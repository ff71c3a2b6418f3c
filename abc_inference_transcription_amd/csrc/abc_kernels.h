// abc_kernels.h -- device data layouts and kernel declarations shared by the
// kernels and the C-ABI host layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abc_device.h"

#define ABC_MAX_BINS 8
#define ABC_MAX_CONDS 16
#define ABC_MAX_STATS 256
#define ABC_MAX_PARTS 16
#define SSA_WG 256
#define ABC_NBUCKET 256      // cost buckets of the queue order
#define ABC_SCAN_BLOCK 1024  // ranks per block of the item scan

// Experimental design in device memory (read through uniform scalar loads).
struct DesignDev {
  double T;                        // cycle length (h)
  int nb, J, C;                    // age bins, conditions, cells per particle
  int age_mode, cap_mode;          // 0 bin centres / 1 uniform; capture 0/1/2
  int beta_a, beta_b;              // integer Beta(a, b) for cap_mode 1
  int bmin, bmax;                  // burn-in cycle bounds
  int prior_mode, pad0;
  double burn_eps;                 // residual survival target of the burn-in
  double splice_log10;             // splicing rate for the reference prior
  double capture_mean;             // mean capture efficiency (cost estimate)
  double split_target;             // molecule events per sub-lane (0 = no split)
  double split_scale;              // capture scale used by the split rule
  double wmax;                     // longest labelling window (pulse + chase), h
  int split_max, pad1;
  double pulse[ABC_MAX_CONDS], chase[ABC_MAX_CONDS];
  int is_chase[ABC_MAX_CONDS];
  double bound[ABC_MAX_BINS];      // T * (nb + j + 1) / nb (thinning bound, h)
  double weights[ABC_MAX_BINS * ABC_MAX_CONDS];  // age weights, col-major nb x J
};

// Per-particle rates (1/h) for each age bin.
struct __align__(16) PRates {
  double kon[ABC_MAX_BINS], koff[ABC_MAX_BINS], amax[ABC_MAX_BINS], gam[ABC_MAX_BINS];
  // 1 / kon, 1 / koff, 1 / (gam + splicing), 1 / gam per bin (the same IEEE
  // quotients the event loops would form at every bin edge)
  double ikon[ABC_MAX_BINS], ikoff[ABC_MAX_BINS], iku[ABC_MAX_BINS], ikt[ABC_MAX_BINS];
  double bs;                       // splicing rate
  double brate;                    // mean over the bins of alpha_max * P(gene on)
  uint64_t thr_lam;                // floor(lambda * 2^32): labelling threshold
  uint64_t thr_pi;                 // floor(P(gene on at cycle start) * 2^32)
  int32_t B, step, scaling;        // burn-in cycles, grid stride, size scaling
  int32_t gene_varies;             // kon differs between bins (burst-frequency model)
  int32_t M, pad[3];               // typical sub-lanes per cell (the queue-order key only)
};

// A cell at the start of its labelling window (burn_kernel -> ssa_kernel).
struct __align__(16) CellStart {
  double age, bc;   // sampling age (h), capture efficiency
  double tg;        // next gene switch (absolute time)
  double lu, lt;    // Poisson means of the unspliced / all unlabelled molecules
  int64_t k;        // last grid edge passed
  uint64_t gctr;    // gene-stream counter
  uint8_t bin, g, gneed;
  uint8_t msh;      // log2 of the cell's sub-lanes M
};

struct ErrSpec {
  int K, n_parts;
  int parts[ABC_MAX_PARTS];
  double clamp;
};

// One entry of a gene's acceptance pool (abc_accept.hip): the FP64 error
// and the global 0-based particle index (the batch row until rescored).
struct PoolEntry {
  double err;
  int64_t idx;
};
#define ABC_POOL_SORT_MAX 2048  // entries per gene sorted on the device

struct Candidate {
  double err;
  int64_t particle;  // global 0-based particle index
  int32_t gene;
  int32_t pad;
};

int abc_layout_n_params(int model, int nb);

__global__ void prior_kernel(const DesignDev* D, int model, uint32_t k0, uint32_t k1,
                             uint32_t first, int64_t n, double* theta);
__global__ void rates_kernel(const DesignDev* D, int model, int64_t n, const double* theta,
                             PRates* R, float* cost);
__global__ void bucket_hist_kernel(const float* cost, int64_t n, unsigned* hist);
__global__ void bucket_scan_kernel(const unsigned* hist, unsigned* cursor);
__global__ void bucket_scatter_kernel(const float* cost, int64_t n, unsigned* cursor,
                                      int32_t* order);
__global__ void items_scan_kernel(const int64_t* pitems, const int32_t* order, int64_t n,
                                  int64_t* off, int64_t* bsum);
__global__ void bsum_scan_kernel(int64_t* bsum, int64_t nblk);
__global__ void add_bsum_kernel(int64_t* off, int64_t n, const int64_t* bsum);
__global__ void fill_ids_kernel(int64_t n, int64_t first, const int64_t* src, uint32_t mtag,
                                int64_t* pidx, uint32_t* mtags);
__global__ void copy_u64_kernel(const unsigned long long* src, unsigned long long* dst,
                                int64_t n);
__global__ void cell_items_kernel(int64_t n, int C, const CellStart* rec, int32_t* coff,
                                  int64_t* pitems);
__global__ void burn_kernel(const DesignDev* D, const PRates* R, const int32_t* order, int64_t n,
                            uint32_t first, uint32_t k0, uint32_t k1, uint32_t mtag_all,
                            const uint32_t* mtags, const double* betas,
                            const int64_t* goff, unsigned long long* queue, CellStart* rec,
                            unsigned long long* steps_out, unsigned long long* react_out,
                            const int64_t* pidx, unsigned long long* burn_tally);
__global__ void ssa_kernel(const DesignDev* D, const PRates* R, const int32_t* order,
                           const int64_t* off, int64_t n, uint32_t first, uint32_t k0,
                           uint32_t k1, uint32_t mtag_all, const uint32_t* mtags,
                           const CellStart* rec,
                           const int32_t* coff,
                           unsigned long long* queue,
                           uint32_t* out_u, uint32_t* out_l, uint8_t* out_bin,
                           unsigned long long* steps_out,
                           const int64_t* pidx, uint32_t claim);
// Moment model (the reference simulator, scripts/model.jl): one call's
// configuration, passed by value as a kernel argument (about 1.9 kB).
#define ABC_MOM_MAX_AGES ABC_MAX_BINS
#define ABC_MOM_MAX_CONDS ABC_MAX_CONDS
struct MomCfgDev {
  double cycle;          // cell-cycle length T
  int n_steps;           // rate bins per cycle
  int n_ages, n_conds;
  int transient_mode;    // 0 = 1 % rule (model.jl:114-142), 1 = exact periodic state
  double t0;             // trajectory start
  double iv[9];          // transient start state
  double rtol, atol;     // Dormand-Prince tolerances
  int method;            // 0 = Dormand-Prince on all nine moments, 1 = gene block exact
  double ages[ABC_MOM_MAX_AGES];
  double pulse[ABC_MOM_MAX_CONDS], chase[ABC_MOM_MAX_CONDS];
  int is_chase[ABC_MOM_MAX_CONDS];
  int downsample;        // apply the capture moments below
  int pulse_cond, chase_cond;  // conditions feeding the mean / Fano statistics (-1: none)
  double cap[ABC_MOM_MAX_AGES * 2 * 3];  // per (age, pulse/chase): E b, E b^2, Var b
  double weights[ABC_MOM_MAX_AGES * ABC_MOM_MAX_CONDS];  // age weights, n_ages x n_conds
};
template <int METHOD>
__global__ void moment_cyclemap_kernel(MomCfgDev cfg, int model, int64_t n,
                                       const double* theta, int64_t ld, const int32_t* order,
                                       double* maps, unsigned long long* cost,
                                       unsigned long long* steps);
__global__ void moment_key_kernel(MomCfgDev cfg, int model, int64_t n, const double* theta,
                                  int64_t ld, unsigned long long* cost);
// particle order for the cycle-map and trajectory kernels: descending cost
#define ABC_MOM_ORDER_WG 1024
#define ABC_MOM_COST_BUCKETS 1024
__global__ void moment_order_kernel(int64_t n, const unsigned long long* cost, int32_t* order);
__global__ void moment_transient_kernel(MomCfgDev cfg, int64_t n, const double* maps,
                                        double* ss);
template <int METHOD>
__global__ void moment_traj_kernel(MomCfgDev cfg, int model, int64_t n,
                                   const double* theta, int64_t ld, const double* ss,
                                   const double* maps, const int32_t* order, double* out,
                                   unsigned long long* steps);
__global__ void moment_stats_kernel(MomCfgDev cfg, int64_t n, double* mom,
                                    double* stats);
// Data-side summary statistics with bootstrap SEs (scripts/data_summary_statistics.jl)
#define ABC_DATA_WG 256
#define ABC_DATA_MAX_CLUSTERS 8
#define ABC_DATA_MAX_CONDS 16
__global__ void data_stats_kernel(int64_t G, int64_t g0, int64_t n, const uint32_t* u,
                                  const uint32_t* l,
                                  const int32_t* age, const int32_t* cond, int nc, int J,
                                  const int32_t* pulse, int64_t np, const int32_t* chase,
                                  int64_t nch, const double* w_point, int n_boot, uint32_t k0,
                                  uint32_t k1, double* boot, double* out, unsigned int* bad_out);
#define ABC_DATA_SORTED_MAX 16384  // cells per family for data_stats_sorted_kernel (32 KiB of draw counters)
#define ABC_DATA_SETS_PER_WG 8    // sets per workgroup of data_stats_sorted_kernel
__global__ void data_pack_kernel(int64_t n, const uint32_t* u, const uint32_t* l, int64_t g0,
                                 const int32_t* pulse, int64_t np, const int32_t* chase,
                                 int64_t nch, const int32_t* perm, const int32_t* grp,
                                 int64_t slots, int64_t genes, unsigned long long* xs,
                                 unsigned int* bad_out);
__global__ void data_stats_sorted_kernel(int64_t genes, int nsb, int64_t g0,
                                         const unsigned long long* xs, int64_t slots, int64_t n,
                                         int64_t np, int64_t nch, int nc, int J,
                                         const double* w_point, int n_boot, int spb, uint32_t k0,
                                         uint32_t k1, double* pt, double* boot);
__global__ void data_se_kernel(int64_t G, int64_t g0, int64_t genes, int nc, int J, int n_boot,
                               const double* pt, const double* boot, double* out);
// Model probabilities with bootstrap bounds (scripts/model_probs.jl)
#define ABC_SEL_WG 256
#define ABC_SEL_MAX_GROUPS 8
#define ABC_SEL_MAX_BOOT 256
#define ABC_SEL_BOOT_PER_WG 8  // resamples per workgroup of model_prob_kernel
__global__ void model_prob_kernel(int64_t G, int64_t g0, int M, const int64_t* counts,
                                  uint32_t k0, uint32_t k1, int n_boot, int rpb, double* shares,
                                  double* prob, double* lb, double* ub);
__global__ void model_prob_bounds_kernel(int64_t G, int64_t g0, int M, const int64_t* counts,
                                         int n_boot, double alpha, const double* shares,
                                         double* prob, double* lb, double* ub);
__global__ void stats_kernel(const DesignDev* D, const uint32_t* cu, const uint32_t* cl,
                             const uint8_t* cb, int64_t n, double* stats_ref, double* moments,
                             double* superset);
__global__ void err_fp64_kernel(const double* stats, int64_t n, const double* data,
                                const double* se, int64_t G, int64_t g0, ErrSpec E, double* err);
__global__ void accept_fp64_kernel(const double* stats, int64_t n, const double* data,
                                   const double* se, int64_t G, int64_t g0, ErrSpec E,
                                   double eps, int64_t first, Candidate* cand, unsigned long long* count,
                                   unsigned long long cap);
__global__ void zero_u64_kernel(unsigned long long* p, int64_t n);
__global__ void math_probe_kernel(int kind, int64_t n, const double* x, double* y);

// screening (abc_gemm.hip)
__global__ void gene_prep_kernel(const double* data, const double* se, int64_t G, int K, int Kp,
                                 int64_t Gpad, __bf16* Bh, __bf16* Bl, double* cg, uint8_t* gflag,
                                 float* Wf);
// surrogate screen (abc_mlp.hip)
__global__ void screen_kernel(const float* y, int64_t n, int K, const float* mean, const float* sd,
                              float z, const float* Wf, const double* data, int64_t G,
                              const double* thr, float margin, double clamp, uint8_t* pass, float* score);
struct abc_mlp;
extern "C" int32_t abc_mlp_forward_dev(abc_mlp* p, int64_t n, const float* x_dev, float* y_dev);
extern "C" int32_t abc_mlp_synchronize(abc_mlp* p);
extern "C" int32_t abc_mlp_n_stats(const abc_mlp* p);
__global__ void particle_prep_kernel(const double* stats, int64_t n, int64_t ld, int K, int Kp,
                                     int64_t npad,
                                     __bf16* Ah, __bf16* Al, uint8_t* pflag);
__global__ void gemm_screen_kernel(const __bf16* Ah, const __bf16* Al, int64_t n, int Kp,
                                   const __bf16* Bh, const __bf16* Bl, int64_t G, const double* cg,
                                   const uint8_t* gflag, const uint8_t* pflag, const double* thr,
                                   double margin_rel, double clamp, Candidate* out,
                                   unsigned long long* count, unsigned long long cap,
                                   float* approx_out, float* bound_out, PoolEntry* pool,
                                   int64_t capg, uint32_t* pool_n);
// per-gene top-k pools (abc_accept.hip)
__global__ void pool_rescore_kernel(const double* scol, int64_t ld, const double* data,
                                    const double* se, int64_t G, ErrSpec E, PoolEntry* pool,
                                    int64_t capg, const uint32_t* pool_n,
                                    const uint32_t* pool_base, int64_t first_index,
                                    const int64_t* pidx);
__global__ void pool_select_kernel(PoolEntry* pool, int64_t capg, uint32_t* pool_n,
                                   uint32_t* pool_base, int64_t G, int64_t k, double* thr,
                                   unsigned long long* n_rescored);
__global__ void pool_sort_kernel(PoolEntry* pool, int64_t capg, const uint32_t* pool_n,
                                 int64_t G);
__global__ void thr_min_kernel(double* thr, const double* lower, int64_t G);
__global__ void rescore_kernel(const double* stats, int64_t n, const double* data,
                               const double* se, int64_t G, ErrSpec E, Candidate* cand,
                               const unsigned long long* count, unsigned long long cap);
