/*
 * abc.h -- C ABI of the MI355X ABC simulation engine (libabc.so).
 *
 * The reference has no FFI: its boundary is the Julia script contract
 * (wrapper.jl:59-78): set the globals m, n_trials, submit and include()
 * scripts/abc_simulation.jl, then scripts/compute_errors.jl, then
 * scripts/accepted_particles.jl.  Each entry point below replaces the body of
 * one of those stages; a Julia shim keeps the globals and the output files and
 * calls these functions through ccall.
 *
 * Conventions
 *  - Host arrays are column-major (Julia Array{Float64}): element (i, k) of an
 *    n x K matrix is at [i + n*k].  The library never keeps a pointer past the
 *    call that received it.
 *  - Particle i of a call uses the random stream of global particle
 *    first + i (Philox4x32-10 counter), so results do not depend on how a run is
 *    split into calls or submits.
 *  - Functions return ABC_OK (0) or a negative code; abc_last_error() gives
 *    the message.  The library never aborts the caller.
 *  - Host-pointer functions are synchronous.  *_dev functions take device
 *    pointers, enqueue on the context stream and return without waiting.
 *  - One context per host thread; a context is not re-entrant.
 */
#ifndef ABC_H
#define ABC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ABC_OK 0
#define ABC_E_ARG (-1)      /* invalid argument                       */
#define ABC_E_HIP (-2)      /* HIP runtime error                      */
#define ABC_E_OOM (-3)      /* device allocation failed               */
#define ABC_E_MODEL (-4)    /* model index outside 1..5               */
#define ABC_E_CAPACITY (-5) /* output buffer too small (counts valid) */
#define ABC_E_STATE (-6)    /* call not allowed now (e.g. inside a graph capture) */

typedef struct abc_ctx abc_ctx;

/* Experimental design.  Replaces the globals read by abc_sim
 * (scripts/abc_simulation.jl:13-14, 64-85) and produced by
 * scripts/load_process_data.jl:50-83 (age bins, pulse/chase split,
 * age_id_distribution) and wrapper.jl:36 (betas). */
typedef struct {
  double cycle_h;               /* cell-cycle length (abc_simulation.jl:72)         */
  int32_t n_bins;               /* age bins = rate bins (n_clusters = 5)            */
  int32_t n_conds;              /* labelling conditions (abc_simulation.jl:65)      */
  const double* pulse_h;        /* [n_conds] pulse lengths                          */
  const double* chase_h;        /* [n_conds] chase times                            */
  const int32_t* is_chase;      /* [n_conds] 0 = pulse cells, 1 = chase cells       */
  int32_t cells_per_particle;   /* simulated cells per parameter set                */
  int32_t age_mode;             /* 0 = bin centres (reference), 1 = theta~U[0,1)    */
  int32_t capture_mode;         /* 0 none, 1 Beta(beta_a,beta_b), 2 empirical list  */
  int32_t beta_a, beta_b;       /* integer Beta parameters, beta_a <= 4             */
  const double* betas;          /* [n_betas] capture efficiencies (mode 2)          */
  const int32_t* beta_bin;      /* [n_betas] 0-based age bin of each efficiency     */
  const int32_t* beta_is_chase; /* [n_betas] 0 pulse cell / 1 chase cell            */
  int64_t n_betas;
  const double* age_weights;    /* [n_bins*n_conds] col-major (age_id_distribution) */
  double burn_eps;              /* burn-in residual target, e.g. 5e-3               */
  int32_t burn_max;             /* burn-in cycle cap, e.g. 16                        */
  int32_t prior_mode;           /* 0 = fix_params ranges, 1 = config priors          */
  double splice_log10;          /* splicing rate for prior_mode 0                   */
  double split_target;          /* expected molecule events per sub-lane (0 = one
                                   lane per cell), e.g. 32768                       */
  int32_t split_max;            /* sub-lanes per cell cap, 1..4096, e.g. 1024, with
                                   cells_per_particle * 2 * split_max < 2^31       */
} abc_design;

/* Context: owns the HIP stream, device buffers and launch shapes.
 * New entry point (the Julia host has no context; the shim creates one per
 * run from the globals above). */
int32_t abc_create(abc_ctx** out, int32_t device, const abc_design* design);
int32_t abc_destroy(abc_ctx* ctx);
const char* abc_last_error(const abc_ctx* ctx);

/* Columns of theta for a model: reference columns (5 or 9) + log10 splicing. */
int32_t abc_n_params(const abc_ctx* ctx, int32_t model);
/* Length of the reference statistics vector: 4*n_bins + 3*n_conds (53). */
int32_t abc_n_stats(const abc_ctx* ctx);

/* Replaces fix_params (scripts/abc_simulation.jl:3-11) for n particles.
 * theta: n x abc_n_params(model), log10 units, reference column order
 * [kon..., koff, alpha..., gamma..., lambda, splice]. */
int32_t abc_sample_prior(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t first, int64_t n,
                         double* theta);

/* Replaces the per-particle body of scripts/abc_simulation.jl:88-97
 * (abc_sim, scripts/abc_simulation.jl:13-62, and the model.jl solves under it).
 *   stats    : n x abc_n_stats col-major, reference order
 *              [pulse mean(n_bins), pulse FF(n_bins), chase mean, chase FF,
 *               ratio(n_conds), mean_corr(n_conds), corr_mean(n_conds)]
 *   moments  : optional n x (n_bins*n_conds*6): per (bin,cond)
 *              [n, mean_u, mean_l, var_u, cov_ul, var_l]
 *   superset : optional n x (2*n_bins*n_conds): mean and Fano factor of u+l
 *   steps    : optional n, Philox draws spent per particle (work counter)
 *   cell_u, cell_l, cell_bin : optional n*cells observed counts per cell */
int32_t abc_simulate(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t first, int64_t n,
                     const double* theta, double* stats, double* moments, double* superset,
                     int64_t* steps, uint32_t* cell_u, uint32_t* cell_l, uint8_t* cell_bin);

/* Replaces compute_trunc_errors (scripts/compute_errors.jl:45-70): the FP64
 * error of every (particle, gene) pair in the reference operation order.
 *   stats: n x K, data/se: G x K, parts: sizes of the summed blocks
 *   ({5,5,5,5,11,11,11} for the reference), err: n x G, clamped at clamp. */
int32_t abc_errors(abc_ctx* ctx, int64_t n, int32_t K, const double* stats, int64_t G,
                   const double* data, const double* se, const int32_t* parts, int32_t n_parts,
                   double clamp, double* err);

/* Replaces the loop of scripts/accepted_particles.jl:12-32 for one model.
 * mode 0: per gene, particles with err <= eps, sorted by (err, index).
 * mode 1: per gene, the k particles with smallest (err, index).
 * Errors are screened with split-BF16 MFMA GEMMs and every particle within a
 * proven margin of the threshold is rescored in FP64, so the sets equal the
 * FP64 ones.  counts: G; idx: 1-based first_index+i+1, gene-major, cap
 * entries; err_out: FP64 errors (optional).  ABC_E_CAPACITY if the total
 * exceeds cap (counts are still written). */
int32_t abc_accept(abc_ctx* ctx, int64_t n, int32_t K, const double* stats, int64_t G,
                   const double* data, const double* se, const int32_t* parts, int32_t n_parts,
                   double clamp, int32_t mode, double eps, int64_t k, int64_t first_index,
                   int64_t* counts, int64_t* idx, double* err_out, int64_t cap);

/* Streaming acceptance over many batches (the loop of accepted_particles.jl
 * run as particles are simulated, so the particle x gene matrix is never
 * stored).  begin uploads the genes' data/se (G x K host arrays) and sets the
 * rule; each add screens one batch (stats n x K; device or host pointer) with
 * the split-BF16 GEMM, rescores its candidates in FP64 and merges them;
 * end writes the per-gene lists like abc_accept and the number of FP64
 * rescorings (optional).  margin_rel widens the screening band by
 * margin_rel * |threshold| on top of the proven bound (0.02 in abc_accept). */
int32_t abc_accept_begin(abc_ctx* ctx, int32_t K, int64_t G, const double* data,
                         const double* se, const int32_t* parts, int32_t n_parts, double clamp,
                         int32_t mode, double eps, int64_t k, double margin_rel);
int32_t abc_accept_add(abc_ctx* ctx, int64_t n, const double* stats, int64_t first_index);
int32_t abc_accept_add_dev(abc_ctx* ctx, int64_t n, const double* stats_dev,
                           int64_t first_index);
int32_t abc_accept_end(abc_ctx* ctx, int64_t* counts, int64_t* idx, double* err_out,
                       int64_t cap, int64_t* n_candidates);

/* Batches whose rows are an arbitrary subset of particles (e.g. those a
 * screen let through): row i is global particle pidx_host[i]. */
int32_t abc_accept_add_idx_dev(abc_ctx* ctx, int64_t n, const double* stats_dev,
                               const int64_t* pidx_host);
/* Lower the per-gene thresholds of the open session to thr[g] (G values):
 * known upper bounds of the k-th error, or the thresholds a screen tests. */
int32_t abc_accept_set_thresholds(abc_ctx* ctx, const double* thr);

/* The screening GEMM alone: approximate error and its proven bound for every
 * (particle, gene) pair (n x G col-major, float).  For validation. */
int32_t abc_screen_errors(abc_ctx* ctx, int64_t n, int32_t K, const double* stats, int64_t G,
                          const double* data, const double* se, const int32_t* parts,
                          int32_t n_parts, double clamp, float* approx, float* bound);

/* The device's elementary functions on n inputs (device pointers x, y), for
 * validation against the CPU restatement: kind 0 = e^x as most kernels take
 * it, 1 = e^x as the burn-in and moment-model loops take it (arguments up to
 * 709.78; three-operand FMA chain), 2 = ln x (table-driven).  No counterpart
 * in the reference: the tests use it to compare the device's bits with the
 * oracle's, the subnormal range included. */
int32_t abc_math_probe_dev(abc_ctx* ctx, int32_t kind, int64_t n, const double* x, double* y);

/* ---- device-resident variants (benchmarks, pipelines that keep data in HBM) */
int32_t abc_set_stream(abc_ctx* ctx, void* hip_stream);   /* NULL = own stream */
int32_t abc_synchronize(abc_ctx* ctx);
int32_t abc_sample_prior_dev(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t first,
                             int64_t n, double* theta_dev);
int32_t abc_simulate_dev(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t first, int64_t n,
                         const double* theta_dev, double* stats_dev, double* moments_dev,
                         double* superset_dev, int64_t* steps_dev);

/* Several models in one call (no reference counterpart; the bodies of several
 * abc_simulation.jl jobs, one per model, at once): the particles of all
 * models share one burn-in launch and one SSA launch, so the machine drains
 * once instead of once per model.  Arrays of n_models (<= 16) entries: model
 * index, first particle, particle count and the device pointers of each
 * model's theta / stats / moments / superset / steps (moments, superset and
 * steps may be NULL, as may the whole moments, superset or steps array).
 * Results equal those of one abc_simulate_dev call per model, bit for bit. */
int32_t abc_simulate_multi_dev(abc_ctx* ctx, int32_t n_models, const int32_t* models,
                               uint64_t seed, const int64_t* first, const int64_t* n,
                               const double* const* theta_dev, double* const* stats_dev,
                               double* const* moments_dev, double* const* superset_dev,
                               int64_t* const* steps_dev);
/* abc_simulate_multi_dev over a subset of each model's particles (no reference
 * counterpart; the particles a screen passes): model m's row i uses the random
 * streams of global particle pidx_dev[m][i] (device int64, < 2^32), so its
 * results equal those of abc_simulate_idx_dev per model, bit for bit. */
int32_t abc_simulate_multi_idx_dev(abc_ctx* ctx, int32_t n_models, const int32_t* models,
                                   uint64_t seed, const int64_t* n,
                                   const int64_t* const* pidx_dev,
                                   const double* const* theta_dev, double* const* stats_dev,
                                   double* const* moments_dev, double* const* superset_dev,
                                   int64_t* const* steps_dev);
/* abc_simulate_dev without device allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results, bit for bit, as abc_simulate_dev; its workspace (rates, queue
 * order, burn-in records, sub-lane offsets, cell counts, counters) is carved
 * from one 256 MiB array per GPU held by the library until the process ends,
 * so a call needs about 77 * cells + 400 bytes per particle: at most about
 * 1,390 particles of 2,500 cells (ABC_E_ARG beyond the limit, with the limit
 * in abc_last_error).  Every call on a GPU uses the same array, and each call
 * clears what it needs cleared: calls that overlap on one GPU (two streams,
 * two contexts or two threads) race and must be serialised by the caller.
 * Its work is not counted by abc_ssa_time. */
int32_t abc_simulate_noalloc_dev(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t first,
                                 int64_t n, const double* theta_dev, double* stats_dev,
                                 double* moments_dev, double* superset_dev, int64_t* steps_dev);
/* Simulation of an arbitrary subset of particles: row i uses the random
 * streams of global particle pidx_dev[i] (device int64, < 2^32), so its
 * results equal those of the same particle in a contiguous run. */
int32_t abc_simulate_idx_dev(abc_ctx* ctx, int32_t model, uint64_t seed, int64_t n,
                             const int64_t* pidx_dev, const double* theta_dev, double* stats_dev,
                             double* moments_dev, double* superset_dev, int64_t* steps_dev);
int32_t abc_errors_dev(abc_ctx* ctx, int64_t n, int32_t K, const double* stats_dev, int64_t G,
                       const double* data_dev, const double* se_dev, const int32_t* parts,
                       int32_t n_parts, double clamp, double* err_dev);

/* Milliseconds of the last simulate/errors call, measured with HIP events on
 * the context stream: [0] prior+rates+order, [1] SSA kernel, [2] statistics
 * kernel, [3] error kernel.  Waits for those events. */
int32_t abc_last_timing(abc_ctx* ctx, double* ms4);

/* SSA work since the last reset, over every simulate call made outside graph
 * capture: ms2[0] milliseconds in burn_kernel (each cell's gene path and
 * Poisson means up to its labelling window), ms2[1] milliseconds in
 * ssa_kernel (the window), both from HIP events on the context stream;
 * *burn_steps the Philox draws burn_kernel spent (graph replays included;
 * the per-particle steps output holds both kernels' draws, so the window's
 * are its sum minus this); *launches the simulate calls timed.  Waits for the
 * pending work; reset != 0 clears all of them. */
int32_t abc_ssa_time(abc_ctx* ctx, int32_t reset, double* ms2, int64_t* burn_steps,
                     int64_t* launches);

/* ---- moment model: the reference's own simulator (scripts/model.jl), the
 * exact first and second moments of the telegraph gene and its unlabelled and
 * labelled mRNA, integrated per particle with adaptive Dormand-Prince 5(4)
 * (the reference uses CVODE_BDF, model.jl:89-96) split at every rate
 * discontinuity.  Results are deterministic and do not depend on the batch.
 * The context supplies the stream and error reporting; the design it was
 * created with is not used. */
typedef struct {
  double cycle;             /* cell-cycle length (h): 20.0 in the reference       */
  int32_t n_steps;          /* rate bins per cycle (n_clusters): 5, <= 8         */
  int32_t n_ages;           /* sampling ages (agevec), <= 8                      */
  const double* ages;       /* n_ages ages (h), host memory                     */
  int32_t n_conds;          /* labelling conditions, <= 16                       */
  const double* pulse;      /* n_conds pulse lengths (h)                         */
  const double* chase;      /* n_conds chase times (h)                           */
  const int32_t* is_chase;  /* n_conds: capture group 0 (pulse) / 1 (chase); NULL
                               = all pulse (only used when downsampling)         */
  double t0;                /* trajectory start: -3 * cycle in the reference     */
  double iv[9];             /* transient start state (E g, E u, E l, Var g, Cov gu,
                               Cov gl, Var u, Cov ul, Var l): abc_simulation.jl:71
                               sets E u = 1/2, recover_statistics.jl:34 E g = 1/2 */
  int32_t transient_mode;   /* 0 = the reference's 1 % rule (model.jl:114-142),
                               1 = the exact periodic state                      */
  double rtol, atol;        /* integrator tolerances, e.g. 1e-10, 1e-12          */
  int32_t method;           /* 0 = Dormand-Prince 5(4) on all nine moments; 1 = the
                               stiff gene block (E g, Var g, Cov gu, Cov gl) in
                               closed form on each smooth piece, Dormand-Prince on
                               the other five (several times fewer steps)          */
} abc_moment_cfg;

/* run_part_sim without downsampling (recover_statistics.jl:2-11, the producer
 * of data/recovered_statistics): theta_dev is n x (reference columns of model:
 * 5 for models 1-2, 9 for 3-5) column-major with leading dimension ld >= n
 * (the engine's theta, whose extra splicing column is ignored, fits as is).
 * out_dev is n x (n_conds * n_ages * 5) column-major; column
 * (cond * n_ages + age) * 5 + k holds k = mean_u, mean_l, var_u, cov_ul, var_l.
 * steps_dev (optional, n) receives the Dormand-Prince steps attempted. */
int32_t abc_recover_dev(abc_ctx* ctx, int32_t model, int64_t n, const double* theta_dev,
                        int64_t ld, const abc_moment_cfg* cfg, double* out_dev,
                        int64_t* steps_dev);
int32_t abc_recover(abc_ctx* ctx, int32_t model, int64_t n, const double* theta, int64_t ld,
                    const abc_moment_cfg* cfg, double* out, int64_t* steps);

/* abc_sim (abc_simulation.jl:13-62) on the moment model: per particle the
 * recovered moments, capture downsampling per (age, pulse/chase) group with
 * capture[(age * 2 + group) * 3 + {0, 1, 2}] = E b, E b^2, Var b (model.jl:
 * 221-239; NULL = no downsampling), then the 4 * n_ages + 3 * n_conds
 * statistics in the reference order (s_pulse mean, FF; s_chase mean, FF;
 * s_ratios; s_mean_corr; s_corr_mean) with the age weights
 * weights[age + n_ages * cond] (age_id_distribution).  The mean/FF rows come
 * from conditions pulse_cond and chase_cond (0-based; 5 and 6 in the
 * reference).  stats_dev is n x K column-major; moments_dev (optional) gets
 * the downsampled moments in the abc_recover_dev layout. */
int32_t abc_moment_sim_dev(abc_ctx* ctx, int32_t model, int64_t n, const double* theta_dev,
                           int64_t ld, const abc_moment_cfg* cfg, const double* capture,
                           const double* weights, int32_t pulse_cond, int32_t chase_cond,
                           double* stats_dev, double* moments_dev, int64_t* steps_dev);
int32_t abc_moment_sim(abc_ctx* ctx, int32_t model, int64_t n, const double* theta, int64_t ld,
                       const abc_moment_cfg* cfg, const double* capture, const double* weights,
                       int32_t pulse_cond, int32_t chase_cond, double* stats, double* moments,
                       int64_t* steps);

/* Replaces get_model_probs and its per-gene branches (scripts/model_probs.jl:
 * 1-63; constant_model_probs.jl and non_constant_model_probs.jl are the same
 * with 2 and 3 groups).  counts is G x M column-major: the particles of model
 * group m accepted for gene g (err <= eps, e.g. the counts of abc_accept).
 * No group accepted: prob, lb, ub = 0; exactly one: 1 for it; otherwise
 * prob = counts / sum and lb, ub = Julia's quantile at 1 - alpha and alpha
 * of the group shares over n_boot resamples of the accepted labels with
 * replacement (the reference: 100 and 0.95).  Resamples come from Philox
 * streams keyed by seed (the reference's Julia RNG is unseeded).  M <= 8,
 * n_boot <= 256, at most 2^32 - 1 accepted particles per gene; outputs are
 * G x M column-major host arrays. */
int32_t abc_model_probs(abc_ctx* ctx, int64_t G, int32_t M, const int64_t* counts, uint64_t seed,
                        int32_t n_boot, double alpha, double* prob, double* lb, double* ub);

/* Replaces get_summary_stats and the functions it calls
 * (scripts/data_summary_statistics.jl:1-196: get_ccp_data, get_ccp_se,
 * get_ratios, get_ratio_se, get_correlations, get_correlation_se): the data's
 * summary statistics and their bootstrap standard errors, for every gene.
 *   u, l         n_cells x n_genes, gene-major (column g = gene g), integer
 *                counts < 2^20 (the reference's u_data, l_data);
 *   age          n_cells age clusters, 0-based (age .- 1), < n_clusters <= 8;
 *   cond         n_cells condition of each cell, 0..n_conds-1 (the position
 *                of experiment[i] in cond_vec), or -1 for none; n_conds <= 16;
 *   pulse, chase 0-based cells of pulse_idx and chase_idx;
 *   age_id_dist  n_clusters x n_conds column-major: the weights of the data's
 *                own correlations (the reference's age_id_distribution);
 *   n_boot       resamples per gene (the reference: 100), seed their Philox
 *                streams (the reference's Julia RNG is unseeded);
 *   out          n_genes x (8 n_clusters + 6 n_conds) column-major, column
 *                blocks in the order of the 14 summary files: pulse_mean,
 *                pulse_ff, pulse_mean_se, pulse_ff_se, chase_mean, chase_ff,
 *                chase_mean_se, chase_ff_se (n_clusters columns each),
 *                ratio_data, ratio_se, mean_corr_data, mean_corr_se,
 *                corr_mean_data, corr_mean_se (n_conds columns each).
 * As in the reference, each resample draws four independent samples of
 * cells: the pulse cells, the chase cells, all cells for the ratios and all
 * cells for the correlations, whose weights are that sample's age-cluster
 * shares within each condition.  Variances use n - 1 and are NaN for one
 * cell; a missing cluster gives 0.  n_cells <= 2^21.  abc_data_stats takes
 * host arrays; abc_data_stats_dev takes u, l and out in device memory (the
 * metadata stays on the host) and returns ABC_E_ARG if a count is 2^20 or
 * more.  Both wait for the result. */
int32_t abc_data_stats(abc_ctx* ctx, int64_t n_genes, int64_t n_cells, const uint32_t* u,
                       const uint32_t* l, int32_t n_clusters, const int32_t* age, int32_t n_conds,
                       const int32_t* cond, const int32_t* pulse, int64_t n_pulse,
                       const int32_t* chase, int64_t n_chase, const double* age_id_dist,
                       int32_t n_boot, uint64_t seed, double* out);
int32_t abc_data_stats_dev(abc_ctx* ctx, int64_t n_genes, int64_t n_cells, const uint32_t* u_dev,
                           const uint32_t* l_dev, int32_t n_clusters, const int32_t* age,
                           int32_t n_conds, const int32_t* cond, const int32_t* pulse,
                           int64_t n_pulse, const int32_t* chase, int64_t n_chase,
                           const double* age_id_dist, int32_t n_boot, uint64_t seed,
                           double* out_dev);

/* ---- surrogate (no reference counterpart: the reference simulates every
 * prior particle, scripts/abc_simulation.jl:88-97).  An MLP maps a
 * particle's features to the mean and log-variance of each statistic (in the
 * standardised asinh space); it is trained on simulator output with BF16
 * MFMA forward/backward kernels, FP32 master weights, Adam and a Gaussian
 * negative log-likelihood, and screens prior particles so that the exact
 * simulation runs only on those that can still be accepted. */
typedef struct abc_mlp abc_mlp;
/* n_in features -> depth x [width, SiLU] -> 2*n_stats outputs (width % 128 == 0). */
int32_t abc_mlp_create(abc_mlp** out, int32_t device, int32_t n_in, int32_t width,
                       int32_t depth, int32_t n_stats, int32_t max_batch, uint64_t seed);
int32_t abc_mlp_destroy(abc_mlp* mlp);
const char* abc_mlp_last_error(const abc_mlp* mlp);
int64_t abc_mlp_n_params(const abc_mlp* mlp);
/* Make the first n_counts statistics counts: their mean output is a log-rate
 * o and their loss the Poisson negative log-likelihood exp(o) - t o on
 * [-30, 8], continued linearly outside; their log-variance output is unused.
 * Default 0. */
int32_t abc_mlp_set_count_outputs(abc_mlp* mlp, int32_t n_counts);
/* Rows of the GEMM workgroup tile: 0 (default) picks 64 when 128-row tiles
 * would give fewer than two workgroups per CU, else 128; 64 or 128 force it.
 * A launch shape only: the results do not depend on it. */
int32_t abc_mlp_set_tile_rows(abc_mlp* mlp, int32_t rows);
/* Columns of the GEMM workgroup tile for outputs of a multiple of 128
 * columns: 0 (default) picks 64 when 128-column tiles would give fewer than
 * four workgroups per CU, else 128; 64 or 128 force it.  A launch shape only. */
int32_t abc_mlp_set_tile_cols(abc_mlp* mlp, int32_t cols);
int32_t abc_mlp_n_stats(const abc_mlp* mlp);
int32_t abc_mlp_set_stream(abc_mlp* mlp, void* hip_stream);
int32_t abc_mlp_synchronize(abc_mlp* mlp);
/* Flat FP32 parameters: per layer W (out_pad x in_pad, row-major) then b (out_pad). */
int32_t abc_mlp_get_params(abc_mlp* mlp, float* host);
int32_t abc_mlp_set_params(abc_mlp* mlp, const float* host);
/* x: n x n_in FP32 row-major (device); y: n x 2*n_stats [mean | log-variance]. */
int32_t abc_mlp_forward_dev(abc_mlp* mlp, int64_t n, const float* x_dev, float* y_dev);
/* Mean NLL and its gradient (flat like the parameters) for one batch, no update. */
int32_t abc_mlp_grad_dev(abc_mlp* mlp, int64_t n, const float* x_dev, const float* t_dev,
                         float* grad_host, float* loss_host);
/* One Adam step on a batch (n <= max_batch); t: n x n_stats targets.  The
 * batch loss is copied to loss_dev when it is not NULL. */
int32_t abc_mlp_train_step_dev(abc_mlp* mlp, int64_t n, const float* x_dev, const float* t_dev,
                               float lr, float beta1, float beta2, float eps, float weight_decay,
                               float* loss_dev);
/* Features of n particles of a model (theta n x abc_n_params, device):
 * one-hot(model) + per-bin log10 rates / 3 -> x (n x (5 + 3*n_bins + 3)). */
int32_t abc_features_dev(int32_t model, int64_t n, const double* theta_dev, int32_t n_bins,
                         float* x_dev, void* hip_stream);
/* Targets: (asinh(stat) - mean) / sd per statistic; stats n x K col-major. */
int32_t abc_targets_dev(int64_t n, const double* stats_dev, int32_t K, const float* mean_dev,
                        const float* sd_dev, float* t_dev, void* hip_stream);
/* Screen n particles against the genes and thresholds of the open acceptance
 * session.  The network's output y = forward(x) gives each statistic a
 * Gaussian in the standardised asinh space; mapped back (7-point
 * Gauss-Hermite) it yields the mean and variance of every gene's error.
 * pass[i] = 1 when for some gene E[err] - z*sd[err] <= thr*(1 + margin), when
 * a threshold reaches the clamp, or when a prediction is not finite. */
int32_t abc_screen_dev(abc_ctx* ctx, abc_mlp* mlp, int64_t n, const float* x_dev,
                       const float* mean_dev, const float* sd_dev, double z, double margin,
                       uint8_t* pass_dev, int64_t* n_pass);

/* The screen's two scores per particle instead of a pass flag, for choosing
 * the pass rule and its threshold: score[2 i] = the expected number of genes
 * that accept particle i, sum over genes of P(err <= thr * (1 + margin))
 * under the same Gaussian approximation of each gene's error, and
 * score[2 i + 1] = the largest standardised margin (thr * (1 + margin) - E)
 * / sd over the genes (abc_screen_dev passes a particle when it is >= -z).
 * A threshold that reaches the clamp, or a prediction that is not finite,
 * gives +inf.  score_dev: n x 2 floats, device. */
int32_t abc_screen_score_dev(abc_ctx* ctx, abc_mlp* mlp, int64_t n, const float* x_dev,
                             const float* mean_dev, const float* sd_dev, double margin,
                             float* score_dev);

/* Launch shape of the persistent SSA kernel: workgroups (0 = occupancy
 * default).  Set by the autotuner. */
int32_t abc_set_ssa_blocks(abc_ctx* ctx, int32_t blocks);
int32_t abc_get_ssa_blocks(const abc_ctx* ctx);
/* Work items (cell sub-lanes) a lane of the SSA kernel claims per atomic on
 * its work queue: 1, 2, 4 or 8 consecutive items (default 2, the fastest on
 * the configured workload).  The first item
 * of a claim is located by binary search, the others by a step from the
 * previous one.  Items are independent, so this changes the schedule only.
 * Set by the autotuner. */
int32_t abc_set_ssa_claim(abc_ctx* ctx, int32_t claim);
int32_t abc_get_ssa_claim(const abc_ctx* ctx);

/* Autotuner of the persistent SSA launch shape (no reference counterpart: the
 * reference runs one particle at a time, scripts/abc_simulation.jl:88).  Draws
 * n prior particles (from first) of each listed model, simulates them (all
 * models in one multi-model call, as abc_simulate_multi_dev) once per
 * candidate grid (1, 2, 3, 4, 6, 8, 12, 16 workgroups per CU, up to
 * occupancy) and then, on the fastest grid, twice per queue claim (2, 4 and 8
 * work items per atomic against 1), times ssa_kernel with HIP events and keeps
 * the fastest shape for later calls.  Outputs never depend on the shape, so
 * tuning changes speed only.  best_blocks / best_ms (optional) receive the
 * grid and its time; abc_get_ssa_claim gives the claim. */
int32_t abc_autotune_ssa(abc_ctx* ctx, int32_t n_models, const int32_t* models, uint64_t seed,
                         int64_t first, int64_t n, int32_t* best_blocks, double* best_ms);

/* HIP graphs (no reference counterpart).  abc_graph_begin starts capturing the
 * context stream; the *_dev calls that follow (sample_prior_dev, simulate_dev,
 * simulate_idx_dev, errors_dev) are recorded instead of run; abc_graph_end
 * closes the capture into *graph: a new executable graph when *graph is NULL,
 * else an in-place update of *graph (same call sequence with new particles,
 * seeds or buffers), re-instantiated if the sequence changed.  Every device
 * buffer a captured call needs must already have its size: run the same calls
 * once uncaptured first (ABC_E_STATE otherwise).  Host-pointer and waiting
 * entry points return ABC_E_STATE during a capture.  Timing events are not
 * recorded inside graphs.  Graphs replay on the context stream. */
typedef struct abc_graph abc_graph;
int32_t abc_graph_begin(abc_ctx* ctx);
int32_t abc_graph_end(abc_ctx* ctx, abc_graph** graph);
int32_t abc_graph_launch(abc_ctx* ctx, abc_graph* graph);
int32_t abc_graph_destroy(abc_graph* graph);

#ifdef __cplusplus
}
#endif
#endif
