/*
 * oracle.h -- CPU oracle for the ABC hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 *
 * It holds two independent restatements:
 *   1. The reference moment model (scripts/model.jl, scripts/abc_simulation.jl,
 *      scripts/compute_errors.jl, scripts/accepted_particles.jl) in plain C:
 *      moment ODE + periodic boundary + transient phase + trajectories +
 *      downsample + the 53 summary statistics + reference-order FP64 errors +
 *      fixed-epsilon acceptance.
 *   2. A scalar CPU replay of the engine's stochastic simulation (Gillespie SSA
 *      over four mRNA species, Philox4x32-10 streams, integer-exact cell
 *      statistics).  The GPU kernels must reproduce its outputs bit for bit.
 *
 * Every function cites the reference file:line it restates.  The SSA replay
 * has no reference counterpart (the reference integrates moments); it is
 * pinned statistically against restatement 1 and against closed forms.
 */
#ifndef ABC_ORACLE_H
#define ABC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_BINS 8
#define ORC_MAX_CONDS 16

/* ------------------------------------------------------------------ RNG */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01_32(uint32_t w);
double orc_u01_53(uint32_t a, uint32_t b);
double orc_log(double x);
float orc_exp1_32(uint32_t w);   /* -ln(((float)w + 0.5f) 2^-32), single precision */
double orc_exp(double x);
int64_t orc_exp_scaling_mismatches(double lo, double hi, int64_t n, uint64_t seed);
double orc_exp10(double x);
double orc_u128_to_f64(uint64_t hi, uint64_t lo);

/* ------------------------------------------------------------- design */
typedef struct {
  double cycle;             /* cell-cycle length T in hours                 */
  int32_t n_bins;           /* age bins (5 in the reference)                */
  int32_t n_conds;          /* labelling conditions                         */
  const double* pulse;      /* [n_conds] pulse length (h)                   */
  const double* chase;      /* [n_conds] chase length (h)                   */
  const int32_t* is_chase;  /* [n_conds] 0 = pulse group, 1 = chase group   */
  int32_t cells;            /* cells per particle                           */
  int32_t age_mode;         /* 0 = bin centres, 1 = theta ~ U[0,1)          */
  int32_t capture_mode;     /* 0 = none, 1 = Beta(a,b), 2 = empirical list  */
  int32_t beta_a, beta_b;   /* integer Beta parameters (capture_mode 1)     */
  const double* betas;      /* grouped empirical list (capture_mode 2)      */
  const int64_t* group_off; /* [2*n_bins+1], group = bin*2 + is_chase      */
  const double* weights;    /* [n_bins*n_conds] col-major age weights       */
  double burn_eps;          /* residual survival target for burn-in         */
  int32_t burn_max;         /* cap on burn-in cycles                        */
  int32_t prior_mode;       /* 0 = reference fix_params, 1 = config priors  */
  double splice_log10;      /* splicing rate used by prior_mode 0           */
  double split_target;      /* molecule events per sub-lane (0 = no split)  */
  int32_t split_max;        /* largest number of sub-lanes per cell         */
} orc_design;

int orc_n_params(int model);   /* reference columns + 1 (splicing) */

/* Prior draws, theta is n x n_params column-major, log10 units. */
void orc_sample_prior(const orc_design* d, int model, uint64_t seed, int64_t first,
                      int64_t n, double* theta);

/* Stochastic simulation replay.  Outputs (all optional except stats):
 *   cell_u, cell_l, cell_bin : n*cells observed counts and age bin per cell
 *   stats_ref   : n x (4*n_bins + 3*n_conds) col-major, reference layout
 *   moments     : n x (n_bins*n_conds*6) col-major: per (bin,cond)
 *                 [n, mean_u, mean_l, var_u, cov_ul, var_l]
 *   superset    : n x (2*n_bins*n_conds): [mean_t(b,j)...] then [ff_t(b,j)...]
 *   steps, reactions : n per-particle work counters
 * Returns 0 or a negative error code. */
int orc_simulate(const orc_design* d, int model, uint64_t seed, int64_t first, int64_t n,
                 const double* theta, int n_threads, uint32_t* cell_u, uint32_t* cell_l,
                 uint8_t* cell_bin, double* stats_ref, double* moments, double* superset,
                 int64_t* steps, int64_t* reactions);

/* Statistics of one particle from its cells (exposed for tests). */
void orc_cell_stats(const orc_design* d, const uint32_t* u, const uint32_t* l,
                    const uint8_t* bin, double* stats_ref, double* moments, double* superset);

/* ---------------------------------------------------- moment model (ref) */
typedef struct {
  double cycle; int32_t n_steps;            /* T and number of rate bins    */
  int32_t n_ages; const double* ages;       /* sampled ages (h)             */
  int32_t n_conds; const double* pulse; const double* chase;
  double t0;                                /* start of trajectories        */
  double iv[9];                             /* initial moments              */
  int32_t transient_mode;                   /* 0 = reference 1% rule,
                                               1 = iterate to fixed point   */
  double rtol, atol;                        /* integrator tolerances        */
  int32_t method;                           /* 0 = Dormand-Prince on all 9,
                                               1 = gene block in closed form */
} orc_moment_cfg;

/* y' = f(y) at time t (model.jl:67-86). */
void orc_moment_rhs(int model, const double* theta, double cycle, int n_steps,
                    double texp, double pulse, double t, const double* y, double* dy);
/* model(): integrate [tmin,tmax] (model.jl:89-96). Returns steps taken. */
int64_t orc_moment_solve(int model, const double* theta, double cycle, int n_steps,
                         double texp, double pulse, double tmin, double tmax, double* y,
                         double rtol, double atol, int method);
void orc_periodic_boundary(const double* in, double* out);
/* n draws of the SSA's Poisson sampler (inversion below 10, PTRS above). */
void orc_poisson_sample(double mu, uint64_t seed, int64_t n, uint32_t* out);
/* burn-in records of cells 0..n_cells-1 of one particle: 8 doubles per cell
 * (age, capture, lu, lt, gene, bin, B, window start) */
int orc_burn_cells(const orc_design* d, int model, uint64_t seed, uint32_t pidx,
                   const double* theta, int n_cells, double* out);
/* one burn-in interval applied to the Poisson means io = {lu, lt} */
void orc_burn_step(double* io, int g, double c0, double c1, double ku, double kt, double D);
/* transient_phase (model.jl:114-142); returns iterations. */
int orc_transient(int model, const double* theta, const orc_moment_cfg* c, double* ss);
/* the 1 % rule by integrating every cycle (checks the affine form) */
int orc_transient_direct(int model, const double* theta, const orc_moment_cfg* c, double* ss);
/* run_part_sim without downsampling (recover_statistics.jl:2-11):
 * out[cond][age][5] = [mean_u, mean_l, var_u, cov_ul, var_l]. */
void orc_recover(int model, const double* theta, const orc_moment_cfg* c, double* out);
/* the same, integrating every cycle of every trajectory (checks the affine
 * shortcut for whole unlabelled cycles) */
void orc_recover_direct(int model, const double* theta, const orc_moment_cfg* c, double* out);
/* orc_recover plus the integration steps of each trajectory [cond * n_ages + age] */
void orc_recover_steps(int model, const double* theta, const orc_moment_cfg* c, double* out,
                       int64_t* traj_steps);
/* downsample (model.jl:221-239): data n_ages x 5 row-major in place. */
void orc_downsample(double* data, int n_ages, const double* mean_b, const double* m2_b,
                    const double* var_b);
/* abc_sim statistics (abc_simulation.jl:23-46) from per-condition moments
 * s[cond][age][5]; w is n_ages x n_conds col-major; pulse_cond/chase_cond are
 * the 0-based conditions feeding s_pulse / s_chase. out has 4*n_ages+3*n_conds. */
void orc_moment_stats(const double* s, int n_ages, int n_conds, const double* w,
                      int pulse_cond, int chase_cond, double* out);
/* batch of abc_sim on the moment model (OpenMP over particles) */
int orc_moment_sim(int model, int64_t n, const double* theta, int64_t ld,
                   const orc_moment_cfg* c, const int32_t* is_chase, const double* cap,
                   const double* w, int pulse_cond, int chase_cond, int n_threads,
                   double* stats, double* moments);

/* ------------------------------------------------ model probabilities */
/* get_model_probs (model_probs.jl:1-29) per gene: counts is G x M col-major
 * (accepted particles per model group); prob, lb, ub are G x M col-major;
 * boot (optional) receives the resampled counts, [gene][resample][group]. */
/* Data-side summary statistics with bootstrap SEs (data_summary_statistics.jl):
 * u, l n x G (gene-major) integer counts <= 2^20, n <= 2^21; age 0..nc-1; cond
 * 0..J-1 or -1; pulse / chase: 0-based cell lists; w_point nc x J.  out:
 * G x (8 nc + 6 J) column-major in the order of the 14 summary files; boot
 * (optional): G x n_boot x (4 nc + 3 J) resampled statistics. */
#define ORC_DATA_MAX_CLUSTERS 8
#define ORC_DATA_MAX_CONDS 16
int orc_data_stats(int64_t G, int64_t n, const uint32_t* u, const uint32_t* l,
                   const int32_t* age, const int32_t* cond, int nc, int J,
                   const int32_t* pulse, int64_t np, const int32_t* chase, int64_t nch,
                   const double* w_point, int n_boot, uint64_t seed, double* out, double* boot,
                   int n_threads);
/* the positions (0..m-1) drawn by resample b of family f for gene g */
void orc_data_resample(int64_t g, int b, int f, int64_t m, uint64_t seed, int64_t* pos);
int orc_model_probs(int64_t G, int M, const int64_t* counts, uint64_t seed, int n_boot,
                    double alpha, double* prob, double* lb, double* ub, int64_t* boot);
/* Julia's quantile(v, p) of a sorted vector (type 7) */
double orc_jl_quantile(const double* v, int n, double p);

/* ------------------------------------------------- errors and acceptance */
/* nlsqerror_part + compute_trunc_errors for one (particle, gene) pair in the
 * reference operation order (compute_errors.jl:30-64). */
double orc_error_pair(const double* d, const double* se, const double* s, int K,
                      const int32_t* parts, int n_parts, double clamp);
/* Full error matrix, err is n x G column-major. */
void orc_errors(int64_t n, const double* stats, int K, int64_t G, const double* data,
                const double* se, const int32_t* parts, int n_parts, double clamp,
                double* err, int n_threads);
/* accepted_particles.jl:20-22 for one gene: indices (1-based) with err <= eps
 * sorted by (err, index).  Returns count; idx/err_out sized n. */
int64_t orc_accept_gene(const double* err_col, int64_t n, double eps, int64_t* idx,
                        double* err_out);

#ifdef __cplusplus
}
#endif
#endif
