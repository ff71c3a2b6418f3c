/*
 * orc_ssa.c -- scalar CPU replay of the engine's stochastic simulation
 * (TEST INFRASTRUCTURE ONLY: the checker for the GPU kernels, and the CPU
 * baseline leg of bench.py).
 *
 * Semantics restated from the reference:
 *   prior draws ........ scripts/abc_simulation.jl:3-11 (fix_params)
 *   parameter layout ... scripts/model.jl:30-43 (get_vary_map), abc_simulation.jl:82-85
 *   rates at time t .... scripts/model.jl:1-27 (get_rate, size_scaling)
 *   labelling window ... scripts/model.jl:58-64, 183
 *   reaction network ... scripts/model.jl:74-86 (the moment equations of the
 *                        telegraph model with labelled/unlabelled mRNA)
 *   division ........... scripts/model.jl:98-111 (binomial partitioning)
 *   capture ............ scripts/model.jl:221-239 (downsample: binomial
 *                        thinning with a per-cell efficiency)
 *   cell statistics .... scripts/data_summary_statistics.jl:2-38, 70-77,
 *                        110-147, 179-181 (estimators on cells)
 *
 * The reference integrates the exact first and second moments of this
 * network; here the network itself is simulated cell by cell.  Two mRNA
 * species of the reference (u, l) are split into unspliced/spliced; both decay
 * at gamma so u = Uu+Us and l = Lu+Ls follow the reference exactly.
 * Capture thinning is applied to births (a birth is captured with
 * probability beta_c), which has the same law as Binomial(N, beta_c) at
 * sampling because molecules never feed back on the gene.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ABC_LOGTAB_QUAL static const
#include "abc_logtab.h"
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

typedef struct {
  int n_kon, n_al, n_ga;
  int o_kon, o_koff, o_al, o_ga, o_lam, o_spl, n_par;
} layout_t;

static layout_t layout_of(int model, int nb) {
  layout_t L;
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

int orc_n_params(int model) { return layout_of(model, 5).n_par; }

/* key (seed_lo, seed_hi) for every model; key[2] = model << 24 goes into the
 * counter's cell word (cells < 2^24), 0xFF000000 | model into the prior's */
static void key_of(uint64_t seed, int model, uint32_t key[3]) {
  key[0] = (uint32_t)seed;
  key[1] = (uint32_t)(seed >> 32);
  key[2] = (uint32_t)model << 24;
}

/* ---------------------------------------------------------------- prior */
typedef struct {
  uint32_t key[3]; uint32_t pidx; int q; uint32_t w[4];
} prior_stream;

static double prior_u(prior_stream* s) {
  int call = s->q >> 1, half = s->q & 1;
  if (half == 0) {
    uint32_t ctr[4] = {(uint32_t)call, 0u, 0xFF000000u | (s->key[2] >> 24), s->pidx};
    orc_philox4x32_10(ctr, s->key, s->w);
  }
  s->q++;
  return orc_u01_53(s->w[2 * half], s->w[2 * half + 1]);
}

static double unif(prior_stream* s, double lo, double hi) {
  double u = prior_u(s);
  return lo + (hi - lo) * u;
}

#define FC_LO (-0.698970004336018804786)
#define FC_HI (0.698970004336018804786)

void orc_sample_prior(const orc_design* d, int model, uint64_t seed, int64_t first,
                      int64_t n, double* theta) {
  int nb = d->n_bins;
  layout_t L = layout_of(model, nb);
  for (int64_t i = 0; i < n; ++i) {
    prior_stream s;
    key_of(seed, model, s.key);
    s.pidx = (uint32_t)(first + i);
    s.q = 0;
    double th[2 * ORC_MAX_BINS + 8];
    if (d->prior_mode == 0) {
      /* fix_params: abc_simulation.jl:4-10, column order kon, koff, alpha, gamma, lambda */
      for (int c = 0; c < L.n_kon; ++c) th[L.o_kon + c] = unif(&s, -3.0, 3.0);
      th[L.o_koff] = unif(&s, -3.0, 3.0);
      for (int c = 0; c < L.n_al; ++c) th[L.o_al + c] = unif(&s, -3.0, 3.0);
      for (int c = 0; c < L.n_ga; ++c) th[L.o_ga + c] = unif(&s, -3.0, 2.0);
      th[L.o_lam] = unif(&s, -0.7, 0.0);
      th[L.o_spl] = d->splice_log10;
    } else {
      /* burst frequency, koff, burst size (alpha = size*koff), decay with
       * per-bin log-uniform fold changes 0.2..5 for the varying rate. */
      double kb = unif(&s, -2.0, 2.0);
      if (L.n_kon > 1)
        for (int c = 0; c < nb; ++c) th[L.o_kon + c] = kb + unif(&s, FC_LO, FC_HI);
      else
        th[L.o_kon] = kb;
      double koff = unif(&s, -3.0, 3.0);
      th[L.o_koff] = koff;
      double bb = unif(&s, -1.0, 3.0);
      if (L.n_al > 1)
        for (int c = 0; c < nb; ++c) th[L.o_al + c] = (bb + unif(&s, FC_LO, FC_HI)) + koff;
      else
        th[L.o_al] = bb + koff;
      double gb = unif(&s, -2.0, 1.0);
      if (L.n_ga > 1)
        for (int c = 0; c < nb; ++c) th[L.o_ga + c] = gb + unif(&s, FC_LO, FC_HI);
      else
        th[L.o_ga] = gb;
      th[L.o_lam] = unif(&s, -0.7, 0.0);
      th[L.o_spl] = unif(&s, -1.0, 2.0);
    }
    for (int c = 0; c < L.n_par; ++c) theta[i + n * c] = th[c];
  }
}

/* ------------------------------------------------------ particle rates */
typedef struct {
  double kon[ORC_MAX_BINS], koff[ORC_MAX_BINS], amax[ORC_MAX_BINS], gam[ORC_MAX_BINS];
  double bound[ORC_MAX_BINS];  /* T * (1 + (j+1)/nb): thinning bound in hours */
  double bs;
  uint64_t thr_lam, thr_pi;
  int B, step, scaling, gene_varies;
  double brate; /* mean over the bins of alpha_max * P(gene on): the split rule */
} prates;

static double window_max(const orc_design* d) {
  double wmax = 0.0;
  for (int j = 0; j < d->n_conds; ++j) {
    double w = d->pulse[j] + d->chase[j];
    if (w > wmax) wmax = w;
  }
  return wmax;
}

static int burn_min(const orc_design* d) {
  int b = (int)ceil(window_max(d) / d->cycle);
  return b < 1 ? 1 : b;
}

static void make_rates(const orc_design* d, int model, const double* th, prates* r) {
  int nb = d->n_bins;
  layout_t L = layout_of(model, nb);
  double h = d->cycle / (double)nb;
  r->scaling = model != 2;
  r->step = model == 2 ? nb : 1;
  r->gene_varies = L.n_kon > 1;
  for (int j = 0; j < nb; ++j) {
    r->kon[j] = orc_exp10(th[L.o_kon + (L.n_kon > 1 ? j : 0)]);
    r->koff[j] = orc_exp10(th[L.o_koff]);
    double al = orc_exp10(th[L.o_al + (L.n_al > 1 ? j : 0)]);
    /* size scaling (model.jl:5,25-27): alpha(t) = alpha * (1 + a/T) within a
     * cycle; bin j is thinned from its end-of-bin bound alpha*(nb+j+1)/nb */
    r->amax[j] = r->scaling ? (al * (double)(nb + j + 1)) / (double)nb : al;
    r->bound[j] = (d->cycle * (double)(nb + j + 1)) / (double)nb;
    r->gam[j] = orc_exp10(th[L.o_ga + (L.n_ga > 1 ? j : 0)]);
  }
  r->bs = orc_exp10(th[L.o_spl]);
  double lam = orc_exp10(th[L.o_lam]);
  r->thr_lam = (uint64_t)(lam * 4294967296.0);
  double pi;
  if (L.n_kon > 1) {
    /* periodic stationary on-probability at cycle start: the one-bin map
     * p -> pi_j + (p - pi_j) e_j composed over the cycle, then its fixed point */
    double p = 0.0, E = 1.0;
    for (int j = 0; j < nb; ++j) {
      double rr = r->kon[j] + r->koff[j];
      double pj = r->kon[j] / rr;
      double e = orc_exp(-(rr * h));
      p = pj + (p - pj) * e;
      E = E * e;
    }
    pi = p / (1.0 - E);
  } else {
    pi = r->kon[0] / (r->kon[0] + r->koff[0]);
  }
  if (pi > 1.0) pi = 1.0;
  if (pi < 0.0) pi = 0.0;
  r->thr_pi = (uint64_t)(pi * 4294967296.0);
  double gsum = r->gam[0];
  for (int j = 1; j < nb; ++j) gsum = gsum + r->gam[j];
  double s = 0.5 * orc_exp(-(gsum * h));
  int bmin = burn_min(d);
  double S = 1.0;
  for (int i = 0; i < bmin; ++i) S = S * s;
  int B = bmin;
  while (S > d->burn_eps && B < d->burn_max) { S = S * s; B++; }
  r->B = B;
  {
    double births = 0.0;
    for (int j = 0; j < nb; ++j) births = births + r->amax[j] * (r->kon[j] / (r->kon[j] + r->koff[j]));
    r->brate = births / (double)nb;
  }
}

/* --------------------------------------------------------------- cells */
static inline uint32_t lemire(uint32_t w, uint32_t n) {
  return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

/* Stream s of a cell: counter (draw_lo, draw_hi[15:0] | s << 16,
 * cell | model << 24, particle).  Stream 0 holds the cell set-up draws and the gene clock and is
 * replayed identically by every sub-lane; stream 1 + sub holds sub-lane sub's
 * molecule events and division coin flips. */
typedef struct {
  uint32_t key[2]; uint32_t cell, pidx, stream; uint64_t ctr;
} cell_stream;

static void draw(cell_stream* s, uint32_t w[4]) {
  uint32_t c[4] = {(uint32_t)s->ctr, ((uint32_t)(s->ctr >> 32) & 0xFFFFu) | (s->stream << 16),
                   s->cell, s->pidx};
  orc_philox4x32_10(c, s->key, w);
  s->ctr++;
}

static inline double grid_t(int64_t k, int64_t nbB, double T, int nb) {
  const double h = T / (double)nb;   /* hours between rate-bin edges */
  return (double)(k - nbB) * h;
}

/* Exponential waiting time with rate a from one word. */

/* ---------------------------------------------------- burn-in, Poisson
 * Before the labelling window molecules are never labelled and never act on
 * the gene or on each other, so given the gene path the observed unspliced
 * and spliced molecules alive at the window start are independent Poisson
 * counts (births: a Poisson process of rate alpha(t) beta / M while the gene is
 * on; decay, splicing and the halving at division thin every molecule
 * independently; Jahnke & Huisinga, J. Math. Biol. 54, 2007).  The burn-in
 * therefore simulates the gene path only and carries the two Poisson means:
 * lu (unspliced, leaving at gamma + splice) and lt (all, leaving at gamma). */

/* E1 = int_0^D e^{-k(D-s)} ds and E2 = int_0^D s e^{-k(D-s)} ds, e = e^{-kD},
 * ik = 1/k */
static void seg_int(double k, double ik, double D, double e, double* E1, double* E2) {
  const double x = k * D;
  if (x < 1e-3) {
    *E1 = D * (1.0 - x * (0.5 - x * (1.0 / 6.0 - x * (1.0 / 24.0))));
    *E2 = (D * D) * (0.5 - x * (1.0 / 6.0 - x * (1.0 / 24.0 - x * (1.0 / 120.0))));
  } else {
    *E1 = (1.0 - e) * ik;
    *E2 = (D - *E1) * ik;
  }
}

/* Advance the Poisson means over [t, t + D]: birth rate c0 + c1 s while the
 * gene is on; ku, kt are the exit rates of the unspliced and of all molecules. */
static void burn_accum(double* lu, double* lt, int g, double c0, double c1, double ku,
                       double kt, double D) {
  if (!(D > 0.0)) return;
  const double eu = orc_exp(-(ku * D)), et = orc_exp(-(kt * D));
  *lu = *lu * eu;
  *lt = *lt * et;
  if (g) {
    double E1u, E2u, E1t, E2t;
    seg_int(ku, 1.0 / ku, D, eu, &E1u, &E2u);
    seg_int(kt, 1.0 / kt, D, et, &E1t, &E2t);
    *lu = *lu + (c0 * E1u + c1 * E2u);
    *lt = *lt + (c0 * E1t + c1 * E2t);
  }
}

typedef struct { uint32_t w[4]; int i; } wordbuf;  /* i = next word; 4 = empty */

static uint32_t next_word(cell_stream* s, wordbuf* b, int64_t* st) {
  if (b->i == 4) { draw(s, b->w); (*st)++; b->i = 0; }
  return b->w[b->i++];
}

/* ln(k!) for integer-valued k >= 0: table below 10, Stirling series above */
static double logfact(double kd) {
  if (kd < 10.0) return abc_logfact[(int)kd];
  const double x = kd + 1.0, ix = 1.0 / x, ix2 = ix * ix;
  return ((x - 0.5) * orc_log(x) - x) +
         (0x1.d67f1c864beb5p-1 + ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0 - ix2 * (1.0 / 1680.0)))));
}

/* Poisson(mu): inversion below 10, else the transformed rejection method with
 * squeeze (Hormann, Insurance: Math. Econ. 12, 1993), at most 1000 tries. */
static uint32_t poisson(cell_stream* s, wordbuf* b, double mu, int64_t* st) {
  if (!(mu > 0.0)) return 0u;
  if (mu < 10.0) {
    const uint32_t w0 = next_word(s, b, st), w1 = next_word(s, b, st);
    const double u = orc_u01_53(w0, w1);
    double p = orc_exp(-mu), F = p;
    uint32_t k = 0;
    while (u > F && k < 200u) {
      k++;
      p = (p * mu) * abc_recip_k[k];   /* RN(1/k), generated table */
      F = F + p;
    }
    return k;
  }
  const double lmu = orc_log(mu), smu = orc_exp(0.5 * lmu);
  const double bb = 0.931 + 2.53 * smu, aa = -0.059 + 0.02483 * bb;
  const double inva = 1.1239 + 1.1328 / (bb - 3.4), vr = 0.9277 - 3.6224 / (bb - 2.0);
  for (int tries = 0; tries < 1000; ++tries) {
    const double U = orc_u01_32(next_word(s, b, st)) - 0.5;
    const double V = orc_u01_32(next_word(s, b, st));
    const double us = 0.5 - fabs(U);
    const double kd = floor(((2.0 * aa) / us + bb) * U + mu + 0.43);
    if (kd < 0.0) continue;
    if (us >= 0.07 && V <= vr) return (uint32_t)kd;
    if (us < 0.013 && V > us) continue;
    if ((orc_log(V) + orc_log(inva)) - orc_log(aa / (us * us) + bb) <= (kd * lmu - mu) - logfact(kd))
      return (uint32_t)kd;
  }
  return (uint32_t)floor(mu + 0.5);
}

/* Test hook: n Poisson(mu) draws from the stream (key, cell = 0, pidx = 0,
 * stream 1), consumed as in a window start. */
void orc_poisson_sample(double mu, uint64_t seed, int64_t n, uint32_t* out) {
  cell_stream s = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, 0u, 0u, 1u, 0};
  wordbuf wb = {{0, 0, 0, 0}, 4};
  int64_t st = 0;
  for (int64_t i = 0; i < n; ++i) out[i] = poisson(&s, &wb, mu, &st);
}

/* Test hook: one burn-in interval of length D applied to the Poisson means
 * (lu, lt) = io[0..1], with the gene on (g != 0) or off. */
void orc_burn_step(double* io, int g, double c0, double c1, double ku, double kt, double D) {
  burn_accum(&io[0], &io[1], g, c0, c1, ku, kt, D);
}

/* A cell at the start of its labelling window, shared by its sub-lanes. */
typedef struct {
  double age, bc, tg, lu, lt;
  int64_t k;         /* last grid edge passed */
  uint64_t gctr;     /* gene-stream counter */
  int bin, g, gneed;
  int M;             /* sub-lanes of this cell */
} cell_start;

/* Burn-in of cell c (once per cell): set-up draws, then the gene path from
 * the start of the burn-in to the window start ws with the Poisson means of
 * the cell's unspliced (lu) and all (lt) observed molecules.  The gene clock
 * is a first-reaction clock on stream 0, redrawn after every switch and at
 * every edge where the gene's rate changes. */
static int burn_cell(const orc_design* d, const prates* r, const uint32_t key[3],
                     uint32_t pidx, int c, cell_start* out, int64_t* steps,
                     int64_t* reactions) {
  const int nb = d->n_bins, J = d->n_conds;
  const double T = d->cycle;
  cell_stream s0 = {{key[0], key[1]}, (uint32_t)c | key[2], pidx, 0u, 0};
  uint32_t w[4], v[4];
  draw(&s0, w);
  int bin;
  double age;
  if (d->age_mode == 1) {
    double th = orc_u01_53(w[0], w[1]);
    age = th * T;
    bin = (int)(th * (double)nb);
    if (bin >= nb) bin = nb - 1;
  } else {
    bin = (c / J) % nb;
    age = ((double)(2 * bin + 1) / (double)(2 * nb)) * T;
  }
  int g = (uint64_t)w[2] < r->thr_pi;
  int cond = c % J;
  double bc = 1.0;
  if (d->capture_mode == 2) {
    int grp = bin * 2 + d->is_chase[cond];
    int64_t cnt = d->group_off[grp + 1] - d->group_off[grp];
    if (cnt <= 0) return -3;
    bc = d->betas[d->group_off[grp] + (int64_t)lemire(w[3], (uint32_t)cnt)];
  } else if (d->capture_mode == 1) {
    /* Beta(a, b) = a-th smallest of a+b-1 uniforms */
    int a = d->beta_a, nu = d->beta_a + d->beta_b - 1;
    uint32_t lo[32];
    int have = 0;
    for (int q = 0; q < nu; ++q) {
      if ((q & 3) == 0) draw(&s0, v);
      uint32_t x = v[q & 3];
      int pos = have < a ? have : a;
      while (pos > 0 && lo[pos - 1] > x) {
        if (pos < a) lo[pos] = lo[pos - 1];
        pos--;
      }
      if (pos < a) lo[pos] = x;
      if (have < a) have++;
    }
    bc = orc_u01_32(lo[a - 1]);
  }
  const double ws = age - d->pulse[cond] - d->chase[cond];
  const int64_t nbB = (int64_t)r->B * nb;
  int64_t k = 0, knext = r->step;
  int j = 0;
  double t = grid_t(0, nbB, T, nb), tg = 0.0, lu = 0.0, lt = 0.0;
  int gneed = 1;
  int64_t st = 0, re = 0;
  for (;;) {
    double tb = age;
    double gk = grid_t(knext, nbB, T, nb);
    if (gk < tb) tb = gk;
    if (ws < tb) tb = ws;
    if (gneed) {
      /* waiting time = Exp(1) * (1 / rate); the reciprocal is a per-bin constant */
      draw(&s0, w);
      st++;
      tg = t + (double)orc_exp1_32(w[0]) * (1.0 / (g ? r->koff[j] : r->kon[j]));
      gneed = 0;
    }
    /* births at rate amax_j bc (T + a) / bound_j (size scaling) or amax_j bc,
     * a = time since the cycle start */
    {
      const double t1 = tg < tb ? tg : tb;
      const double atx = r->amax[j] * bc;
      double c0 = atx, c1 = 0.0;
      if (r->scaling) {
        const double cs = grid_t(k - j, nbB, T, nb);
        c1 = atx / r->bound[j];
        c0 = c1 * (T + (t - cs));
      }
      burn_accum(&lu, &lt, g, c0, c1, r->gam[j] + r->bs, r->gam[j], t1 - t);
    }
    if (tg < tb) {
      t = tg;
      g = !g;
      re++;
      gneed = 1;
      continue;
    }
    t = tb;
    if (tb == gk) {
      k = knext;
      knext += r->step;
      j = (int)(k % nb);
      if (j == 0) {
        /* division halves the Poisson means (each molecule kept w.p. 1/2) */
        lu = lu * 0.5;
        lt = lt * 0.5;
      }
      if (r->gene_varies) gneed = 1;  /* the gene's rate changed: restart its clock */
    }
    if (tb == ws) break;
  }
  out->age = age; out->bc = bc; out->tg = tg; out->lu = lu; out->lt = lt;
  out->k = k; out->gctr = s0.ctr; out->bin = bin; out->g = g; out->gneed = gneed;
  /* Sub-lanes of this cell: the smallest power of two M with the cell's
   * expected molecule events in its window <= split_target * M (up to
   * split_max): the unlabelled molecules at the window start (at most 2
   * events each) and births at the cell's capture over its window (3 events
   * each).  Molecules never act on the gene, so M independent 1/M-thinned
   * birth streams and 1/M Poisson shares of the unlabelled molecules over
   * one shared gene path have the law of the whole cell. */
  {
    int M = 1;
    if (d->split_target > 0.0) {
      const double w = d->pulse[cond] + d->chase[cond];
      const double est = (lu + lt) + ((r->brate * bc) * w) * 3.0;
      while (M < d->split_max && est > d->split_target * (double)M) M *= 2;
    }
    out->M = M;
  }
  *steps += st;
  *reactions += re;
  return 0;
}

/* Labelling window of sub-lane sub of cell c: its unlabelled molecules are
 * Poisson(lu / M) unspliced and Poisson((lt - lu) / M) spliced, drawn from
 * its own stream 1 + sub; then the four-species simulation from ws to the
 * sampling age over the cell's gene path (stream 0, continued from the
 * burn-in) with a 1/M share of the births. */
static int sim_window(const orc_design* d, const prates* r, const uint32_t key[3],
                      uint32_t pidx, int c, int sub, const cell_start* cst, uint32_t* u_out,
                      uint32_t* l_out, int64_t* steps, int64_t* reactions) {
  const int nb = d->n_bins, J = d->n_conds;
  const double T = d->cycle;
  cell_stream s0 = {{key[0], key[1]}, (uint32_t)c | key[2], pidx, 0u, cst->gctr};
  cell_stream sm = {{key[0], key[1]}, (uint32_t)c | key[2], pidx, 1u + (uint32_t)sub, 0};
  uint32_t w[4];
  const int cond = c % J;
  const double age = cst->age, bc = cst->bc;
  const double we = age - d->chase[cond];
  const int64_t nbB = (int64_t)r->B * nb;
  const double inv_share = (double)cst->M;
  int64_t st = 0, re = 0;
  uint32_t n0, n1, n2 = 0, n3 = 0;
  {
    wordbuf wb = {{0, 0, 0, 0}, 4};
    const double ls = cst->lt - cst->lu;
    n0 = poisson(&sm, &wb, cst->lu / inv_share, &st);
    n1 = poisson(&sm, &wb, ls > 0.0 ? ls / inv_share : 0.0, &st);
  }
  int64_t k = cst->k, knext = cst->k + r->step;
  int j = (int)(k % nb);
  double t = age - d->pulse[cond] - d->chase[cond], tg = cst->tg;
  int g = cst->g, in_win = 1;
  /* Per iteration one Philox block is drawn: division coins, the gene clock
   * (first-reaction method on the shared stream 0) or a new molecule
   * candidate (direct method on stream 1 + sub).  A candidate pre-empted by
   * the gene clock or a boundary stays pending: its words are still unused and
   * its Exp(1) hazard, less what the interval used up, is again Exp(1)
   * (memoryless).  A gene switch only marks the clock for redrawing. */
  int gneed = cst->gneed, pend = 0, div = 0, div_s = 0;
  uint32_t div_rem = 0, div_keep = 0, pw[3] = {0, 0, 0};
  double erem = 0.0;
  for (;;) {
    double tb = age;
    double gk = grid_t(knext, nbB, T, nb);
    if (gk < tb) tb = gk;
    if (in_win && we < tb) tb = we;
    /* molecule channels of this sub-lane (direct method) */
    double a_tx = g ? (r->amax[j] * bc) / inv_share : 0.0;
    uint32_t nun = n0 + n2, ntot = n0 + n1 + n2 + n3;
    double a_spl = r->bs * (double)nun;
    double a_dec = r->gam[j] * (double)ntot;
    double c1 = a_tx, c2 = c1 + a_spl, a0 = c2 + a_dec;
    int kind = div ? 1 : (gneed ? 2 : ((!pend && a0 > 0.0) ? 3 : 0));
    if (kind == 1 || kind == 3) { draw(&sm, w); st++; }
    else if (kind == 2) { draw(&s0, w); st++; }
    if (kind == 1) {
      /* binomial(count, 1/2) partition of species div_s, 128 coins per block */
      uint32_t* sp[4] = {&n0, &n1, &n2, &n3};
      uint32_t m = div_rem < 128u ? div_rem : 128u;
      for (int b = 0; b < 4; ++b) {
        int bits = (int)m - 32 * b;
        if (bits <= 0) break;
        uint32_t x = bits >= 32 ? w[b] : (w[b] & ((1u << bits) - 1u));
        div_keep += (uint32_t)__builtin_popcount(x);
      }
      div_rem -= m;
      if (div_rem == 0u) {
        *sp[div_s] = div_keep;
        div_s++;
        while (div_s < 4 && *sp[div_s] == 0u) div_s++;
        if (div_s < 4) { div_rem = *sp[div_s]; div_keep = 0; }
        else div = 0;
      }
      continue;
    }
    if (kind == 2) {
      tg = t + (double)orc_exp1_32(w[0]) * (1.0 / (g ? r->koff[j] : r->kon[j]));
      gneed = 0;
      if (!pend && a0 > 0.0) continue;   /* a molecule candidate comes first */
    } else if (kind == 3) {
      erem = (double)orc_exp1_32(w[0]);
      pw[0] = w[1]; pw[1] = w[2]; pw[2] = w[3];
      pend = 1;
    }
    double tm = (pend && a0 > 0.0) ? t + erem / a0 : INFINITY;
    if (tm < tg && tm < tb) {
      t = tm;
      pend = 0;
      double rr = orc_u01_32(pw[0]) * a0;
      if (rr < c1) {
        int acc = 1;
        if (r->scaling) {
          /* accept with probability (T + a) / bound_j, a = time since cycle start,
           * formed as t + (T - cycle start) */
          double cs = grid_t(k - j, nbB, T, nb);
          acc = orc_u01_32(pw[1]) * r->bound[j] < t + (T - cs);
        }
        if (acc) {
          re++;
          if (in_win && (uint64_t)pw[2] < r->thr_lam) n2++; else n0++;
        }
      } else if (rr < c2) {
        re++;
        if (lemire(pw[1], nun) < n0) { n0--; n1++; } else { n2--; n3++; }
      } else {
        re++;
        uint32_t x = lemire(pw[1], ntot);
        if (x < n0) n0--;
        else if (x < n0 + n1) n1--;
        else if (x < n0 + n1 + n2) n2--;
        else n3--;
      }
      continue;
    }
    if (pend) {
      /* no molecule event before min(tg, tb): a0 times that interval is used up */
      const double t1 = tg < tb ? tg : tb;
      erem = erem - a0 * (t1 - t);
      if (erem < 0.0) erem = 0.0;
    }
    if (tg < tb) {
      t = tg;
      g = !g;
      re++;
      gneed = 1;
      continue;
    }
    /* boundary: sampling age, rate-bin / cycle edge, or the window end */
    t = tb;
    if (tb == age) break;
    if (tb == gk) {
      k = knext;
      knext += r->step;
      j = (int)(k % nb);
      if (j == 0) {
        /* division: partition each non-empty species in order, by blocks of
         * 128 coins drawn by the following iterations */
        uint32_t* sp[4] = {&n0, &n1, &n2, &n3};
        div_s = 0;
        while (div_s < 4 && *sp[div_s] == 0u) div_s++;
        if (div_s < 4) { div = 1; div_rem = *sp[div_s]; div_keep = 0; }
      }
      if (r->gene_varies) gneed = 1;   /* the gene's rate changed: restart its clock */
    }
    if (in_win && tb == we) in_win = 0;
  }
  *u_out = n0 + n1;
  *l_out = n2 + n3;
  *steps += st;
  *reactions += re;
  return 0;
}

/* Test hook: the burn-in of cells 0 .. n_cells-1 of particle pidx (theta: its
 * log10 rates).  out[8 c ...] = age, capture, lu, lt, gene state, bin, burn-in
 * cycles B, window start ws. */
int orc_burn_cells(const orc_design* d, int model, uint64_t seed, uint32_t pidx,
                   const double* theta, int n_cells, double* out) {
  if (model < 1 || model > 5 || n_cells > d->cells) return -1;
  uint32_t key[3];
  key_of(seed, model, key);
  prates r;
  make_rates(d, model, theta, &r);
  for (int c = 0; c < n_cells; ++c) {
    cell_start cst;
    int64_t st = 0, re = 0;
    const int e = burn_cell(d, &r, key, pidx, c, &cst, &st, &re);
    if (e) return e;
    const int cond = c % d->n_conds;
    double* o = out + 8 * (size_t)c;
    o[0] = cst.age;
    o[1] = cst.bc;
    o[2] = cst.lu;
    o[3] = cst.lt;
    o[4] = cst.g;
    o[5] = cst.bin;
    o[6] = r.B;
    o[7] = cst.age - d->pulse[cond] - d->chase[cond];
  }
  return 0;
}

static int sim_cell(const orc_design* d, const prates* r, const uint32_t key[3],
                    uint32_t pidx, int c, uint32_t* u_out, uint32_t* l_out,
                    uint8_t* bin_out, int64_t* steps, int64_t* reactions) {
  cell_start cst;
  int e = burn_cell(d, r, key, pidx, c, &cst, steps, reactions);
  if (e) return e;
  uint32_t u = 0, l = 0;
  for (int sub = 0; sub < cst.M; ++sub) {
    uint32_t su, sl;
    e = sim_window(d, r, key, pidx, c, sub, &cst, &su, &sl, steps, reactions);
    if (e) return e;
    u += su;
    l += sl;
  }
  *u_out = u;
  *l_out = l;
  *bin_out = (uint8_t)cst.bin;
  return 0;
}

/* --------------------------------------------------------- statistics */
static double f128(u128 v) { return orc_u128_to_f64((uint64_t)(v >> 64), (uint64_t)v); }

static double var_of(uint64_t n, uint64_t S, u128 SS) {
  u128 num = (u128)n * SS - (u128)S * S;
  return f128(num) / ((double)n * (double)(n - 1));
}

static double cov_of(uint64_t n, uint64_t Sa, uint64_t Sb, u128 Sab) {
  u128 a = (u128)n * Sab, b = (u128)Sa * Sb;
  double den = (double)n * (double)(n - 1);
  if (a >= b) return f128(a - b) / den;
  return -(f128(b - a) / den);
}

/* weighted_cov, data_summary_statistics.jl:179-181 */
static double wcov(const double* x, const double* y, const double* w, int n) {
  double mx = w[0] * x[0], my = w[0] * y[0];
  for (int b = 1; b < n; ++b) { mx = mx + w[b] * x[b]; }
  for (int b = 1; b < n; ++b) { my = my + w[b] * y[b]; }
  double acc = w[0] * ((x[0] - mx) * (y[0] - my));
  for (int b = 1; b < n; ++b) acc = acc + w[b] * ((x[b] - mx) * (y[b] - my));
  return acc;
}

void orc_cell_stats(const orc_design* d, const uint32_t* u, const uint32_t* l,
                    const uint8_t* bin, double* stats_ref, double* moments, double* superset) {
  const int nb = d->n_bins, J = d->n_conds, C = d->cells;
  uint64_t n[ORC_MAX_BINS * ORC_MAX_CONDS] = {0}, Su[ORC_MAX_BINS * ORC_MAX_CONDS] = {0},
           Sl[ORC_MAX_BINS * ORC_MAX_CONDS] = {0};
  u128 Suu[ORC_MAX_BINS * ORC_MAX_CONDS] = {0}, Sll[ORC_MAX_BINS * ORC_MAX_CONDS] = {0},
       Sul[ORC_MAX_BINS * ORC_MAX_CONDS] = {0};
  for (int c = 0; c < C; ++c) {
    int idx = bin[c] + nb * (c % J);
    uint64_t a = u[c], b = l[c];
    n[idx]++; Su[idx] += a; Sl[idx] += b;
    Suu[idx] += (u128)a * a; Sll[idx] += (u128)b * b; Sul[idx] += (u128)a * b;
  }
  /* per (bin, cond) moments and the mean / Fano superset */
  for (int j = 0; j < J; ++j)
    for (int b = 0; b < nb; ++b) {
      int idx = b + nb * j;
      double* m = moments ? moments + 6 * idx : NULL;
      double mt = 0.0, fft = 0.0;
      if (n[idx] == 0) {
        if (m) for (int q = 0; q < 6; ++q) m[q] = 0.0;
      } else {
        double nn = (double)n[idx];
        if (m) {
          m[0] = nn;
          m[1] = (double)Su[idx] / nn;
          m[2] = (double)Sl[idx] / nn;
          m[3] = var_of(n[idx], Su[idx], Suu[idx]);
          m[4] = cov_of(n[idx], Su[idx], Sl[idx], Sul[idx]);
          m[5] = var_of(n[idx], Sl[idx], Sll[idx]);
        }
        uint64_t St = Su[idx] + Sl[idx];
        u128 Stt = Suu[idx] + 2 * Sul[idx] + Sll[idx];
        mt = (double)St / nn;
        double vt = var_of(n[idx], St, Stt);
        fft = vt / (mt + (mt == 0.0 ? 0.0001 : 0.0));
      }
      if (superset) { superset[idx] = mt; superset[nb * J + idx] = fft; }
    }
  if (!stats_ref) return;
  /* pooled pulse / chase Fano per bin (data_summary_statistics.jl:2-32, 315-317) */
  for (int grp = 0; grp < 2; ++grp)
    for (int b = 0; b < nb; ++b) {
      uint64_t nn = 0, St = 0;
      u128 Stt = 0;
      for (int j = 0; j < J; ++j) {
        if (d->is_chase[j] != grp) continue;
        int idx = b + nb * j;
        nn += n[idx]; St += Su[idx] + Sl[idx];
        Stt += Suu[idx] + 2 * Sul[idx] + Sll[idx];
      }
      double mean = 0.0, ff = 0.0;
      if (nn > 0) {
        mean = (double)St / (double)nn;
        double v = var_of(nn, St, Stt);
        ff = v / (mean + (mean == 0.0 ? 0.0001 : 0.0));
      }
      stats_ref[(2 * grp) * nb + b] = mean;
      stats_ref[(2 * grp + 1) * nb + b] = ff;
    }
  /* ratios (data_summary_statistics.jl:61-77) and correlations (110-147) */
  for (int j = 0; j < J; ++j) {
    uint64_t nn = 0, su = 0, sl = 0;
    double m1[ORC_MAX_BINS] = {0}, m2[ORC_MAX_BINS] = {0}, v1[ORC_MAX_BINS] = {0},
           v2[ORC_MAX_BINS] = {0}, c12[ORC_MAX_BINS] = {0};
    int anynz = 0;
    for (int b = 0; b < nb; ++b) {
      int idx = b + nb * j;
      nn += n[idx]; su += Su[idx]; sl += Sl[idx];
      if (n[idx] == 0) {
        m1[b] = m2[b] = v1[b] = v2[b] = c12[b] = 0.0;
      } else {
        double q = (double)n[idx];
        m1[b] = (double)Su[idx] / q;
        m2[b] = (double)Sl[idx] / q;
        v1[b] = var_of(n[idx], Su[idx], Suu[idx]);
        v2[b] = var_of(n[idx], Sl[idx], Sll[idx]);
        c12[b] = cov_of(n[idx], Su[idx], Sl[idx], Sul[idx]);
      }
      if (c12[b] != 0.0) anynz = 1;
    }
    double ratio = 0.0;
    if (nn > 0) {
      double a1 = (double)su / (double)nn, a2 = (double)sl / (double)nn;
      if (a1 + a2 > 0.0) ratio = a2 / (a1 + a2);
    }
    const double* w = d->weights + nb * j;
    double s1 = w[0] * v1[0], s2 = w[0] * v2[0], s12 = w[0] * c12[0];
    for (int b = 1; b < nb; ++b) s1 = s1 + w[b] * v1[b];
    for (int b = 1; b < nb; ++b) s2 = s2 + w[b] * v2[b];
    for (int b = 1; b < nb; ++b) s12 = s12 + w[b] * c12[b];
    double tv1 = s1 + wcov(m1, m1, w, nb);
    double tv2 = s2 + wcov(m2, m2, w, nb);
    double mc = 0.0, cm = 0.0;
    if (anynz && tv1 != 0.0 && tv2 != 0.0) mc = s12 / sqrt(fabs(tv1 * tv2));
    if (tv1 != 0.0 && tv2 != 0.0) cm = wcov(m1, m2, w, nb) / sqrt(fabs(tv1 * tv2));
    stats_ref[4 * nb + j] = ratio;
    stats_ref[4 * nb + J + j] = mc;
    stats_ref[4 * nb + 2 * J + j] = cm;
  }
}

/* ----------------------------------------------------- batch entry point */
int orc_simulate(const orc_design* d, int model, uint64_t seed, int64_t first, int64_t n,
                 const double* theta, int n_threads, uint32_t* cell_u, uint32_t* cell_l,
                 uint8_t* cell_bin, double* stats_ref, double* moments, double* superset,
                 int64_t* steps, int64_t* reactions) {
  if (model < 1 || model > 5) return -1;
  if (d->n_bins > ORC_MAX_BINS || d->n_conds > ORC_MAX_CONDS || d->cells <= 0) return -1;
  if (first < 0 || first + n > 0xFFFFFFFFll || d->cells > 0x7FFFFFFF) return -1;
  const int C = d->cells, nb = d->n_bins, J = d->n_conds;
  const int K = 4 * nb + 3 * J, M = nb * J * 6, KS = 2 * nb * J;
  const int np = layout_of(model, nb).n_par;
  uint32_t key[3];
  key_of(seed, model, key);
  int err = 0;
  prates* R = (prates*)malloc(sizeof(prates) * (n > 0 ? n : 1));
  uint32_t* U = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * C);
  uint32_t* Lc = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * C);
  uint8_t* BN = (uint8_t*)malloc((size_t)(n > 0 ? n : 1) * C);
  int64_t* ST = (int64_t*)calloc((size_t)(n > 0 ? n : 1) * C, sizeof(int64_t));
  int64_t* RE = (int64_t*)calloc((size_t)(n > 0 ? n : 1) * C, sizeof(int64_t));
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  for (int64_t i = 0; i < n; ++i) {
    double th[2 * ORC_MAX_BINS + 8];
    for (int c = 0; c < np; ++c) th[c] = theta[i + n * c];
    make_rates(d, model, th, &R[i]);
  }
  /* every (particle, cell) is independent: schedule them across threads */
  const int64_t total = n * (int64_t)C;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int64_t q = 0; q < total; ++q) {
    const int64_t i = q / C;
    const int c = (int)(q - i * C);
    int e = sim_cell(d, &R[i], key, (uint32_t)(first + i), c, U + q, Lc + q, BN + q, ST + q, RE + q);
    if (e) {
#ifdef _OPENMP
#pragma omp critical
#endif
      err = e;
    }
  }
  if (!err) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i) {
      double sr[4 * ORC_MAX_BINS + 3 * ORC_MAX_CONDS];
      double mo[ORC_MAX_BINS * ORC_MAX_CONDS * 6];
      double ss[2 * ORC_MAX_BINS * ORC_MAX_CONDS];
      const size_t o = (size_t)i * C;
      orc_cell_stats(d, U + o, Lc + o, BN + o, sr, mo, ss);
      if (stats_ref) for (int k = 0; k < K; ++k) stats_ref[i + n * k] = sr[k];
      if (moments) for (int k = 0; k < M; ++k) moments[i + n * k] = mo[k];
      if (superset) for (int k = 0; k < KS; ++k) superset[i + n * k] = ss[k];
      if (cell_u) memcpy(cell_u + o, U + o, sizeof(uint32_t) * C);
      if (cell_l) memcpy(cell_l + o, Lc + o, sizeof(uint32_t) * C);
      if (cell_bin) memcpy(cell_bin + o, BN + o, C);
      int64_t st = 0, re = 0;
      for (int c = 0; c < C; ++c) { st += ST[o + c]; re += RE[o + c]; }
      if (steps) steps[i] = st;
      if (reactions) reactions[i] = re;
    }
  }
  free(R); free(U); free(Lc); free(BN); free(ST); free(RE);
  return err;
}
