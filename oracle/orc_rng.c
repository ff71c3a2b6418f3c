/*
 * orc_rng.c -- oracle RNG and elementary functions (TEST INFRASTRUCTURE ONLY).
 *
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11.  Pinned by the Random123 known-answer vectors in
 * tests/test_oracle_rng.py.
 *
 * log/exp: Cody & Waite, "Software Manual for the Elementary Functions"
 * (1980): argument reduction by powers of two (log) and by multiples of ln 2
 * with a two-word constant (exp), followed by a truncated series
 * (log: 2*atanh(s) series in s=(m-1)/(m+1); exp: Taylor series to degree 13).
 * Written with plain IEEE +,-,*,/ and compiled with -ffp-contract=off, so any
 * other implementation of the same steps gets the same bits.  Pinned against
 * libm/mpmath to a few ulp in tests.
 *
 * The reference draws its randomness from Julia's unseeded global RNG
 * (scripts/abc_simulation.jl:4-9); there is no bit-level counterpart.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

#define ABC_LOGTAB_QUAL static const
#include "abc_logtab.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += PHILOX_W0; k1 += PHILOX_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* (w + 1/2) * 2^-32: exactly representable, strictly inside (0,1). */
double orc_u01_32(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

/* 53-bit uniform on [0,1) from two words (27 + 26 bits). */
double orc_u01_53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1p-53;
}

static const double LN2_HI = 0x1.62e42fee00000p-1;
static const double LN2_LO = 0x1.a39ef35793c76p-33;
static const double INV_LN2 = 0x1.71547652b82fep+0;
static const double LN10 = 0x1.26bb1bbb55516p+1;

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double bitsd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* Natural log, table-driven (Tang, ACM TOMS 16(4), 1990): x = 2^k z with z in
 * [0.70703125, 1.4140625); the bits of x minus ABC_LOG_OFF select subinterval
 * j (bits 45..51), r = z * INV[j] - 1 by one fused multiply-add (|r| <= 2^-7),
 * ln x = k ln2 + L[j] + log1p(r) with L[j] = -ln(INV[j]) as hi + lo and
 * log1p(r) = r + r^2 (-1/2 + r/3 - r^2/4 + ... - r^6/8).  The table is the
 * generated abc_logtab.h of the engine (pure data; its entries are pinned by
 * the accuracy tests, which cover every subinterval). */
double orc_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int adj = 0;
  if (x < 0x1p-1022) { x = x * 0x1p54; adj = -54; }   /* subnormal */
  const uint64_t ix = dbits(x);
  const uint64_t tmp = ix - ABC_LOG_OFF;
  const int j = (int)((tmp >> 45) & (ABC_LOG_N - 1));
  const int k = (int)((int64_t)tmp >> 52) + adj;
  const double z = bitsd(ix - (tmp & (0xfffull << 52)));
  const double r = fma(z, abc_log_inv[j], -1.0);
  const double r2 = r * r;
  double p = -0x1p-3;                         /* -1/8 */
  p = fma(p, r, 0x1.2492492492492p-3);        /*  1/7 */
  p = fma(p, r, -0x1.5555555555555p-3);       /* -1/6 */
  p = fma(p, r, 0x1.999999999999ap-3);        /*  1/5 */
  p = fma(p, r, -0x1p-2);                     /* -1/4 */
  p = fma(p, r, 0x1.5555555555555p-2);        /*  1/3 */
  p = fma(p, r, -0x1p-1);                     /* -1/2 */
  const double lp = fma(p, r2, r);
  const double kd = (double)k;
  const double hi = kd * LN2_HI + abc_log_hi[j];
  return hi + ((kd * LN2_LO + abc_log_lo[j]) + lp);
}

static inline uint32_t fbits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float bitsf(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* Standard exponential variate from one 32-bit word in single precision, the
 * waiting times of the event loops: E = -ln(u), u = ((float)w + 0.5f) 2^-32.
 * The same table scheme on the bits of a float: the bits of u minus
 * ABC_LOG32_OFF select subinterval j (bits 16..22), r = z * INV32[j] - 1 by
 * one fused multiply-add (|r| <= 2^-8), log1p(r) = r + r^2 (r/3 - 1/2),
 * ln u = k ln2 + L32[j] + log1p(r) with ln2 as two floats.  Every step is one
 * IEEE single-precision operation or fused multiply-add. */
float orc_exp1_32(uint32_t w) {
  const float u = ((float)w + 0.5f) * 0x1p-32f;
  const uint32_t iu = fbits(u);
  const uint32_t tmp = iu - ABC_LOG32_OFF;
  const int j = (int)((tmp >> 16) & (uint32_t)(ABC_LOG_N - 1));
  const int k = (int32_t)tmp >> 23;
  const float z = bitsf(iu - (tmp & 0xFF800000u));
  const float r = fmaf(z, abc_log32_inv[j], -1.0f);
  const float p = fmaf(r, 0x1.555556p-2f, -0.5f);   /* 1/3, -1/2 */
  const float lp = fmaf(p * r, r, r);
  const float kf = (float)k;
  return -(fmaf(kf, 0x1.62e430p-1f, abc_log32_l[j]) + fmaf(kf, -0x1.05c610p-29f, lp));
}

/* e^r of the reduction x = k ln2 + r (ln2 in two parts): the degree-13
   Taylor polynomial, evaluated by Horner's rule in FMAs */
static double exp_reduced(double x, int* k_out) {
  double kd = floor(x * INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 0x1.6124613a86d09p-33;          /* 1/13! */
  p = fma(p, r, 0x1.1eed8eff8d898p-29);         /* 1/12! */
  p = fma(p, r, 0x1.ae64567f544e4p-26);         /* 1/11! */
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);         /* 1/10! */
  p = fma(p, r, 0x1.71de3a556c734p-19);         /* 1/9!  */
  p = fma(p, r, 0x1.a01a01a01a01ap-16);         /* 1/8!  */
  p = fma(p, r, 0x1.a01a01a01a01ap-13);         /* 1/7!  */
  p = fma(p, r, 0x1.6c16c16c16c17p-10);         /* 1/6!  */
  p = fma(p, r, 0x1.1111111111111p-7);          /* 1/5!  */
  p = fma(p, r, 0x1.5555555555555p-5);          /* 1/4!  */
  p = fma(p, r, 0x1.5555555555555p-3);          /* 1/3!  */
  p = fma(p, r, 0.5);                           /* 1/2!  */
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  *k_out = k;
  return p;
}

double orc_exp(double x) {
  if (x != x) return x;
  if (x > 709.78) return INFINITY;
  if (x < -745.2) return 0.0;
  int k;
  double p = exp_reduced(x, &k);
  /* scale by 2^k in two steps: the first exact, the second rounding once
     where the result is subnormal */
  if (k < -1000) {
    double s1 = bitsd((uint64_t)(k + 1000 + 1023) << 52);
    return (p * s1) * 0x1p-1000;
  }
  if (k > 1000) {
    double s1 = bitsd((uint64_t)(k - 1000 + 1023) << 52);
    return (p * s1) * 0x1p1000;
  }
  return p * bitsd((uint64_t)(k + 1023) << 52);
}

/* The device code scales by 2^k with one ldexp, which is exact or rounds
   once into the subnormal range like the second step above.  Returns how
   many of n pseudo-random x in [lo, hi] (xorshift64 from seed) give
   different bits under the two scalings. */
int64_t orc_exp_scaling_mismatches(double lo, double hi, int64_t n, uint64_t seed) {
  uint64_t s = seed ? seed : 1u;
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const double x = lo + (hi - lo) * ((double)(s >> 11) * 0x1p-53);
    if (x < -745.2 || x > 709.78) continue;
    int k;
    const double p = exp_reduced(x, &k);
    const double a = orc_exp(x), b = ldexp(p, k);
    uint64_t ua, ub;
    memcpy(&ua, &a, 8);
    memcpy(&ub, &b, 8);
    bad += ua != ub;
  }
  return bad;
}

double orc_exp10(double x) { return orc_exp(x * LN10); }

/* Correctly rounded (nearest-even) conversion of hi*2^64+lo to double. */
double orc_u128_to_f64(uint64_t hi, uint64_t lo) {
  unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
  if (v == 0) return 0.0;
  int p = hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll(lo);
  if (p < 53) return (double)(uint64_t)v;          /* exact */
  int sh = p - 52;
  unsigned __int128 mant = v >> sh;
  unsigned __int128 rem = v & ((((unsigned __int128)1) << sh) - 1);
  unsigned __int128 half = ((unsigned __int128)1) << (sh - 1);
  if (rem > half || (rem == half && (mant & 1))) {
    mant += 1;
    if (mant >> 53) { mant >>= 1; sh += 1; }
  }
  return ldexp((double)(uint64_t)mant, sh);
}
