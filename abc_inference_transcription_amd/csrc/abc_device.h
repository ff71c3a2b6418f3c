// abc_device.h -- device-side random numbers and elementary functions for
// the gfx950 kernels.
//
// Philox4x32-10 (Salmon et al., SC'11): one independent stream per
// (particle, cell), counter = (draw_lo, draw_hi, cell | model << 24,
// particle), key = (seed_lo, seed_hi), the same in every lane of a launch so
// that the round keys stay scalar.  The reference draws from Julia's
// unseeded global RNG (scripts/abc_simulation.jl:4-9); counter-based streams
// make every particle reproducible and independent of the batch split.
//
// log is table-driven (Tang, ACM TOMS 16(4), 1990): 128 subintervals of
// [0.707, 1.414), reduction r = z * (1/c) - 1 by one fused multiply-add and a
// degree-8 log1p polynomial (the single-precision waiting-time log exp1_32
// uses the same scheme with a degree-3 polynomial); exp uses Cody & Waite's two-word ln2 reduction
// and a degree-13 Taylor polynomial.  Only IEEE-rounded +,-,*,/ and fused
// multiply-adds written as fma() are used, and the library is compiled with
// -ffp-contract=off, so results are identical on every IEEE machine that
// performs the same steps; this is what makes GPU runs replayable on a CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ABC_DEV __device__ __forceinline__

#define ABC_LOGTAB_QUAL __device__ static const
#include "abc_logtab.h"

namespace abcd {

struct u4 { uint32_t x, y, z, w; };

ABC_DEV u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 product per multiplier (a single v_mad_u64_u32 each)
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    // three-input xor in one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
    const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
    const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  return u4{c0, c1, c2, c3};
}

ABC_DEV double u01_32(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

ABC_DEV double u01_53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1p-53;
}

ABC_DEV uint32_t lemire(uint32_t w, uint32_t n) {
  return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

// Table of the log in LDS: entry j = {1/c_j, -ln(1/c_j) hi, lo, unused}.
struct LogEntry { double inv, hi, lo, pad; };

// fills a workgroup's LDS copy of the table (call before __syncthreads())
ABC_DEV void load_log_table(LogEntry* t) {
  for (int q = threadIdx.x; q < ABC_LOG_N; q += blockDim.x)
    t[q] = LogEntry{abc_log_inv[q], abc_log_hi[q], abc_log_lo[q], 0.0};
}

// natural log for finite x > 0 (the callers never pass anything else)
ABC_DEV double ln(double x, const LogEntry* tab) {
  uint64_t ix = (uint64_t)__double_as_longlong(x);
  int adj = 0;
  if (x < 0x1p-1022) {
    ix = (uint64_t)__double_as_longlong(x * 0x1p54);
    adj = -54;
  }
  const uint64_t tmp = ix - ABC_LOG_OFF;
  const int j = (int)((tmp >> 45) & (uint64_t)(ABC_LOG_N - 1));
  const int k = (int)((int64_t)tmp >> 52) + adj;
  const double z = __longlong_as_double((int64_t)(ix - (tmp & (0xfffull << 52))));
  const LogEntry e = tab[j];
  const double r = fma(z, e.inv, -1.0);
  const double r2 = r * r;
  double p = -0x1p-3;
  p = fma(p, r, 0x1.2492492492492p-3);
  p = fma(p, r, -0x1.5555555555555p-3);
  p = fma(p, r, 0x1.999999999999ap-3);
  p = fma(p, r, -0x1p-2);
  p = fma(p, r, 0x1.5555555555555p-2);
  p = fma(p, r, -0x1p-1);
  const double lp = fma(p, r2, r);
  const double kd = (double)k;
  const double hi = kd * 0x1.62e42fee00000p-1 + e.hi;
  return hi + ((kd * 0x1.a39ef35793c76p-33 + e.lo) + lp);
}

// Table of the single-precision waiting-time log in LDS: entry j =
// {1/c_j, -ln(1/c_j)} (float).
struct LogEntry32 { float inv, l; };

ABC_DEV void load_log32_table(LogEntry32* t) {
  for (int q = threadIdx.x; q < ABC_LOG_N; q += blockDim.x)
    t[q] = LogEntry32{abc_log32_inv[q], abc_log32_l[q]};
}

// A standard exponential variate from one 32-bit word, in single precision:
// E = -ln(u), u = ((float)w + 0.5f) * 2^-32 (0 < u <= 1).  The same table
// scheme as ln on the bits of a float (128 subintervals of [0.707, 1.414)),
// r = z * (1/c) - 1 with |r| <= 2^-8 by one fused multiply-add, log1p(r) =
// r + r^2 (r/3 - 1/2) (truncation below r^4/4 <= 2^-34), and k ln2 with ln2
// as two floats.  These are the waiting times of the event loops, added to
// their FP64 clocks; about 18 instructions against 38 for the FP64 log.
ABC_DEV float exp1_32(uint32_t w, const LogEntry32* tab) {
  // one rounding of ((float)w + 1/2) 2^-32, the same bits as the sum then the
  // exact power-of-two scaling (checked for every 32-bit w)
  const float u = fmaf((float)w, 0x1p-32f, 0x1p-33f);
  const uint32_t iu = __float_as_uint(u);
  const uint32_t tmp = iu - ABC_LOG32_OFF;
  const int j = (int)((tmp >> 16) & (uint32_t)(ABC_LOG_N - 1));
  const int k = (int)tmp >> 23;
  const float z = __uint_as_float(iu - (tmp & 0xFF800000u));
  const LogEntry32 e = tab[j];
  const float r = fmaf(z, e.inv, -1.0f);
  const float p = fmaf(r, 0x1.555556p-2f, -0.5f);
  const float lp = fmaf(p * r, r, r);
  const float kf = (float)k;
  return -(fmaf(kf, 0x1.62e430p-1f, e.l) + fmaf(kf, -0x1.05c610p-29f, lp));
}

// v_fma_f64 with its own destination: the compiler's two-address form
// (v_fmac_f64) writes over the addend and so copies a constant addend first
ABC_DEV double fma3(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// e^x for x <= 709.78 (or NaN): Cody-Waite reduction x = k ln2 + r with
// ln2 split in two, the degree-13 Taylor polynomial of e^r, then the scaling
// by 2^k as one ldexp (exact, or one rounding into the subnormal range).
// F: the fused multiply-add of the Horner chain (fma, or fma3 where the
// constants may stay in registers)
template <double (*F)(double, double, double)>
ABC_DEV double ex_core_with(double x) {
  if (x < -745.2) return 0.0;
  const double kd = floor(x * 0x1.71547652b82fep+0 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
  double p = 0x1.6124613a86d09p-33;
  p = F(p, r, 0x1.1eed8eff8d898p-29);
  p = F(p, r, 0x1.ae64567f544e4p-26);
  p = F(p, r, 0x1.27e4fb7789f5cp-22);
  p = F(p, r, 0x1.71de3a556c734p-19);
  p = F(p, r, 0x1.a01a01a01a01ap-16);
  p = F(p, r, 0x1.a01a01a01a01ap-13);
  p = F(p, r, 0x1.6c16c16c16c17p-10);
  p = F(p, r, 0x1.1111111111111p-7);
  p = F(p, r, 0x1.5555555555555p-5);
  p = F(p, r, 0x1.5555555555555p-3);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, k);
}

ABC_DEV double fma2(double a, double b, double c) { return fma(a, b, c); }

// in loops that evaluate it every iteration: the polynomial's constants stay
// in registers and each step is one instruction
ABC_DEV double ex_core(double x) { return ex_core_with<fma3>(x); }

ABC_DEV double ex(double x) {
  if (x != x) return x;
  if (x > 709.78) return __longlong_as_double(0x7ff0000000000000ll);
  return ex_core_with<fma2>(x);
}

ABC_DEV double ex10(double x) { return ex(x * 0x1.26bb1bbb55516p+1); }

// correctly rounded (nearest-even) conversion of the 128-bit integer hi:lo
ABC_DEV double u128_to_f64(uint64_t hi, uint64_t lo) {
  if (hi == 0 && (lo >> 53) == 0) return (double)lo;  // exact
  const int p = hi ? 127 - __clzll(hi) : 63 - __clzll(lo);
  const int sh = p - 52;  // 1..75: bits shifted out below the 53-bit mantissa
  uint64_t mant, rem_hi, rem_lo;
  if (sh >= 64) {
    mant = hi >> (sh - 64);
    rem_hi = hi & ((1ull << (sh - 64)) - 1ull);
    rem_lo = lo;
  } else {
    mant = (lo >> sh) | (hi << (64 - sh));
    rem_hi = 0ull;
    rem_lo = lo & ((1ull << sh) - 1ull);
  }
  const int hb = sh - 1;  // weight of one half ulp
  const uint64_t half_hi = hb >= 64 ? (1ull << (hb - 64)) : 0ull;
  const uint64_t half_lo = hb >= 64 ? 0ull : (1ull << hb);
  const bool gt = rem_hi > half_hi || (rem_hi == half_hi && rem_lo > half_lo);
  const bool eq = rem_hi == half_hi && rem_lo == half_lo;
  int e = sh;
  if (gt || (eq && (mant & 1ull))) {
    mant += 1ull;
    if (mant >> 53) { mant >>= 1; e += 1; }
  }
  return ldexp((double)mant, e);
}

}  // namespace abcd
