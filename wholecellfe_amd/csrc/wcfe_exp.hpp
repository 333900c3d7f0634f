// exp(x) in FP64, one implementation for the GPU kinetics and the CPU.
//
// The bond kinetics decide every transition by u < 1 - exp(-k dt) with u a
// Philox uniform (UL_growth_explicit.cpp:2139-2148 for the mean-field rates;
// the Bell model of the batched replicas).  With the same exp on both sides,
// and every operation written out (no contraction), the GPU and a CPU
// restatement round each probability identically, so bond states compare bit
// for bit by construction rather than by margin.
//
// Method: argument reduction of Cody and Waite ("Software Manual for the
// Elementary Functions", 1980): x = n ln2 + r, n = rint(x log2(e)),
// |r| <= ln2/2 (up to the rounding of n).  ln2 is split into a head with 32
// significant bits, so n * head is exact for |n| < 2^21 and x - n * head is
// exact, and a tail: r = (x - n head) - n tail with one rounding.  exp(r) is
// the Taylor polynomial of degree 13 (truncation error below 5e-18 for
// |r| <= 0.35), evaluated by Horner's rule in fused multiply-adds
// (correctly rounded on the CPU and the GPU alike), then scaled by 2^n through
// the exponent field (two steps where 2^n is not a normal number).  The
// result is within 1 ulp of exp(x) (tested against the C library's exp).
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define WCFE_HD __host__ __device__
#else
#define WCFE_HD
#endif

namespace wcfe {

// 2^n for -1022 <= n <= 1023, from the exponent field
WCFE_HD inline double exp2_int(int n) {
  const uint64_t bits = (uint64_t)(n + 1023) << 52;
  double d;
  memcpy(&d, &bits, sizeof d);
  return d;
}

WCFE_HD inline double shared_exp(double x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double LOG2E = 0x1.71547652b82fep+0;
  const double LN2_HI = 0x1.62e42feep-1;       // ln 2, 32 significant bits
  const double LN2_LO = 0x1.a39ef35793c76p-33; // ln 2 - LN2_HI
  // 1/k!, k = 2..13, rounded to nearest
  const double C2 = 0x1.0p-1, C3 = 0x1.5555555555555p-3, C4 = 0x1.5555555555555p-5, C5 = 0x1.1111111111111p-7,
               C6 = 0x1.6c16c16c16c17p-10, C7 = 0x1.a01a01a01a01ap-13, C8 = 0x1.a01a01a01a01ap-16,
               C9 = 0x1.71de3a556c734p-19, C10 = 0x1.27e4fb7789f5cp-22, C11 = 0x1.ae64567f544e4p-26,
               C12 = 0x1.1eed8eff8d898p-29, C13 = 0x1.6124613a86d09p-33;
  if (x != x) return x + x;            // NaN
  if (x > 710.0) return x * 0x1.0p+1023;  // overflow to +inf
  if (x < -746.0) return 0.0;          // below half the smallest subnormal
  const double n = rint(x * LOG2E);
  double r = fma(-n, LN2_HI, x);
  r = fma(-n, LN2_LO, r);
  double p = C13;
  p = fma(p, r, C12);
  p = fma(p, r, C11);
  p = fma(p, r, C10);
  p = fma(p, r, C9);
  p = fma(p, r, C8);
  p = fma(p, r, C7);
  p = fma(p, r, C6);
  p = fma(p, r, C5);
  p = fma(p, r, C4);
  p = fma(p, r, C3);
  p = fma(p, r, C2);
  p = fma(p, r, 1.0);
  const double s = r * p;   // exp(r) - 1
  const double y = 1.0 + s; // in [0.70, 1.42]
  const int k = (int)n;
  if (k > 1023) return (y * exp2_int(k - 1)) * 2.0;
  if (k < -1021) return (y * exp2_int(k + 1000)) * exp2_int(-1000);
  return y * exp2_int(k);
}

}  // namespace wcfe
