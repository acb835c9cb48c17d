// fw_math.h -- FP64 sine and cosine for gfx950 device code: the phase factors
// of the propagators (fw_common.h) and the caller's expression strings, whose
// sin and cos the generated sources map to fw_user_sin / fw_user_cos. Embedded
// in libfastwigner.so and compiled by hipRTC; included before the caller's
// expressions.
#pragma once

// FP64 sine and cosine of the phase factors, branch-free so that the factors
// of a thread interleave. Cody-Waite reduction by pi/2 with a two-term split
// (C1 = pi/2 rounded to double, C2 = pi/2 - C1 rounded): for |x| >= 1 every bit
// of x - k C1 is at least 2^-52 and the difference is below 1, so the first
// FMA is exact; the second rounds once; the neglected k (pi/2 - C1 - C2) is
// below 1.5e-33 |k|, i.e. under 1e-17 for |x| < 2^42. The reduced
// |r| <= pi/4 + 2^-50 goes through the Taylor series of sin to r^17 and of cos
// to r^18 (truncation below 1e-19 at pi/4), then the quadrant rotation.
// Accuracy: about 1 ulp for |x| < 2^42 (phases in radians); beyond that the
// neglected term grows as |x| 1e-33.
__device__ __forceinline__ void sincos_fp64(double x, double& s, double& c) {
  const double k = rint(x * 0.63661977236758134308);
  const double C1 = 1.5707963267948966192e+00;
  const double C2 = 6.1232339957367658e-17;
  double r = fma(-k, C1, x);
  r = fma(-k, C2, r);
  const double r2 = r * r;
  double ps = -1.0 / 355687428096000.0;                 // -1/17!
  ps = fma(ps, r2, 1.0 / 1307674368000.0);              // 1/15!
  ps = fma(ps, r2, -1.0 / 6227020800.0);                // -1/13!
  ps = fma(ps, r2, 1.0 / 39916800.0);                   // 1/11!
  ps = fma(ps, r2, -1.0 / 362880.0);                    // -1/9!
  ps = fma(ps, r2, 1.0 / 5040.0);                       // 1/7!
  ps = fma(ps, r2, -1.0 / 120.0);                       // -1/5!
  ps = fma(ps, r2, 1.0 / 6.0);                          // 1/3!
  const double sn = fma(-r * r2, ps, r);                // r - r^3/3! + ...
  double pc = 1.0 / 6402373705728000.0;                 // 1/18!
  pc = fma(pc, r2, -1.0 / 20922789888000.0);            // -1/16!
  pc = fma(pc, r2, 1.0 / 87178291200.0);                // 1/14!
  pc = fma(pc, r2, -1.0 / 479001600.0);                 // -1/12!
  pc = fma(pc, r2, 1.0 / 3628800.0);                    // 1/10!
  pc = fma(pc, r2, -1.0 / 40320.0);                     // -1/8!
  pc = fma(pc, r2, 1.0 / 720.0);                        // 1/6!
  pc = fma(pc, r2, -1.0 / 24.0);                        // -1/4!
  pc = fma(pc, r2, 0.5);                                // 1/2!
  const double cs = fma(-r2, pc, 1.0);                  // 1 - r^2/2! + ...
  const int q = (int)((long long)k & 3);
  c = (q == 0) ? cs : (q == 1) ? -sn : (q == 2) ? -cs : sn;
  s = (q == 0) ? sn : (q == 1) ? cs : (q == 2) ? -sn : -cs;
}

// sin and cos for the caller's expression strings and functions. Branch-free
// and inlined, so the evaluations of a time-only subexpression such as
// sin(omega * t) in the many factor evaluations of one pass are straight-line
// code the compiler computes once (the library's sin and cos carry a
// large-argument branch per call, which keeps every copy alive). Accuracy as
// sincos_fp64: about 1 ulp for |x| < 2^42.
__device__ __forceinline__ double fw_user_sin(double x) {
  double s, c;
  sincos_fp64(x, s, c);
  return s;
}
__device__ __forceinline__ double fw_user_cos(double x) {
  double s, c;
  sincos_fp64(x, s, c);
  return c;
}
__device__ __forceinline__ float fw_user_sin(float x) { return sinf(x); }
__device__ __forceinline__ float fw_user_cos(float x) { return cosf(x); }
__device__ __forceinline__ double fw_user_sin(int x) { return fw_user_sin((double)x); }
__device__ __forceinline__ double fw_user_cos(int x) { return fw_user_cos((double)x); }
__device__ __forceinline__ double fw_user_sin(long x) { return fw_user_sin((double)x); }
__device__ __forceinline__ double fw_user_cos(long x) { return fw_user_cos((double)x); }
