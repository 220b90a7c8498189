// mrm_math.h -- natural logarithms for the hot loops of the step (FP64: the
// EOS lookup's ln rho and ln eps, the TVD point's ln rho and ln eps), written
// from the textbook reduction so that it costs about 40 instead of 98 VALU
// instructions of the general library routine on gfx950 (the exponential of
// the library is already short and stays).  Accuracy: within one ulp of the
// C library for every positive argument, denormals included; special values
// as the library's: log(0) = -inf, log(x < 0) = log(NaN) = NaN,
// log(inf) = inf.  tests/test_math.py checks it against the C library on the
// host over the ranges the step uses and beyond.
//
// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
// s = f / (2 + f), z = s^2; log(m) = 2 atanh(s) = f - s (f - R(z)) with
// R(z) = sum_{k=1..10} 2 z^k / (2k + 1) (|s| <= 0.1716, the omitted terms are
// below 4e-19 of the result); log(x) = e ln2_hi + (log(m) + e ln2_lo), ln2
// split so that e ln2_hi is exact (Cody and Waite's argument reduction).
//
// mlogf is the same reduction in FP32 for the FP32-mixed TVD limiter (ln rho
// and ln eps of every point): four terms of R (omitted terms below 2e-9 of
// the result) and ln2 split into 15 + 24 bits, so that for the large
// arguments of the step (|ln x| of 10 to 45) the result is within about half
// an ulp of the exact logarithm, where the library's FP32 log built on the
// hardware log2 estimate is off by up to several ulps, and systematically
// so between neighbouring points.
#pragma once
#include <math.h>

#ifndef MRM_HD
#define MRM_HD __host__ __device__ __forceinline__
#endif

namespace mrm_math {

constexpr double LN2_HI = 0x1.62e42ffp-1;          // ln 2 to 29 significant bits
constexpr double LN2_LO = -0x1.718432a1b0e26p-35;  // ln 2 - LN2_HI

MRM_HD double mlog(double x) {
  int e;
  double m = frexp(x, &e);   // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 21.0;
  R = fma(R, z, 2.0 / 19.0);
  R = fma(R, z, 2.0 / 17.0);
  R = fma(R, z, 2.0 / 15.0);
  R = fma(R, z, 2.0 / 13.0);
  R = fma(R, z, 2.0 / 11.0);
  R = fma(R, z, 2.0 / 9.0);
  R = fma(R, z, 2.0 / 7.0);
  R = fma(R, z, 2.0 / 5.0);
  R = fma(R, z, 2.0 / 3.0);
  R = R * z;
  const double lm = fma(-s, f - R, f);
  const double de = (double)e;
  double r = fma(de, LN2_HI, fma(de, LN2_LO, lm));
  // zero, negative, NaN and infinite arguments
  if (!(x > 0.0)) r = (x == 0.0) ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) r = x;
  return r;
}

constexpr float LN2F_HI = 0x1.62e4p-1f;          // ln 2 to 15 significant bits
constexpr float LN2F_LO = 0x1.7f7d1cp-20f;       // ln 2 - LN2F_HI to 24 bits

MRM_HD float mlogf(float x) {
  int e;
  float m = frexpf(x, &e);   // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678f) {
    m = m + m;
    e -= 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float R = 2.0f / 9.0f;
  R = fmaf(R, z, 2.0f / 7.0f);
  R = fmaf(R, z, 2.0f / 5.0f);
  R = fmaf(R, z, 2.0f / 3.0f);
  R = R * z;
  const float lm = fmaf(-s, f - R, f);
  const float de = (float)e;
  float r = fmaf(de, LN2F_HI, fmaf(de, LN2F_LO, lm));
  if (!(x > 0.0f)) r = (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) r = x;
  return r;
}

}  // namespace mrm_math
