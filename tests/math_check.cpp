// Host check of mrm_math.h's mlog against the C library and of its FP32
// mlogf against the exact logarithm (the C library's FP64 log): the largest
// error in units of the last place over sweeps of arguments, and the special
// values.  Prints one line per sweep: name, count, max ulp error.
#define MRM_HD inline
#include "../muram_main_amd/csrc/mrm_math.h"
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>

// distance in units of the last place of the correctly rounded value ref
static double ulps(double got, double ref) {
  if (got == ref) return 0.0;
  if (!std::isfinite(got) || !std::isfinite(ref)) return 1e300;
  int e;
  std::frexp(ref, &e);
  return std::fabs(got - ref) / std::ldexp(1.0, e - 53);
}

// distance in FP32 units of the last place of the exact value ref
static double ulpsf(float got, double ref) {
  if (!std::isfinite(got) || !std::isfinite(ref)) return got == ref ? 0.0 : 1e300;
  int e;
  std::frexp(ref, &e);
  return std::fabs((double)got - ref) / std::ldexp(1.0, e - 24);
}

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  int bad = 0;
  auto sweep_log = [&](const char* name, double lo, double hi, int n, bool logscale) {
    double worst = 0.0, at = 0.0;
    for (int i = 0; i < n; i++) {
      double x = logscale ? std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * u(rng)) : lo + (hi - lo) * u(rng);
      double d = ulps(mrm_math::mlog(x), std::log(x));
      if (d > worst) { worst = d; at = x; }
    }
    std::printf("log %s %d %.3f %.17g\n", name, n, worst, at);
  };
  sweep_log("near_one", 0.5, 2.0, 2000000, false);
  sweep_log("density", 1e-22, 1e-3, 2000000, true);     // ln rho
  sweep_log("specific_energy", 1e10, 1e17, 2000000, true);   // ln eps
  sweep_log("all_normal", 1e-300, 1e300, 2000000, true);
  sweep_log("denormal", 4.9e-324, 2.2e-308, 200000, true);
  auto sweep_logf = [&](const char* name, double lo, double hi, int n) {
    double worst = 0.0, at = 0.0;
    for (int i = 0; i < n; i++) {
      float x = (float)std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * u(rng));
      double d = ulpsf(mrm_math::mlogf(x), std::log((double)x));
      if (d > worst) { worst = d; at = x; }
    }
    std::printf("logf %s %d %.3f %.9g\n", name, n, worst, at);
  };
  sweep_logf("near_one", 0.5, 2.0, 2000000);
  sweep_logf("density", 1e-22, 1e-3, 2000000);
  sweep_logf("specific_energy", 1e10, 1e17, 2000000);
  sweep_logf("all_normal", 1.2e-38, 3.4e38, 2000000);
  // special values
  bad += !(std::isinf(mrm_math::mlog(0.0)) && mrm_math::mlog(0.0) < 0);
  bad += !std::isnan(mrm_math::mlog(-1.0));
  bad += !std::isnan(mrm_math::mlog(NAN));
  bad += !(std::isinf(mrm_math::mlog(INFINITY)) && mrm_math::mlog(INFINITY) > 0);
  bad += !(mrm_math::mlog(1.0) == 0.0);
  bad += !(std::isinf(mrm_math::mlogf(0.0f)) && mrm_math::mlogf(0.0f) < 0);
  bad += !std::isnan(mrm_math::mlogf(-1.0f));
  bad += !std::isnan(mrm_math::mlogf(NAN));
  bad += !(std::isinf(mrm_math::mlogf(INFINITY)) && mrm_math::mlogf(INFINITY) > 0);
  bad += !(mrm_math::mlogf(1.0f) == 0.0f);
  std::printf("special %d\n", bad);
  return 0;
}
