// Sine and cosine for the w-phase screens of the column passes.
//
// Written as host+device code so that the CPU check (tests/test_trig.py) runs
// the same function against the C library.  Method: Cody-Waite reduction of x
// by pi/2, with pi/2 split in three parts whose first two have 28 significant
// bits (n * part is exact for |n| < 2^25, i.e. |x| < 5.2e7), then Taylor
// polynomials of sin and cos on [-pi/4, pi/4] (terms through r^17 and r^18;
// the first omitted term is below 1e-19).  No branches: the quadrant is applied
// by selects.
#pragma once

#ifdef __HIPCC__
#define TRIG_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define TRIG_HD inline
#endif

namespace curig {

constexpr double kTrigMaxArg = 5.0e7;

TRIG_HD void sincos_cw(double x, double *sp, double *cp) {
  const double n = rint(x * 0x1.45f306dc9c883p-1);  // x * 2 / pi
  double r = x - n * 0x1.921fb54p+0;
  r = r - n * 0x1.10b4612p-30;
  r = r - n * -0x1.676733ae8fe48p-60;
  const double r2 = r * r;
  // sin r = r - r^3/3! + ... + r^17/17!
  double ps = 1.0 / 355687428096000.0;
  ps = ps * r2 - 1.0 / 1307674368000.0;
  ps = ps * r2 + 1.0 / 6227020800.0;
  ps = ps * r2 - 1.0 / 39916800.0;
  ps = ps * r2 + 1.0 / 362880.0;
  ps = ps * r2 - 1.0 / 5040.0;
  ps = ps * r2 + 1.0 / 120.0;
  ps = ps * r2 - 1.0 / 6.0;
  const double s = r + r * r2 * ps;
  // cos r = 1 - r^2/2! + ... - r^18/18!
  double pc = -1.0 / 6402373705728000.0;
  pc = pc * r2 + 1.0 / 20922789888000.0;
  pc = pc * r2 - 1.0 / 87178291200.0;
  pc = pc * r2 + 1.0 / 479001600.0;
  pc = pc * r2 - 1.0 / 3628800.0;
  pc = pc * r2 + 1.0 / 40320.0;
  pc = pc * r2 - 1.0 / 720.0;
  pc = pc * r2 + 1.0 / 24.0;
  const double c = 1.0 - 0.5 * r2 + r2 * r2 * pc;
  const int q = (int)((long long)n & 3);
  const double s1 = (q & 1) ? c : s, c1 = (q & 1) ? s : c;
  *sp = (q & 2) ? -s1 : s1;
  *cp = ((q + 1) & 2) ? -c1 : c1;
}

// Same reduction in FP64, then sin and cos of the reduced argument rounded to
// FP32 (Taylor through r^9 and r^10, truncation below 2e-9): the phase of a
// w-plane in the FP32-mixed path, where the plane values are FP32 themselves.
// The argument x and its reduction stay FP64 (x reaches ~1e4 rad).
TRIG_HD void sincos_rf(double x, float *sp, float *cp) {
  const double n = rint(x * 0x1.45f306dc9c883p-1);
  double r = x - n * 0x1.921fb54p+0;
  r = r - n * 0x1.10b4612p-30;
  r = r - n * -0x1.676733ae8fe48p-60;
  const float f = (float)r, f2 = f * f;
  float ps = 1.0f / 362880.0f;
  ps = ps * f2 - 1.0f / 5040.0f;
  ps = ps * f2 + 1.0f / 120.0f;
  ps = ps * f2 - 1.0f / 6.0f;
  const float s = f + f * f2 * ps;
  float pc = -1.0f / 3628800.0f;
  pc = pc * f2 + 1.0f / 40320.0f;
  pc = pc * f2 - 1.0f / 720.0f;
  pc = pc * f2 + 1.0f / 24.0f;
  const float c = 1.0f - 0.5f * f2 + f2 * f2 * pc;
  const int q = (int)((long long)n & 3);
  const float s1 = (q & 1) ? c : s, c1 = (q & 1) ? s : c;
  *sp = (q & 2) ? -s1 : s1;
  *cp = ((q + 1) & 2) ? -c1 : c1;
}

// Phase e^{i x} at the precision of the plane values T.
template <typename T> TRIG_HD void plane_phase(double x, double *s, double *c) {
  if (sizeof(T) == 4) {
    float sf, cf;
    sincos_rf(x, &sf, &cf);
    *s = sf;
    *c = cf;
  } else {
    sincos_cw(x, s, c);
  }
}

}  // namespace curig
