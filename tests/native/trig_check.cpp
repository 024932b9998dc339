// Host build of trig.h for tests/test_trig.py.
#include "trig.h"

extern "C" void trig_check_run(long n, const double *x, double *s, double *c) {
  for (long i = 0; i < n; ++i) curig::sincos_cw(x[i], &s[i], &c[i]);
}

extern "C" void trig_check_rf(long n, const double *x, float *s, float *c) {
  for (long i = 0; i < n; ++i) curig::sincos_rf(x[i], &s[i], &c[i]);
}
