// Host run of the LDS FFT stages (fft_core.h) with the workgroup's threads
// emulated by loops, for the CPU tests: same code, same indexing, plus bounds
// checks on every buffer access.  Built with g++ by tests/test_fft_core.py.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fft_core.h"

using namespace curig::fftc;

namespace {

// bounds-checked view used in place of the LDS / table pointers
template <typename T> struct Checked {
  cpx<T> *p;
  long n;
  const char *what;
  cpx<T> &operator[](long i) const {
    if (i < 0 || i >= n) {
      fprintf(stderr, "fft_core_check: %s index %ld outside [0, %ld)\n", what, i, n);
      abort();
    }
    return p[i];
  }
};

template <typename T, int E, typename Tw>
int run_stages(int n, int count, int seq_stride, int elem_stride, int seq_fast, int sign, double *data, int n_tw,
               int maxr, const Tw &tw);

template <typename T, int E>
int run_e(int n, int count, int seq_stride, int elem_stride, int seq_fast, int sign, double *data, int n_tw,
          int maxr, int two_level) {
  std::vector<cpx<T>> twv(n_tw);
  for (int t = 0; t < n_tw; ++t) twv[t] = mk((T)std::cos(2 * M_PI * t / n_tw), (T)std::sin(2 * M_PI * t / n_tw));
  if (!two_level) {
    Checked<T> tw{twv.data(), (long)twv.size(), "twiddle"};
    return run_stages<T, E>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, tw);
  }
  // two-level table as the kernels hold it in LDS: hi (n_tw >> lb) | lo (2^lb)
  const int lb = twiddle_split(n_tw), nhi = n_tw >> lb, nlo = 1 << lb;
  std::vector<cpx<T>> hv(nhi), lv(nlo);
  for (int j = 0; j < nhi; ++j)
    hv[j] = mk((T)std::cos(2 * M_PI * ((double)j * nlo) / n_tw), (T)std::sin(2 * M_PI * ((double)j * nlo) / n_tw));
  for (int j = 0; j < nlo; ++j) lv[j] = mk((T)std::cos(2 * M_PI * j / n_tw), (T)std::sin(2 * M_PI * j / n_tw));
  TwoLevel<T, Checked<T>> tw{Checked<T>{hv.data(), nhi, "twiddle hi"}, Checked<T>{lv.data(), nlo, "twiddle lo"}, lb};
  return run_stages<T, E>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, tw);
}

template <typename T, int E, typename Tw>
int run_stages(int n, int count, int seq_stride, int elem_stride, int seq_fast, int sign, double *data, int n_tw,
               int maxr, const Tw &tw) {
  Layout L{n, count, seq_stride, elem_stride, seq_fast != 0};
  const int total = n * count;
  if (total % E) return 1;
  const int nthreads = total / E;
  // extent of the (unpadded) buffer
  int extent = (count - 1) * seq_stride + (n - 1) * elem_stride + 1;
  std::vector<cpx<T>> bufv(padded_size(extent));
  Checked<T> buf{bufv.data(), (long)bufv.size(), "LDS"};
  for (int s = 0; s < count; ++s)
    for (int k = 0; k < n; ++k) {
      int i = s * seq_stride + k * elem_stride;
      buf[pad(i)] = mk((T)data[2 * (s * n + k)], (T)data[2 * (s * n + k) + 1]);
    }
  std::vector<cpx<T>> regs((size_t)nthreads * E);
  for (int Ns = 1; Ns < n;) {
    int R = stage_radix(n, Ns, maxr);
    for (int t = 0; t < nthreads; ++t) {
      cpx<T> *v = &regs[(size_t)t * E];
      switch (R) {
        case 16: stage_load<T, E, 16>(buf, v, L, Ns, sign, tw, n_tw, t, nthreads); break;
        case 8: stage_load<T, E, 8>(buf, v, L, Ns, sign, tw, n_tw, t, nthreads); break;
        case 4: stage_load<T, E, 4>(buf, v, L, Ns, sign, tw, n_tw, t, nthreads); break;
        case 2: stage_load<T, E, 2>(buf, v, L, Ns, sign, tw, n_tw, t, nthreads); break;
        default: return 2;
      }
    }
    for (int t = 0; t < nthreads; ++t) {
      cpx<T> *v = &regs[(size_t)t * E];
      switch (R) {
        case 16: stage_store<T, E, 16>(buf, v, L, Ns, t, nthreads); break;
        case 8: stage_store<T, E, 8>(buf, v, L, Ns, t, nthreads); break;
        case 4: stage_store<T, E, 4>(buf, v, L, Ns, t, nthreads); break;
        case 2: stage_store<T, E, 2>(buf, v, L, Ns, t, nthreads); break;
      }
    }
    Ns *= R;
  }
  for (int s = 0; s < count; ++s)
    for (int k = 0; k < n; ++k) {
      int i = s * seq_stride + k * elem_stride;
      data[2 * (s * n + k)] = buf[pad(i)].x;
      data[2 * (s * n + k) + 1] = buf[pad(i)].y;
    }
  return 0;
}

}  // namespace

extern "C" int fftc_check_run2(int n, int count, int seq_stride, int elem_stride, int seq_fast, int sign,
                               int dbl, double *data, int n_tw, int E, int maxr, int two_level) {
  if (n < 16 || (n & (n - 1)) || (maxr != 8 && maxr != 16)) return 3;
  if (E == 32)
    return dbl ? run_e<double, 32>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, two_level)
               : run_e<float, 32>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, two_level);
  return dbl ? run_e<double, 16>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, two_level)
             : run_e<float, 16>(n, count, seq_stride, elem_stride, seq_fast, sign, data, n_tw, maxr, two_level);
}

extern "C" int fftc_check_run(int n, int count, int seq_stride, int elem_stride, int seq_fast, int sign,
                              int dbl, double *data, int n_tw, int E, int maxr) {
  return fftc_check_run2(n, count, seq_stride, elem_stride, seq_fast, sign, dbl, data, n_tw, E, maxr, 0);
}
