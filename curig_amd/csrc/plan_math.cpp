// Host plan math: kernel width / beta, grid sizes, w-plane count, quadrature and
// the kernel's Fourier series.  Pure host code (no HIP calls).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "curig_internal.h"

namespace curig {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

// src/fourier/conv_interp_invoker.cu:46-116.  Returns 0, 1 (eps clamped: the
// C entry points then stop, as cugridder.cu:403-406 does) or 2 (sigma <= 1).
int setup_conv_opts(KernelOpts &o, double eps, double upsampfac) {
  if (!(upsampfac > 1.0)) {
    set_error("upsampling factor %.3g too small (must be > 1)", upsampfac);
    return kErrSigma;
  }
  int ier = kOk;
  if (!(eps >= kEpsilonMin)) {
    set_error("epsilon %.3g below %.3g", eps, kEpsilonMin);
    eps = kEpsilonMin;
    ier = kErrEpsTooSmall;
  }
  int kw = (int)std::ceil(-std::log10(eps / 10.0));
  if (upsampfac != 2.0)
    kw = (int)std::ceil(-std::log(eps) / (M_PI * std::sqrt(1.0 - 1.0 / upsampfac)));
  if (kw < 2) kw = 2;
  if (kw > kMaxKernelWidth) kw = kMaxKernelWidth;
  double betaoverns = 2.30;
  if (kw == 2) betaoverns = 2.20;
  if (kw == 3) betaoverns = 2.26;
  if (kw == 4) betaoverns = 2.38;
  if (upsampfac != 2.0) betaoverns = 0.97 * M_PI * (1 - 1 / (2 * upsampfac));
  o.kw = kw;
  o.halfwidth = kw / 2.0;
  o.c = 4.0 / (double)(kw * kw);
  o.beta = betaoverns * kw;
  o.upsampfac = upsampfac;
  return ier;
}

// src/utils.cu:251-276
int next235beven(int n, int b) {
  if (n <= 2) return 2;
  if (n % 2 == 1) n += 1;
  int nplus = n - 2, numdiv = 2;
  while (numdiv > 1 || nplus % b != 0) {
    nplus += 2;
    numdiv = nplus;
    while (numdiv % 2 == 0) numdiv /= 2;
    while (numdiv % 3 == 0) numdiv /= 3;
    while (numdiv % 5 == 0) numdiv /= 5;
  }
  return nplus;
}

// src/fourier/conv_interp_invoker.cu:29-44
int get_num_cells(int ms, const KernelOpts &o) {
  int nf = (int)(o.upsampfac * ms);
  if (nf < 2 * o.kw) nf = 2 * o.kw;
  return next235beven(nf, 1);
}

// contrib/common.cpp:161-196 (S: half-width of n-1, X: half-width of w, radians)
void set_nhg_type3(double S, double X, const KernelOpts &o, int &nf, double &h, double &gam) {
  int nss = o.kw + 1;
  double Xsafe = X, Ssafe = S;
  if (X == 0.0) {
    if (S == 0.0) {
      Xsafe = 1.0;
      Ssafe = 1.0;
    } else {
      Xsafe = std::max(Xsafe, 1 / S);
    }
  } else {
    Ssafe = std::max(Ssafe, 1 / X);
  }
  double nfd = 2.0 * o.upsampfac * Ssafe * Xsafe / M_PI + nss;
  if (!std::isfinite(nfd)) nfd = 0.0;
  if (nfd > 2.0e9) nfd = 2.0e9;  // rejected by the caller (too many w planes)
  nf = (int)nfd;
  if (nf >= 1000000000) {
    nf = 1000000000;
    h = 2 * M_PI / nf;
    gam = (double)nf / (2.0 * o.upsampfac * Ssafe);
    return;
  }
  if (nf < 2 * o.kw) nf = 2 * o.kw;
  nf = next235beven(nf, 1);
  h = 2 * M_PI / nf;
  gam = (double)nf / (2.0 * o.upsampfac * Ssafe);
}

// n-point Gauss-Legendre rule on [-1,1], nodes ascending.  The reference uses
// the Glaser-Liu-Rokhlin solver (contrib/legendre_rule_fast.c:15-92) for the
// same unique rule; here: Newton on the three-term recurrence.
void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Scaled nodes and weighted kernel values of the 2p-node rule
// (src/fourier/deconv.cu:165-181: alpha = kw/2, p = (int)(2 + 3 alpha)).
// phi is the normalised kernel exp(beta (sqrt(1 - c x^2) - 1)).
int kernel_quadrature(const KernelOpts &o, double *x, double *g) {
  double alpha = o.kw / 2.0;
  int p = (int)(2 + 3.0 * alpha);
  int n = 2 * p;
  double t[kMaxQuadNodes], wq[kMaxQuadNodes];
  gauss_legendre(n, t, wq);
  for (int i = 0; i < n; ++i) {
    double xi = t[i] * alpha, phi = 0.0;
    if (std::fabs(xi) <= o.halfwidth) {
      double s = 1.0 - o.c * xi * xi;
      phi = std::exp(o.beta * (std::sqrt(s > 0 ? s : 0.0) - 1.0));
    }
    x[i] = xi;
    g[i] = alpha * wq[i] * phi;
  }
  return n;
}

// src/fourier/deconv.cu:13-24, 60-102: phihat(k) (-1)^k, k = 0..nf/2
void fseries_grid(const KernelOpts &o, int nf, double *out) {
  double x[kMaxQuadNodes], g[kMaxQuadNodes];
  int n = kernel_quadrature(o, x, g);
  int half = nf / 2;
  for (int k = 0; k <= half; ++k) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += g[i] * std::cos((x[i] - half) / half * M_PI * k);
    out[k] = s;
  }
}

// Piecewise-polynomial form of the normalised kernel (the reference's
// gpu_kerevalmeth = 1, declared in include/curafft_opts.h:23 but not
// implemented there).  On each unit interval i, x in [-kw/2 + i, -kw/2 + i + 1),
// phi is the polynomial sum_k a_k T_k(z), z = 2 t - 1 with t = x + kw/2 - i,
// interpolating phi at deg + 1 Chebyshev nodes (coefficients by the discrete
// cosine sum, so the fit is well conditioned), then rewritten in powers of z
// for Horner's rule.  deg is the smallest degree from kw - 1 up to kMaxHornerDeg
// whose largest error on 256 points per interval is below tol (phi peaks at 1).
// Fit tolerance: a quarter of the requested accuracy, not below what the
// evaluation precision resolves.  (The kernel's square-root edge at |x| = kw/2
// limits any polynomial fit of the outer intervals to ~1.3 e^-beta, about a
// tenth of epsilon, whatever the degree.)
double horner_tol(double eps, bool dbl) { return std::max(0.25 * eps, dbl ? 1e-15 : 3e-8); }

int horner_fit(const KernelOpts &o, double tol, std::vector<double> &coef) {
  const int kw = o.kw;
  auto phi = [&](double x) {
    const double s = 1.0 - o.c * x * x;
    return (std::fabs(x) < o.halfwidth && s > 0) ? std::exp(-o.beta * o.c * x * x / (std::sqrt(s) + 1.0)) : 0.0;
  };
  for (int deg = std::max(1, kw - 1); deg <= kMaxHornerDeg; ++deg) {
    const int n = deg + 1;
    coef.assign((size_t)kw * n, 0.0);
    double err = 0.0;
    for (int i = 0; i < kw; ++i) {
      long double a[kMaxHornerDeg + 1] = {0};
      for (int k = 0; k < n; ++k) {
        long double sum = 0;
        for (int j = 0; j < n; ++j) {
          const long double th = M_PI * (j + 0.5L) / n;
          const double z = (double)std::cos(th);
          sum += phi(-0.5 * kw + i + 0.5 * (z + 1.0)) * std::cos(k * th);
        }
        a[k] = sum * (k == 0 ? 1.0L : 2.0L) / n;
      }
      // sum a_k T_k(z) in powers of z: T_{k+1} = 2 z T_k - T_{k-1}
      long double tm[kMaxHornerDeg + 2] = {0}, t0[kMaxHornerDeg + 2] = {0}, p[kMaxHornerDeg + 2] = {0};
      t0[0] = 1;  // T_0
      tm[1] = 1;  // T_1 (held in tm, T_0 in t0)
      p[0] += a[0];
      if (n > 1) p[1] += a[1];
      for (int k = 2; k < n; ++k) {
        long double tn[kMaxHornerDeg + 2] = {0};
        for (int q = 0; q <= k; ++q) tn[q] = (q > 0 ? 2 * tm[q - 1] : 0) - t0[q];
        for (int q = 0; q <= k; ++q) p[q] += a[k] * tn[q];
        for (int q = 0; q <= k; ++q) {
          t0[q] = tm[q];
          tm[q] = tn[q];
        }
      }
      for (int q = 0; q < n; ++q) coef[(size_t)i * n + q] = (double)p[deg - q];  // highest power first
      for (int s2 = 0; s2 <= 256; ++s2) {
        const double z = -1.0 + 2.0 * s2 / 256.0;
        double r = coef[(size_t)i * n];
        for (int q = 1; q < n; ++q) r = r * z + coef[(size_t)i * n + q];
        err = std::max(err, std::fabs(r - phi(-0.5 * kw + i + 0.5 * (z + 1.0))));
      }
    }
    if (err < tol) return deg;
  }
  coef.clear();
  return -1;
}

}  // namespace curig
