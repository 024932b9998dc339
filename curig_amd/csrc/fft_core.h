// Stockham radix-16/8/4/2 FFT stages over a (padded) LDS buffer.
//
// Written as host+device code: the GPU kernels run it with a workgroup's threads,
// and the CPU check (fft_core_check.cpp) runs the same stages with a loop over
// thread ids, so the indexing is verified against numpy on the host.
//
// A buffer holds `count` independent sequences of length n (power of 2); element
// k of sequence s lives at pad(s * seq_stride + k * elem_stride).  Each stage is
// split in two phases separated by a barrier: every thread loads its butterflies'
// inputs into registers and computes (phase 1), then stores (phase 2), so one
// buffer suffices.  Each thread owns E elements per stage (E / R butterflies).
#pragma once

#ifdef __HIPCC__
#define FC_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define FC_HD inline
#endif

namespace curig {
namespace fftc {

template <typename T> struct cpx {
  T x, y;
};

template <typename T> FC_HD cpx<T> mk(T x, T y) {
  cpx<T> r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T> FC_HD cpx<T> add(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> FC_HD cpx<T> sub(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> FC_HD cpx<T> mul(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by (c + i s sign)
template <typename T> FC_HD cpx<T> mulw(cpx<T> a, T c, T s) { return mk(a.x * c - a.y * s, a.x * s + a.y * c); }

// LDS padding: one slot every 16 elements breaks power-of-two strides
FC_HD constexpr int pad(int i) { return i + (i >> 4); }
FC_HD constexpr int padded_size(int n) { return n + (n >> 4) + 1; }

// Padded positions of a butterfly's R elements e0 + r * st as p0 + r * pst when
// that holds exactly: st a multiple of 16 (each step adds st / 16 pad slots), or
// all R elements inside one 16-element block (no pad slot between them).  With
// compile-time sizes both tests fold, and a butterfly needs one address.
FC_HD bool run_offsets(int e0, int st, int R, int &p0, int &pst) {
  p0 = pad(e0);
  if ((st & 15) == 0) {
    pst = st + (st >> 4);
    return true;
  }
  if ((e0 & 15) + (R - 1) * st < 16) {
    pst = st;
    return true;
  }
  pst = 0;
  return false;
}



// cos / sin of 2 pi t / 16 as constant expressions (always folded at compile time)
FC_HD constexpr double cos16(int t) {
  return t == 0 ? 1.0 : t == 1 ? 0.92387953251128674 : t == 2 ? 0.70710678118654752 : t == 3 ? 0.38268343236508977
       : t == 4 ? 0.0 : t == 5 ? -0.38268343236508977 : t == 6 ? -0.70710678118654752 : t == 7 ? -0.92387953251128674
       : t == 8 ? -1.0 : t == 9 ? -0.92387953251128674 : t == 10 ? -0.70710678118654752
       : t == 11 ? -0.38268343236508977 : t == 12 ? 0.0 : t == 13 ? 0.38268343236508977
       : t == 14 ? 0.70710678118654752 : 0.92387953251128674;
}
FC_HD constexpr double sin16(int t) { return cos16((t + 12) & 15); }

FC_HD int bitrev(int x, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
  return r;
}

// a e^{sign 2 pi i t / 16} for a constant t (after unrolling): the multiples of
// pi/4 are formed without general products (IEEE rules keep the compiler from
// dropping y * 0, so 1, +-i and (1 +- i)/sqrt2 are written out).
template <typename T> FC_HD cpx<T> rot16(cpx<T> a, int t, int sign) {
  const int q = ((sign > 0 ? t : 16 - t) & 15);
  const T h = (T)0.70710678118654752440;
  switch (q) {
    case 0: return a;
    case 4: return mk(-a.y, a.x);
    case 8: return mk(-a.x, -a.y);
    case 12: return mk(a.y, -a.x);
    case 2: return mk((a.x - a.y) * h, (a.x + a.y) * h);
    case 6: return mk(-(a.x + a.y) * h, (a.x - a.y) * h);
    case 10: return mk((a.y - a.x) * h, -(a.x + a.y) * h);
    case 14: return mk((a.x + a.y) * h, (a.y - a.x) * h);
    default: return mulw(a, (T)cos16(q), (T)sin16(q));
  }
}

// In-register DFT of size R (2, 4, 8, 16): X[q] = sum_r v[r] e^{sign 2 pi i q r / R}
template <typename T, int R> FC_HD void dft_small(cpx<T> *v, int sign) {
  constexpr int L = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  cpx<T> a[R];
#pragma unroll
  for (int i = 0; i < R; ++i) a[bitrev(i, L)] = v[i];
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int st = 0; st < R; st += len) {
#pragma unroll
      for (int j = 0; j < len / 2; ++j) {
        const int t = j * (16 / len);
        cpx<T> u = a[st + j], w = rot16(a[st + j + len / 2], t, sign);
        a[st + j] = add(u, w);
        a[st + j + len / 2] = sub(u, w);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = a[i];
}

struct Layout {
  int n;            // sequence length (power of 2)
  int count;        // independent sequences
  int seq_stride;   // element offset between sequences
  int elem_stride;  // element offset between consecutive elements of a sequence
  bool seq_fast;    // consecutive butterflies walk sequences first (for elem_stride > 1)
};

// Phase 1 of a Stockham stage with sub-transform size Ns and radix R.
// tw: table of n_tw entries, tw[t] = e^{+2 pi i t / n_tw}, n_tw a multiple of n.
// Twiddles w^r, r = 1..R-1, of w = tw[k twstep]: w, w^2, w^4, w^8 are read from
// the table and the others formed by at most three products.
template <typename T, int R, typename Tw>
FC_HD void twiddle_powers(cpx<T> *w, const Tw tw, int base) {
  w[1] = tw[base];
  if (R > 2) w[2] = tw[2 * base];
  if (R > 4) w[4] = tw[4 * base];
  if (R > 8) w[8] = tw[8 * base];
  if (R > 2) w[3] = mul(w[1], w[2]);
  if (R > 4) {
    w[5] = mul(w[4], w[1]);
    w[6] = mul(w[4], w[2]);
    w[7] = mul(w[4], w[3]);
  }
  if (R > 8) {
#pragma unroll
    for (int r = 9; r < 16; ++r) w[r] = mul(w[8], w[r - 8]);
  }
}

template <typename T, int E, int R, typename Buf, typename Tw>
FC_HD void stage_load(const Buf buf, cpx<T> *v, const Layout &L, int Ns, int sign, const Tw tw, int n_tw,
                      int tid, int nthreads) {
  const int nbf = L.n / R;  // butterflies per sequence
  const int twstep = n_tw / (Ns * R);
#pragma unroll
  for (int b = 0; b < E / R; ++b) {
    const int beta = tid + b * nthreads;
    const int seq = L.seq_fast ? beta % L.count : beta / nbf;
    const int j = L.seq_fast ? beta / L.count : beta % nbf;
    const int base = seq * L.seq_stride;
    cpx<T> *u = v + b * R;
    int p0, pst;
    if (run_offsets(base + j * L.elem_stride, nbf * L.elem_stride, R, p0, pst)) {
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = buf[p0 + r * pst];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = buf[pad(base + (j + r * nbf) * L.elem_stride)];
    }
    const int k = j & (Ns - 1);
    if (Ns > 1) {
      cpx<T> w[R];
      twiddle_powers<T, R>(w, tw, k * twstep);
#pragma unroll
      for (int r = 1; r < R; ++r) u[r] = mulw(u[r], w[r].x, (T)sign * w[r].y);
    }
    dft_small<T, R>(u, sign);
    // keep one butterfly's addresses and twiddles live at a time
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
  }
}

// Phase 2: store the butterfly outputs at their Stockham positions.
template <typename T, int E, int R, typename Buf>
FC_HD void stage_store(Buf buf, const cpx<T> *v, const Layout &L, int Ns, int tid, int nthreads) {
  const int nbf = L.n / R;
#pragma unroll
  for (int b = 0; b < E / R; ++b) {
    const int beta = tid + b * nthreads;
    const int seq = L.seq_fast ? beta % L.count : beta / nbf;
    const int j = L.seq_fast ? beta / L.count : beta % nbf;
    const int base = seq * L.seq_stride;
    const int k = j & (Ns - 1);
    const int d = (j - k) * R + k;
    const cpx<T> *u = v + b * R;
    int p0, pst;
    if (run_offsets(base + d * L.elem_stride, Ns * L.elem_stride, R, p0, pst)) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[p0 + r * pst] = u[r];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad(base + (d + r * Ns) * L.elem_stride)] = u[r];
    }
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
  }
}

// Twiddle table e^{+2 pi i t / n} (0 <= t < n) as the product of two short
// tables, hi[t >> lb] = e^{2 pi i (t >> lb) 2^lb / n} and lo[t mod 2^lb] =
// e^{2 pi i (t mod 2^lb) / n}, small enough for LDS (2 sqrt(n) entries); each
// twiddle is one complex product of two correctly rounded values.
template <typename T, typename Ptr> struct TwoLevel {
  Ptr hi, lo;
  int lb;
  FC_HD cpx<T> operator[](int t) const { return mul(hi[t >> lb], lo[t & ((1 << lb) - 1)]); }
};
// lb for a length-n table: ceil(log2(n) / 2); hi has n >> lb entries, lo 2^lb
FC_HD int twiddle_split(int n) {
  int m = 0;
  while ((1 << m) < n) ++m;
  return (m + 1) / 2;
}
FC_HD int twiddle_entries(int n) { return (n >> twiddle_split(n)) + (1 << twiddle_split(n)); }

// Radices used for a length-n transform: maxr, maxr, ..., then the remainder.
FC_HD int stage_radix(int n, int Ns, int maxr = 16) { return (Ns * maxr <= n) ? maxr : n / Ns; }

}  // namespace fftc
}  // namespace curig
