// Pruned LDS FFT path (power-of-two grids): the row and column kernels of
// both directions, included by one source per precision and kernel group
// (kern_lds_{spread,interp,cols}_{f32,f64}.hip), each of which instantiates
// its launchers.
#pragma once
#include "dev_common.h"
#include "trig.h"

namespace curig {

// --------------------------------------------------------------------------
// pruned LDS FFT path (power-of-two grids)
//
// ms2dirty per batch of planes: (1) one workgroup per grid row of a plane
// spreads the row's visibilities into LDS, runs the row FFT (length nf1) and
// writes the N1 image columns to the column-tiled buffer H; (2) one workgroup
// per image column runs the column FFTs (length nf2) of all the batch's planes
// and accumulates the N2 image rows with their w phases in FP64.  dirty2ms runs
// the mirror image: (1) per image column, w-fill + inverse column FFT of every
// plane -> H; (2) per grid row, inverse row FFT (the row expanded to nf1 cells)
// and interpolation.  Each plane crosses HBM twice (written once, read once) at
// half the grid size.
// --------------------------------------------------------------------------

// Record ranges of one row workgroup (plane p, grid rows [row0, row0 + nrows)):
// the bins of every layer holding a first plane in [p - kw + 1, p] and of every
// strip overlapping rows [row0 - kw + 1, row0 + nrows - 1] (both cyclic), with
// exclusive prefix counts, so that the workgroup walks them as one flat index
// range and issues all its record loads at once.
struct WalkTable {
  int nb, total;
  int start[kMaxWalk], pre[kMaxWalk];
};

// Called by the 64 lanes of wave 0; the table is read after a barrier.
__device__ __forceinline__ void walk_table_build(WalkTable &wt, const int *__restrict__ off, const Geom &g, int p,
                                                 int row0, int nrows) {
  const int lane = threadIdx.x;
  const int al = imod(p - g.kww + 1, g.P), as = imod(row0 - g.kw + 1, g.nf2);
  const int nl = tiles_count(al, g.kww, g.P, g.Pbin, g.nlayer);
  const int ns = tiles_count(as, nrows + g.kw - 1, g.nf2, g.T2, g.nty);
  int nb = nl * ns;
  if (nb > kMaxWalk) nb = kMaxWalk;  // the plan keeps nb within kMaxWalk
  const int l0 = al / g.Pbin, s0 = as / g.T2;
  int carry = 0;
  for (int base = 0; base < nb; base += 64) {
    const int k = base + lane;
    int st = 0, len = 0;
    if (k < nb) {
      const int j = k / ns, i = k - j * ns;
      int layer = l0 + j, strip = s0 + i;
      if (layer >= g.nlayer) layer -= g.nlayer;
      if (strip >= g.nty) strip -= g.nty;
      const int b = layer * g.ntiles + strip;
      st = off[b];
      len = off[b + 1] - st;
    }
    int x = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (k < nb) {
      wt.start[k] = st;
      wt.pre[k] = carry + x - len;
    }
    carry += __shfl(x, 63, 64);
  }
  if (lane == 0) {
    wt.nb = nb;
    wt.total = carry;
  }
}

// Record index of flat position t (0 <= t < total): the last range starting at or before t.
__device__ __forceinline__ int walk_record(const WalkTable &wt, int t) {
  int lo = 0, hi = wt.nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wt.pre[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return wt.start[lo] + (t - wt.pre[lo]);
}

// Bin offsets of a walk table (rows [row0, row0 + nrows)) loaded ahead by the 64 lanes of wave 0
// (bins k = lane and lane + 64), completed into a WalkTable later
// (walk_finish), so the load's latency hides behind other work.
struct WalkAhead {
  int st[2], en[2];
  int nb;
};

__device__ __forceinline__ void walk_issue(WalkAhead &wa, const int *__restrict__ off, const Geom &g, int p,
                                           int row0, int nrows = 1) {
  const int lane = threadIdx.x & 63;
  const int al = imod(p - g.kww + 1, g.P), as = imod(row0 - g.kw + 1, g.nf2);
  const int nl = tiles_count(al, g.kww, g.P, g.Pbin, g.nlayer);
  const int ns = tiles_count(as, nrows + g.kw - 1, g.nf2, g.T2, g.nty);
  int nb = nl * ns;
  if (nb > kMaxWalk) nb = kMaxWalk;  // the plan keeps nb within kMaxWalk
  wa.nb = nb;
  const int l0 = al / g.Pbin, s0 = as / g.T2;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = r * 64 + lane;
    wa.st[r] = wa.en[r] = 0;
    if (k < nb) {
      const int j = k / ns, i = k - j * ns;
      int layer = l0 + j, strip = s0 + i;
      if (layer >= g.nlayer) layer -= g.nlayer;
      if (strip >= g.nty) strip -= g.nty;
      const int b = layer * g.ntiles + strip;
      wa.st[r] = off[b];
      wa.en[r] = off[b + 1];
    }
  }
}

__device__ __forceinline__ void walk_finish(WalkTable &wt, const WalkAhead &wa) {
  const int lane = threadIdx.x & 63;
  int carry = 0;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = r * 64 + lane;
    const int len = k < wa.nb ? wa.en[r] - wa.st[r] : 0;
    int x = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (k < wa.nb) {
      wt.start[k] = wa.st[r];
      wt.pre[k] = carry + x - len;
    }
    carry += __shfl(x, 63, 64);
  }
  if (lane == 0) {
    wt.nb = wa.nb;
    wt.total = carry;
  }
}

// Copies a two-level twiddle table (hi | lo, fftc::twiddle_entries(n) entries)
// from global memory to LDS at dst; the caller's next barrier publishes it.
template <typename T>
__device__ __forceinline__ fftc::TwoLevel<T, const cpx<T> *> lds_twiddles(cpx<T> *dst,
                                                                         const cpx<T> *__restrict__ src, int n) {
  const int lb = fftc::twiddle_split(n), nhi = n >> lb, cnt = nhi + (1 << lb);
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
  return fftc::TwoLevel<T, const cpx<T> *>{dst, dst + nhi, lb};
}

template <typename T, int E, int MAXR = 16, typename Tw>
__device__ __forceinline__ void lds_fft(cpx<T> *buf, const fftc::Layout &L, int sign, const Tw tw, int n_tw) {
  cpx<T> v[E];
  const int nt = blockDim.x;
  for (int Ns = 1; Ns < L.n;) {
    const int R = fftc::stage_radix(L.n, Ns, MAXR);
    // opaque copy of the thread id: keeps each stage's address arithmetic inside
    // the stage (hoisted out of the loop it is held for all stages and spills)
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    if (R == 16)
      fftc::stage_load<T, E, 16>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else if (R == 8)
      fftc::stage_load<T, E, 8>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else if (R == 4)
      fftc::stage_load<T, E, 4>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else
      fftc::stage_load<T, E, 2>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    __syncthreads();
    if (R == 16)
      fftc::stage_store<T, E, 16>(buf, v, L, Ns, tid, nt);
    else if (R == 8)
      fftc::stage_store<T, E, 8>(buf, v, L, Ns, tid, nt);
    else if (R == 4)
      fftc::stage_store<T, E, 4>(buf, v, L, Ns, tid, nt);
    else
      fftc::stage_store<T, E, 2>(buf, v, L, Ns, tid, nt);
    __syncthreads();
    Ns *= R;
  }
}

// Stockham stages of `count` contiguous length-N sequences (buffer index
// pad(s * N + k)) with N a compile-time constant: every stage's radix, stride
// and twiddle step fold to constants, and the element indexing to shifts.
// (the twiddle table has TWM * N entries)
template <typename T, int E, int MAXR, int N, int NS, int TWM = 1, typename Tw>
__device__ __forceinline__ void fixed_stages(cpx<T> *buf, int count, int sign, const Tw &tw) {
  if constexpr (NS < N) {
    constexpr int R = (NS * MAXR <= N) ? MAXR : N / NS;
    const fftc::Layout L{N, count, N, 1, false};
    const int nt = blockDim.x;
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    cpx<T> v[E];
    fftc::stage_load<T, E, R>(buf, v, L, NS, sign, tw, N * TWM, tid, nt);
    __syncthreads();
    fftc::stage_store<T, E, R>(buf, v, L, NS, tid, nt);
    __syncthreads();
    fixed_stages<T, E, MAXR, N, NS * R, TWM>(buf, count, sign, tw);
  }
}

// FFT of one sequence of length m = 2^lgm (m = nf / S of a row split S ways)
// with the twiddle table of the full length nf = S m.
template <typename T, int E, int MAXR, int S, typename Tw>
__device__ __forceinline__ void lds_fft_part(cpx<T> *buf, int lgm, int sign, const Tw &tw) {
  switch (lgm) {
    case 10: fixed_stages<T, E, MAXR, 1024, 1, S>(buf, 1, sign, tw); break;
    case 11: fixed_stages<T, E, MAXR, 2048, 1, S>(buf, 1, sign, tw); break;
    case 12: fixed_stages<T, E, MAXR, 4096, 1, S>(buf, 1, sign, tw); break;
    default: fixed_stages<T, E, MAXR, 8192, 1, S>(buf, 1, sign, tw); break;
  }
}

// lds_fft_part with the first radix-16 stage on the thread's registers: v[e]
// holds x[tid + nt e] (m = 16 nt, the first Stockham stage's butterfly), so
// the inputs need not pass through LDS before it.
template <typename T, int E, int MAXR, int N, int S, typename Tw>
__device__ __forceinline__ void part_stages_from_regs(cpx<T> *buf, cpx<T> *v, int sign, const Tw &tw) {
  static_assert(E == 16 && MAXR == 16, "one radix-16 butterfly per thread");
  fftc::dft_small<T, 16>(v, sign);
  const fftc::Layout L{N, 1, N, 1, false};
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  fftc::stage_store<T, E, 16>(buf, v, L, 1, tid, blockDim.x);
  __syncthreads();
  fixed_stages<T, E, MAXR, N, 16, S>(buf, 1, sign, tw);
}
template <typename T, int E, int MAXR, int S, typename Tw>
__device__ __forceinline__ void lds_fft_part_from_regs(cpx<T> *buf, cpx<T> *v, int lgm, int sign, const Tw &tw) {
  switch (lgm) {
    case 11: part_stages_from_regs<T, E, MAXR, 2048, S>(buf, v, sign, tw); break;
    case 12: part_stages_from_regs<T, E, MAXR, 4096, S>(buf, v, sign, tw); break;
    default: part_stages_from_regs<T, E, MAXR, 8192, S>(buf, v, sign, tw); break;
  }
}

// A 4096-point part (256 threads, three radix-16 stages) whose last stage ends
// in registers: thread j returns X[j + 256 r] in v[r], having cleared the LDS
// elements it read last (every element of the part, for the next spread).
// Thread j's stage twiddles are the same for every plane: tw2 and tw3 hold
// w, w^2, w^4, w^8 of stages 2 (w = W256^(j mod 16)) and 3 (w = W4096^j),
// sign applied (part_twiddles); the other powers are products of these.
template <typename T> struct PartTw {
  cpx<T> tw2[4], tw3[4];
};

template <typename T, typename Tw>
__device__ __forceinline__ void part_twiddles(PartTw<T> &pt, const Tw &tw, int nf1, int sign) {
  const int j = threadIdx.x, n = nf1 - 1, st = nf1 / 4096;  // the table has nf1 = S 4096 entries
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const cpx<T> a = tw[((((j & 15) * 16) << i) * st) & n];
    const cpx<T> b = tw[((j << i) * st) & n];
    pt.tw2[i] = fftc::mk(a.x, (T)sign * a.y);
    pt.tw3[i] = fftc::mk(b.x, (T)sign * b.y);
  }
}

// u[r] *= w^r for r = 1..15 from p = (w, w^2, w^4, w^8); the powers pass
// through an empty asm so the eleven products stay at their use (hoisted out
// of a plane loop they would hold 22 more registers)
template <typename T> __device__ __forceinline__ void apply_tw16(cpx<T> *u, const cpx<T> *pin) {
  cpx<T> p[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    p[i] = pin[i];
    __asm__ volatile("" : "+v"(p[i].x), "+v"(p[i].y));
  }
#pragma unroll
  for (int r = 1; r < 16; ++r) {
    const int lo = __builtin_ctz(r);
    cpx<T> w = p[lo];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (((r >> i) & 1) && i != lo) w = fftc::mul(w, p[i]);
    u[r] = fftc::mul(u[r], w);
  }
}

// A 4096-point part (256 threads) from registers to LDS: v[r] holds
// x[j + 256 r] for thread j (the first stage's butterfly), the stage
// twiddles come from pt (part_twiddles), and X[k] ends at pad(k) in buf.
// The caller has made buf free (a barrier after its last reader); the
// function ends with a barrier.
template <typename T>
__device__ __forceinline__ void fft4096_part_regs_to_lds(cpx<T> *buf, cpx<T> *v, int sign, const PartTw<T> &pt) {
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  fftc::dft_small<T, 16>(v, sign);
  {
    cpx<T> *b = buf + 17 * tid;  // stage 1 outputs: pad(16 j + r) = 17 j + r
#pragma unroll
    for (int r = 0; r < 16; ++r) b[r] = v[r];
  }
  __syncthreads();
  int tid2 = threadIdx.x;
  __asm__ volatile("" : "+v"(tid2));
  {
    const cpx<T> *b = buf + fftc::pad(tid2);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = b[272 * r];
  }
  apply_tw16(v, pt.tw2);
  fftc::dft_small<T, 16>(v, sign);
  __syncthreads();
  {
    const int k = tid2 & 15;
    cpx<T> *b = buf + fftc::pad(16 * (tid2 - k) + k);  // stage 2 outputs: pad(d) + 17 r
#pragma unroll
    for (int r = 0; r < 16; ++r) b[17 * r] = v[r];
  }
  __syncthreads();
  int tid3 = threadIdx.x;
  __asm__ volatile("" : "+v"(tid3));
  cpx<T> *own = buf + fftc::pad(tid3);
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = own[272 * r];
  apply_tw16(v, pt.tw3);
  fftc::dft_small<T, 16>(v, sign);
  // X[j + 256 r] goes where this thread read its stage-3 inputs
#pragma unroll
  for (int r = 0; r < 16; ++r) own[272 * r] = v[r];
  __syncthreads();
}

// Cells per block of the block-floating-point spread (kAccFix below) of a
// 4096-cell row part, and the blocks.
constexpr int kFixBlkLog2 = 4, kFixBlks = 4096 >> kFixBlkLog2;
constexpr int kFixNaN = 0x40000000;  // exponent of a block with a non-finite bound

template <typename T, int IN = 0>
__device__ __forceinline__ void fft4096_part_to_regs(cpx<T> *buf, cpx<T> *v, int sign, const PartTw<T> &pt,
                                                     const short *fexp = nullptr) {
  // every stage's butterfly j reads x[j + 256 r] at pad(j) + 272 r
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  if (IN == 2) {
    // block floating point (kAccFix): int32 pairs at the padded positions,
    // cell j + 256 r in block (j + 256 r) >> kFixBlkLog2 of exponent fexp[block]
    const int2 *b = reinterpret_cast<const int2 *>(buf) + fftc::pad(tid);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int2 x = b[272 * r];
      const int e = fexp[(tid >> kFixBlkLog2) + r * (256 >> kFixBlkLog2)];
      v[r] = fftc::mk((T)ldexpf((float)x.x, -e), (T)ldexpf((float)x.y, -e));
    }
  } else {
    const cpx<T> *b = buf + fftc::pad(tid);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = b[272 * r];
  }
  fftc::dft_small<T, 16>(v, sign);
  __syncthreads();
  {
    cpx<T> *b = buf + 17 * tid;  // stage 1 outputs: pad(16 j + r) = 17 j + r
#pragma unroll
    for (int r = 0; r < 16; ++r) b[r] = v[r];
  }
  __syncthreads();
  int tid2 = threadIdx.x;
  __asm__ volatile("" : "+v"(tid2));
  {
    const cpx<T> *b = buf + fftc::pad(tid2);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = b[272 * r];
  }
  apply_tw16(v, pt.tw2);
  fftc::dft_small<T, 16>(v, sign);
  __syncthreads();
  {
    const int k = tid2 & 15;
    cpx<T> *b = buf + fftc::pad(16 * (tid2 - k) + k);  // stage 2 outputs: pad(d) + 17 r
#pragma unroll
    for (int r = 0; r < 16; ++r) b[17 * r] = v[r];
  }
  __syncthreads();
  int tid3 = threadIdx.x;
  __asm__ volatile("" : "+v"(tid3));
  cpx<T> *own = buf + fftc::pad(tid3);
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = own[272 * r];
#pragma unroll
  for (int r = 0; r < 16; ++r) own[272 * r] = fftc::mk((T)0, (T)0);
  apply_tw16(v, pt.tw3);
  fftc::dft_small<T, 16>(v, sign);
}

// FFT of `count` contiguous sequences of length n = 2^lg (6 <= lg <= 14) in LDS.
template <typename T, int E, int MAXR, typename Tw>
__device__ __forceinline__ void lds_fft_rows(cpx<T> *buf, int lg, int count, int sign, const Tw &tw) {
  switch (lg) {
    case 6: fixed_stages<T, E, MAXR, 64, 1>(buf, count, sign, tw); break;
    case 7: fixed_stages<T, E, MAXR, 128, 1>(buf, count, sign, tw); break;
    case 8: fixed_stages<T, E, MAXR, 256, 1>(buf, count, sign, tw); break;
    case 9: fixed_stages<T, E, MAXR, 512, 1>(buf, count, sign, tw); break;
    case 10: fixed_stages<T, E, MAXR, 1024, 1>(buf, count, sign, tw); break;
    case 11: fixed_stages<T, E, MAXR, 2048, 1>(buf, count, sign, tw); break;
    case 12: fixed_stages<T, E, MAXR, 4096, 1>(buf, count, sign, tw); break;
    case 13: fixed_stages<T, E, MAXR, 8192, 1>(buf, count, sign, tw); break;
    default: fixed_stages<T, E, MAXR, 16384, 1>(buf, count, sign, tw); break;
  }
}

__device__ __forceinline__ int image_index(int k, int N, int nf) {
  // grid index (FFT order) -> image index (centred), or -1 outside the image
  const int h = N / 2;
  return k < N - h ? k + h : (k >= nf - h ? k - nf + h : -1);
}
__device__ __forceinline__ int grid_index(int i, int N, int nf) {
  const int h = N / 2;
  return i >= h ? i - h : nf + i - h;
}

// A record's kernel taps, evaluated once when the rows were binned
// (bin_scatter_kernel): x along nf1 (KWM cells from ix), y along nf2, z along
// the w planes; zero beyond the kernel width.
template <typename T, int KWM> struct Taps {
  T x[KWM], y[KWM], z[KWM];
};

template <typename T, int KWM>
__device__ __forceinline__ void load_taps(const T *__restrict__ taps, int r, Taps<T, KWM> &t) {
  const T *p = taps + (size_t)r * (3 * KWM);
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    t.x[i] = p[i];
    t.y[i] = p[KWM + i];
    t.z[i] = p[2 * KWM + i];
  }
}

// a[q] for a runtime q < KWM, by selects.  Each value passes through an empty
// asm so the compiler cannot fold the select chain back into an indexed load
// from a stack copy of the array (which would go to scratch memory).
template <typename T, int KWM> __device__ __forceinline__ T sel_tap(const T *a, int q) {
  T r = a[0];
  __asm__ volatile("" : "+v"(r));
#pragma unroll
  for (int i = 1; i < KWM; ++i) {
    T v = a[i];
    __asm__ volatile("" : "+v"(v));
    r = (q == i) ? v : r;
  }
  return r;
}

// Spread of one record into the workgroup's grid rows [row0, row0 + rpw) of plane p (LDS).
template <typename T, int KWM>
__device__ __forceinline__ void spread_rows(const Rec<T> rc, const Taps<T, KWM> &tp, T *ldsf, const Geom &g, int p,
                                            int row0) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2, rpw = g.rpw;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int ry = rc.iy - row0;
  if (ry < 0) ry += nf2;
  if (!(ry < rpw || ry + kw - 1 >= nf2)) return;
  const T vw = sel_tap<T, KWM>(tp.z, k);
  const T vre = rc.re * vw, vim = rc.im * vw;
  for (int jj = 0; jj < kw; ++jj) {
    int cy = ry + jj;
    if (cy >= nf2) cy -= nf2;
    if (cy >= rpw) continue;
    const T ky = sel_tap<T, KWM>(tp.y, jj);
    const T are = vre * ky, aim = vim * ky;
    const int base = cy * nf1;
#pragma unroll
    for (int i = 0; i < KWM; ++i) {
      if (i < kw) {
        int cx = rc.ix + i;
        if (cx >= nf1) cx -= nf1;
        const int a = 2 * fftc::pad(base + cx);
        atomicAdd(ldsf + a, are * tp.x[i]);
        atomicAdd(ldsf + a + 1, aim * tp.x[i]);
      }
    }
  }
}

// Interpolation of one record from the workgroup's rows of plane p (LDS): its
// partial sum over those rows, added to the record's FP64 accumulator.
template <typename T, int KWM>
__device__ __forceinline__ void interp_rows(const Rec<T> rc, const Taps<T, KWM> &tp, int idx, const cpx<T> *lds,
                                            const Geom &g, int p, int row0, double2 *__restrict__ vacc) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2, rpw = g.rpw;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int ry = rc.iy - row0;
  if (ry < 0) ry += nf2;
  if (!(ry < rpw || ry + kw - 1 >= nf2)) return;
  T sr = 0, si = 0;
  for (int jj = 0; jj < kw; ++jj) {
    int cy = ry + jj;
    if (cy >= nf2) cy -= nf2;
    if (cy >= rpw) continue;
    const int base = cy * nf1;
    T rr = 0, ri = 0;
#pragma unroll
    for (int i = 0; i < KWM; ++i) {
      if (i < kw) {
        int cx = rc.ix + i;
        if (cx >= nf1) cx -= nf1;
        const cpx<T> f = lds[fftc::pad(base + cx)];
        rr += f.x * tp.x[i];
        ri += f.y * tp.x[i];
      }
    }
    const T ky = sel_tap<T, KWM>(tp.y, jj);
    sr += rr * ky;
    si += ri * ky;
  }
  const T kz = sel_tap<T, KWM>(tp.z, k);
  atomicAdd(&vacc[idx].x, (double)(sr * kz));
  atomicAdd(&vacc[idx].y, (double)(si * kz));
}


// The intermediate planes live in a column-tiled layout: element (plane k,
// grid row k2, column slot s) at (k * nct + s / kTile) * tstride + k2 * kTile
// + s % kTile with tstride = nf2 * kTile + a pad, so a row's kTile-column
// chunks are contiguous for the row kernels and a column's elements are kTile
// apart for the column kernels (the kTile columns of a tile run on one XCD,
// whose L2 serves the shared sectors).

// Offsets in H: the first element of plane k, and of column slot s in row 0
// of a plane (row k2 adds k2 * kTile).
__device__ __forceinline__ size_t h_plane(const Geom &g, int k) { return (size_t)k * g.nct * g.tstride; }
__device__ __forceinline__ int h_slot(const Geom &g, int s) {
  return (s >> kTileLog2) * g.tstride + (s & (kTile - 1));
}

// Column slots of the tiled layout.  With whole rows the slot of image column
// col is col.  With rows split S ways, the columns of part h (grid index h mod
// S, col = S i + c_h with c_h = (h + N1/2) mod S) take slots h M + i, M = nf1 /
// (2 S) >= ceil(N1 / S), so a part writes whole kTile-column chunks; slots whose
// i is past the part's last column are unused (-1 here).
__device__ __forceinline__ int slot_column(int s, const Geom &g) {
  if (g.rsplit <= 1) return s < g.N1 ? s : -1;
  if (s >= g.nslot) return -1;
  const int S = g.rsplit, h = s >> g.slot_lg, i = s & ((1 << g.slot_lg) - 1);
  const int col = S * i + ((h + g.N1 / 2) & (S - 1));
  return col < g.N1 ? col : -1;
}

// whether a column workgroup's slots [s0, s0 + cpw) hold any image column
__device__ __forceinline__ bool slots_used(int s0, const Geom &g, int cpw) {
  for (int c = 0; c < cpw; ++c)
    if (slot_column(s0 + c, g) >= 0) return true;
  return false;
}

// Workgroup id -> work item such that each XCD (workgroups are dispatched to
// the 8 XCDs round robin) takes one contiguous range of items.
__device__ __forceinline__ int xcd_group(int bid, int n) {
  const int q = n >> 3, r = n & 7, x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Workgroup id -> work item with blocks of blk consecutive items dealt to the
// 8 XCDs in turn (n a multiple of 8 blk): every XCD gets the same mix of the
// items, and a block's items share one L2.
__device__ __forceinline__ int xcd_interleave(int bid, int n, int blk) {
  if (n % (8 * blk)) return xcd_group(bid, n);
  const int x = bid & 7, k = bid >> 3;
  return ((k / blk) * 8 + x) * blk + k % blk;
}

template <typename T> struct MaxThreads { static constexpr int value = sizeof(T) == 8 ? 512 : 1024; };

// ms2dirty: spread into full grid rows held in LDS, then the row FFT (length
// nf1), keeping the N1 image columns, written to the column-tiled layout H.  One
// workgroup owns rpw rows of one w plane (rpw * nf1 = E * blockDim.x); rows are
// binned into full-width strips of T2 rows.
template <typename T, int E, int KWM>
__global__ void __launch_bounds__(MaxThreads<T>::value) spread_row_fft_kernel(const Rec<T> *__restrict__ recs,
                                                                             const T *__restrict__ taps,
                                                                             const int *__restrict__ off,
                                                                             cpx<T> *__restrict__ H,
                                                                             const cpx<T> *__restrict__ tw, Geom g,
                                                                             int p0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  T *ldsf = reinterpret_cast<T *>(smem_raw);
  const int rpw = g.rpw, nf1 = g.nf1, nt = blockDim.x;
  // rows in XCD-contiguous order: the rows sharing a tile's lines are written
  // through one L2, which merges their chunks into whole lines
  const int row0 = xcd_group(blockIdx.x, gridDim.x) * rpw, pl = blockIdx.y, p = p0 + pl;
  const int ntot = fftc::padded_size(rpw * nf1);
  const auto tv = lds_twiddles<T>(lds + ntot, tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, rpw);
  for (int i = threadIdx.x; i < ntot; i += nt) lds[i] = fftc::mk((T)0, (T)0);
  __syncthreads();
  const int total = wt.total;
  const int N1 = g.N1, nout = rpw * N1, h1 = N1 / 2;
  cpx<T> *hb = H + h_plane(g, pl) + row0 * kTile;
  if (total == 0) {
    // no visibility reaches these rows of this plane: the transform is zero
    for (int i = threadIdx.x; i < nout; i += nt) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      hb[h_slot(g, col) + (r << kTileLog2)] = fftc::mk((T)0, (T)0);
    }
    return;
  }
  for (int t0 = threadIdx.x; t0 < total; t0 += nt) {
    const int ia = walk_record(wt, t0);
    const Rec<T> ra = recs[ia];
    Taps<T, KWM> ta;
    load_taps<T, KWM>(taps, ia, ta);
    spread_rows<T, KWM>(ra, ta, ldsf, g, p, row0);
  }
  __syncthreads();
  lds_fft_rows<T, E, 16>(lds, 31 - __clz(nf1), rpw, +1, tv);
  // pruned write: all LDS reads first, then the stores (kTile-column chunks);
  // offsets from the row's base in 32 bits (a plane holds < 2^31 elements)
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  cpx<T> o[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      const int k1 = col >= h1 ? col - h1 : nf1 + col - h1;
      o[e] = lds[fftc::pad(r * nf1 + k1)];
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      hb[h_slot(g, col) + (r << kTileLog2)] = o[e];
    }
  }
}


// dirty2ms: row pass of the column-transformed planes (H as col_inv_kernel
// writes it; each row expanded to nf1 cells, zero outside the image columns),
// inverse row FFT in LDS, then interpolation of every record touching the workgroup's rows of
// plane p: per record a partial sum, added to its FP64 accumulator (atomics),
// so the full grid planes are never written.
template <typename T, int E, int KWM>
__global__ void __launch_bounds__(MaxThreads<T>::value) row_ifft_interp_kernel(const cpx<T> *__restrict__ H,
                                                                              const Rec<T> *__restrict__ recs,
                                                                              const T *__restrict__ taps,
                                                                              const int *__restrict__ off,
                                                                              double2 *__restrict__ vacc,
                                                                              const cpx<T> *__restrict__ tw,
                                                                              Geom g, int p0, int sign) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int rpw = g.rpw, nf1 = g.nf1, N1 = g.N1, nt = blockDim.x;
  // rows in XCD-contiguous order: the 16 rows sharing a column line run together
  const int row0 = xcd_group(blockIdx.x, gridDim.x) * rpw, pl = blockIdx.y, p = p0 + pl;
  const int lg = 31 - __clz(nf1);
  cpx<T> v[E];
  const cpx<T> *hb = H + h_plane(g, pl) + row0 * kTile;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = threadIdx.x + e * nt;  // < rpw * nf1 = E * nt
    const int r = i >> lg, k1 = i & (nf1 - 1);
    const int col = image_index(k1, N1, nf1);
    v[e] = col >= 0 ? hb[h_slot(g, col) + (r << kTileLog2)]
                    : fftc::mk((T)0, (T)0);
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(rpw * nf1), tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, rpw);
#pragma unroll
  for (int e = 0; e < E; ++e) lds[fftc::pad(threadIdx.x + e * nt)] = v[e];
  __syncthreads();
  const int total = wt.total;
  if (total == 0) return;  // no visibility reaches these rows of this plane
  // first record of this thread, loaded while the FFT runs
  int idx0 = 0;
  Rec<T> r0;
  Taps<T, KWM> q0;
  if (threadIdx.x < total) {
    idx0 = walk_record(wt, threadIdx.x);
    r0 = recs[idx0];
    load_taps<T, KWM>(taps, idx0, q0);
  }
  lds_fft_rows<T, E, 16>(lds, lg, rpw, sign, tv);
  if (threadIdx.x < total) interp_rows<T, KWM>(r0, q0, idx0, lds, g, p, row0, vacc);
  for (int t = threadIdx.x + nt; t < total; t += nt) {
    const int idx = walk_record(wt, t);
    Taps<T, KWM> q;
    load_taps<T, KWM>(taps, idx, q);
    interp_rows<T, KWM>(recs[idx], q, idx, lds, g, p, row0, vacc);
  }
}

// Spread of one record into part h of a row split S ways (ms2dirty): part h
// holds z_h[n mod m] = sum over the row's cells n of x[n] w^(sign h n), whose
// m-point FFT gives the row transform at the indices S k + h (radix-S
// decimation in frequency; w = e^{2 pi i / nf1}, m = nf1 / S).
// Accumulation of the spread in the LDS block: kAccNative, the padded buffer
// of the path's type with float atomics; kAccFix, block floating point in the
// same buffer, each component an int32 scaled by 2^e_b for the cells' block b
// of 2^kFixBlkLog2 (e_b from a bound on the block's contributions, so every
// partial sum stays below 2^30: bound_row_part).  LDS FP32 atomics take about
// 100 cycles per wave instruction on this GPU, integer ones about as little
// as stores.
enum { kAccNative = 0, kAccFix = 2 };

// Bound pass of kAccFix for one record (index idx): adds the record's bound
// (|re| + |im|) |y tap| |z tap| for this row and plane to bsum[b] of each
// block b its cells reach (at most two: kw <= 16).  Every component of the
// record's contribution to a cell is at most that bound times an x tap, at
// most 1 (the ES kernel's maximum, e^0 at the centre, which its polynomial
// fits to the requested accuracy), and twiddles have modulus 1; so bsum[b]
// bounds every partial sum of the block's cells.  The factor 2 between the
// scaled bound (2^30) and the int32 range covers x taps slightly above 1,
// the FP32 rounding of the sum and that of the contributions.
template <typename T, int KWM>
__device__ __forceinline__ void bound_row_part(const Rec<T> rc, int idx, const T *__restrict__ taps, const Geom &g,
                                               int p, int row0, int mm, float *bsum) {
  const int kw = g.kw, nf2 = g.nf2;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int jj = row0 - rc.iy;
  if (jj < 0) jj += nf2;
  if (jj >= kw) return;
  const T *tq = taps + (size_t)idx * (3 * KWM);
  const float b = (fabsf((float)rc.re) + fabsf((float)rc.im)) * fabsf((float)tq[KWM + jj]) *
                  fabsf((float)tq[2 * KWM + k]);
  const int b0 = (rc.ix & mm) >> kFixBlkLog2, b1 = ((rc.ix + kw - 1) & mm) >> kFixBlkLog2;
  atomicAdd(bsum + b0, b);
  if (b1 != b0) atomicAdd(bsum + b1, b);
}

// kAccFix: the exponent e of a block from its bound B (< 2^(ilogb(B) + 1)):
// e = 29 - ilogb(B) keeps B 2^e < 2^30.  Contributions are scaled by 2^e with
// ldexp, so data of any FP32 magnitude keep 30 bits.  kFixNaN marks a block
// with a non-finite bound: the row part's transform is then NaN throughout,
// as a non-finite input makes it.
__device__ __forceinline__ int fix_block_exp(float B) {
  if (!(B < __uint_as_float(0x7f800000u))) return kFixNaN;  // NaN or +inf
  if (B == 0.f) return 0;
  return 29 - ilogbf(B);
}

template <typename T, int KWM, bool PAD = true, typename Tw>
__device__ __forceinline__ void spread_row_part(const Rec<T> rc, const Taps<T, KWM> &tp, T *ldsf, const Tw &tv,
                                                const cpx<T> *wh, const Geom &g, int p, int row0, int h, int mm,
                                                int sign, int accm = kAccNative, const float *fixtab = nullptr) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int jj = row0 - rc.iy;
  if (jj < 0) jj += nf2;
  if (jj >= kw) return;
  const T a = sel_tap<T, KWM>(tp.z, k) * sel_tap<T, KWM>(tp.y, jj);
  T vre = rc.re * a, vim = rc.im * a;
  if (h) {
    // w^(sign h n) at the cells n = ix + i: w^(sign h ix) once, w^(sign h i) per tap (wh)
    const cpx<T> w = tv[(h * rc.ix) & (nf1 - 1)];
    const T ws = (T)sign * w.y, t = vre * w.x - vim * ws;
    vim = vre * ws + vim * w.x;
    vre = t;
  }
  // kAccFix: the exponents of the (at most two) blocks of the record's cells,
  // from the blocks' bounds
  int b0 = 0, e0 = 0, e1 = 0;
  if (accm == kAccFix) {
    b0 = (rc.ix & mm) >> kFixBlkLog2;
    const int b1 = ((rc.ix + kw - 1) & mm) >> kFixBlkLog2;
    e0 = fix_block_exp(fixtab[b0]);
    e1 = fix_block_exp(fixtab[b1]);
    e0 = e0 == kFixNaN ? 0 : e0;
    e1 = e1 == kFixNaN ? 0 : e1;
  }
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    if (i < kw) {
      int cx = rc.ix + i;
      if (cx >= nf1) cx -= nf1;
      T re = vre * tp.x[i], im = vim * tp.x[i];
      if (h) {
        const T t = re * wh[i].x - im * wh[i].y;
        im = re * wh[i].y + im * wh[i].x;
        re = t;
      }
      if (accm == kAccFix) {
        const int e = ((cx & mm) >> kFixBlkLog2) == b0 ? e0 : e1;
        int *d = reinterpret_cast<int *>(ldsf) + 2 * (PAD ? fftc::pad(cx & mm) : (cx & mm));
        atomicAdd(d, __float2int_rn(ldexpf((float)re, e)));
        atomicAdd(d + 1, __float2int_rn(ldexpf((float)im, e)));
      } else {
        const int q = 2 * (PAD ? fftc::pad(cx & mm) : (cx & mm));
        atomicAdd(ldsf + q, re);
        atomicAdd(ldsf + q + 1, im);
      }
    }
  }
}


// ms2dirty row pass for rows too long for two workgroups per CU: one workgroup
// per part h of a row split S ways (the row transform's outputs S k + h, see
// spread_row_part), each spreading the row's visibilities into an m-point LDS
// buffer, transforming it and writing its share of the N1 image columns (those
// whose grid index is h mod S).  Several such workgroups fit one CU, so one's
// memory phases overlap another's FFT.  The workgroup keeps its (row, part)
// and walks the batch's nb planes, as the column kernels do: while plane p is
// transformed, the walk table of plane p + 1 is complete and each thread's
// first record of it is in flight, and the twiddle table and the part's
// factors are loaded once (16384^2 grids: 23.8 against 24.3 ms per 32-plane launch for
// one workgroup per (row, part, plane)).  With S = 1 it runs whole FP32 rows
// of 4096 (2048^2 images).  FP32 parts of 4096 accumulate the spread in block
// floating point (kAccFix): int32 LDS atomics instead of float ones, after a
// bound pass over the plane's records.
template <typename T, int E, int KWM, int S>
__global__ void __launch_bounds__(256) spread_row_part_planes_kernel(
    const Rec<T> *__restrict__ recs, const T *__restrict__ taps, const int *__restrict__ off, cpx<T> *__restrict__ H,
    const cpx<T> *__restrict__ tw, Geom g, int p0, int nb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt[2];
  __shared__ cpx<T> whs[KWM];
  // kAccFix block table (FP32): the blocks' bounds, and their exponents for
  // the first FFT stage
  constexpr bool kFixable = sizeof(T) == 4;
  __shared__ float fixt[kFixable ? kFixBlks : 1];
  __shared__ short fixe[kFixable ? kFixBlks : 1];
  __shared__ int fixnan;                         // a block's bound is not finite
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  T *ldsf = reinterpret_cast<T *>(smem_raw);
  const int nf1 = g.nf1, m = nf1 / S, nt = blockDim.x, sign = +1, j = threadIdx.x;
  // the parts of four consecutive rows (whose chunks share lines of H) are
  // consecutive items of one XCD; the XCDs take such blocks in turn
  const int item = xcd_interleave(blockIdx.x, gridDim.x, 4 * S);
  const int row0 = item / S, h = item % S;
  // FP32 parts of 4096 accumulate in block floating point (kAccFix), not
  // with LDS float atomics
  const bool regs = E == 16 && m == 4096, fix = kFixable && regs;
  const int accm = fix ? kAccFix : kAccNative;
  if (kFixable && j < kFixBlks) fixt[j] = 0.f;
  if (kFixable && j == 0) fixnan = 0;
  const int ntot = fftc::padded_size(m);
  const auto tv = lds_twiddles<T>(lds + ntot, tw, nf1);
  if (j < 64) walk_table_build(wt[0], off, g, p0, row0, 1);
  for (int i = j; i < ntot; i += nt) lds[i] = fftc::mk((T)0, (T)0);
  if (j < KWM) {
    // w^(sign h i) for the taps i (part h's decimation factor), from the
    // global twiddle table (the LDS copy is being written)
    const int lb = fftc::twiddle_split(nf1);
    const fftc::TwoLevel<T, const cpx<T> *> tg{tw, tw + (nf1 >> lb), lb};
    const cpx<T> w = tg[h * j];
    whs[j] = fftc::mk(w.x, (T)sign * w.y);
  }
  __syncthreads();
  // this part's image columns: grid_index(col) = col - N1/2 (mod nf1) = h (mod S)
  const int N1 = g.N1, h1 = N1 / 2, c0 = (h + h1) % S, nout = (N1 - c0 + S - 1) / S;
  const int hs = h << g.slot_lg;  // this part's first column slot
  // the stage twiddles of this thread (FP32 parts of 4096: kept across planes)
  PartTw<T> pt;
  if (regs) part_twiddles(pt, tv, nf1, sign);
  // this thread's first record of plane p0 (its taps are loaded when spread)
  Rec<T> r0;
  int idx0 = 0;
  if (j < wt[0].total) {
    idx0 = walk_record(wt[0], j);
    r0 = recs[idx0];
  }
  int cur = 0;
  for (int pl = 0; pl < nb; ++pl, cur ^= 1) {
    const int p = p0 + pl, total = wt[cur].total;
    const bool more = pl + 1 < nb;
    WalkAhead wa;
    if (j < 64 && more) walk_issue(wa, off, g, p + 1, row0);
    if (fix && total > 0) {
      // block bounds (kAccFix); a plane without records leaves the table zero
      if (j < total) bound_row_part<T, KWM>(r0, idx0, taps, g, p, row0, m - 1, fixt);
      for (int t = j + nt; t < total; t += nt) {
        const int ia = walk_record(wt[cur], t);
        bound_row_part<T, KWM>(recs[ia], ia, taps, g, p, row0, m - 1, fixt);
      }
      __syncthreads();
      if (j < kFixBlks) {
        // published to the first FFT stage by the barrier after the spread
        const int e = fix_block_exp(fixt[j]);
        fixe[j] = (short)(e == kFixNaN ? 0 : e);
        if (e == kFixNaN) fixnan = 1;
      }
    }
    if (j < total) {
      Taps<T, KWM> q0;
      load_taps<T, KWM>(taps, idx0, q0);
      spread_row_part<T, KWM>(r0, q0, ldsf, tv, whs, g, p, row0, h, m - 1, sign, accm, fixt);
    }
    for (int t = j + nt; t < total; t += nt) {
      const int ia = walk_record(wt[cur], t);
      Taps<T, KWM> ta;
      load_taps<T, KWM>(taps, ia, ta);
      spread_row_part<T, KWM>(recs[ia], ta, ldsf, tv, whs, g, p, row0, h, m - 1, sign, accm, fixt);
    }
    if (j < 64 && more) walk_finish(wt[cur ^ 1], wa);
    __syncthreads();  // spread done; the next plane's walk table is complete
    // the next plane's first record, in flight during this plane's transform
    if (more && j < wt[cur ^ 1].total) {
      idx0 = walk_record(wt[cur ^ 1], j);
      r0 = recs[idx0];
    }
    cpx<T> *hb = H + h_plane(g, pl) + row0 * kTile;
    if (total == 0) {
      // no visibility reaches this part of this row and plane: its outputs are
      // zero (and the LDS block and the block table still are)
      for (int i = j; i < nout; i += nt) hb[h_slot(g, hs + i)] = fftc::mk((T)0, (T)0);
      continue;
    }
    if (regs) {
      // last stage in registers: output X[j + 256 r] (part index) is image
      // column S (j + 256 r) + h + N1 / 2 (folded by nf1), which can be an
      // image column for r < 4 and r >= 12 only; its slot steps by 64 tiles
      // per r (as in row_ifft_interp_part_kernel's inputs)
      cpx<T> v[16];
      const bool pnan = fix && fixnan;  // read before the transform's barriers; cleared after them
      if (fix)
        fft4096_part_to_regs<T, 2>(lds, v, sign, pt, fixe);
      else
        fft4096_part_to_regs<T, 0>(lds, v, sign, pt);
      if (pnan) {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = fftc::mk((T)NAN, (T)NAN);
      }
      const int tid = threadIdx.x, q = (h + h1) / S, k0 = S * tid + h + h1;
      cpx<T> *a0 = hb + h_slot(g, hs + tid + q);
      const long estep = (long)(256 >> kTileLog2) * g.tstride, mstep = (long)(4096 >> kTileLog2) * g.tstride;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if (r < 4) {
          if (k0 + S * 256 * r < N1) a0[r * estep] = v[r];
        } else if (r >= 12) {
          if (k0 + S * 256 * r >= nf1) a0[r * estep - mstep] = v[r];
        }
      }
      if (fix && j < kFixBlks) fixt[j] = 0.f;  // last read by the spread
      if (fix && j == 0) fixnan = 0;
      __syncthreads();  // the block and the table are clear before the next plane's spread
    } else {
      lds_fft_part<T, E, 16, S>(lds, 31 - __clz(m), sign, tv);
      int tid = threadIdx.x;
      __asm__ volatile("" : "+v"(tid));
      // nout <= m / 2 = (E / 2) nt, since N1 <= nf1 / 2
#pragma unroll
      for (int e = 0; e < E / 2; ++e) {
        const int i = tid + e * nt;
        if (i < nout) {
          const int col = S * i + c0;
          const int k1 = col >= h1 ? col - h1 : nf1 + col - h1;
          hb[h_slot(g, hs + i)] = lds[fftc::pad(k1 / S)];
        }
      }
      __syncthreads();  // every output read: clear the block for the next plane
      for (int i = tid; i < ntot; i += nt) lds[i] = fftc::mk((T)0, (T)0);
      __syncthreads();  // cleared before the next plane's spread
    }
  }
}

// Interpolation of one record from part h of a row split S ways: the row's
// cells are x[n] = sum_h w^(sign h n) Y_h[n mod m] (w = e^{2 pi i / nf1}, m = nf1 / S),
// and part h holds Y_h; its share of the record's sum is added to the record's
// FP64 accumulator.
template <typename T, int KWM, bool PAD = true, typename Tw>
__device__ __forceinline__ void interp_row_part(const Rec<T> rc, const Taps<T, KWM> &tp, int idx,
                                                const cpx<T> *lds, const Tw &tv, const cpx<T> *wh, const Geom &g,
                                                int p, int row0, int h, int mm, int sign,
                                                double2 *__restrict__ vacc) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int jj = row0 - rc.iy;
  if (jj < 0) jj += nf2;
  if (jj >= kw) return;
  T rr = 0, ri = 0;
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    if (i < kw) {
      int cx = rc.ix + i;
      if (cx >= nf1) cx -= nf1;
      cpx<T> f = lds[PAD ? fftc::pad(cx & mm) : (cx & mm)];
      if (h) f = fftc::mulw(f, wh[i].x, wh[i].y);  // w^(sign h i); w^(sign h ix) below
      rr += f.x * tp.x[i];
      ri += f.y * tp.x[i];
    }
  }
  if (h) {
    const cpx<T> w = tv[(h * rc.ix) & (nf1 - 1)];
    const T ws = (T)sign * w.y, t = rr * w.x - ri * ws;
    ri = rr * ws + ri * w.x;
    rr = t;
  }
  const T kyz = sel_tap<T, KWM>(tp.y, jj) * sel_tap<T, KWM>(tp.z, k);
  atomicAdd(&vacc[idx].x, (double)(rr * kyz));
  atomicAdd(&vacc[idx].y, (double)(ri * kyz));
}

// dirty2ms row pass for rows too long for two workgroups per CU: each row is
// split S ways by the index of its inverse-transform input, k = S k' + h
// (radix-S decimation in frequency), one workgroup per part: Y_h = FFT_m over
// X[S k' + h], so the parts of a row load disjoint columns of H, run m-point
// FFTs and add their shares to each record's accumulator.
template <typename T, int E, int KWM, int S>
__global__ void __launch_bounds__(256) row_ifft_interp_part_kernel(
    const cpx<T> *__restrict__ H, const Rec<T> *__restrict__ recs, const T *__restrict__ taps,
    const int *__restrict__ off,
    double2 *__restrict__ vacc, const cpx<T> *__restrict__ tw, Geom g, int p0, int sign) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int nf1 = g.nf1, m = nf1 / S, N1 = g.N1, nt = blockDim.x;
  const int item = xcd_group(blockIdx.x, gridDim.x);
  const int row0 = item / S, h = item % S, pl = blockIdx.y, p = p0 + pl;
  const int lgm = 31 - __clz(m);
  const cpx<T> *hb = H + h_plane(g, pl) + row0 * kTile;
  // Input e of this thread is X[S i + h], i = tid + e nt (< m = E nt).  With
  // N1 <= nf1 / 2 only the first and last quarter of i can be image columns
  // (grid indices within N1 / 2 of 0); their column slots are
  // h nf1 / (2 S) + i + q (first quarter) and the same minus m (last), with
  // q = (h + N1 / 2) / S, so their addresses step by (nt / kTile) tiles.
  const int tid = threadIdx.x, h1 = N1 / 2, q = (h + h1) / S;
  const int k0 = S * tid + h + h1;  // column of e = 0 before the fold
  const cpx<T> *a0 = hb + h_slot(g, (h << g.slot_lg) + tid + q);
  const long estep = (long)(nt >> kTileLog2) * g.tstride, mstep = (long)(m >> kTileLog2) * g.tstride;
  cpx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    v[e] = fftc::mk((T)0, (T)0);
    if (e < E / 4) {
      if (k0 + S * nt * e < N1) v[e] = a0[e * estep];
    } else if (e >= 3 * E / 4) {
      if (k0 + S * nt * e >= nf1) v[e] = a0[e * estep - mstep];
    }
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(m), tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, 1);
  // FP32 keeps the inputs in registers for the first stage; FP64 parks them
  // in LDS (held across the record loads they would cost 64 registers)
  constexpr bool regs = sizeof(T) == 4;
  if (!regs) {
#pragma unroll
    for (int e = 0; e < E; ++e) lds[fftc::pad(threadIdx.x + e * nt)] = v[e];
  }
  __syncthreads();
  const int total = wt.total;
  if (total == 0) return;
  int idx0 = 0;
  Rec<T> r0;
  Taps<T, KWM> q0;
  if (threadIdx.x < total) {
    idx0 = walk_record(wt, threadIdx.x);
    r0 = recs[idx0];
    load_taps<T, KWM>(taps, idx0, q0);
  }
  if constexpr (regs)
    lds_fft_part_from_regs<T, E, 16, S>(lds, v, lgm, sign, tv);
  else
    lds_fft_part<T, E, 16, S>(lds, lgm, sign, tv);
  cpx<T> wh[KWM];
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    const cpx<T> w = tv[h * i];
    wh[i] = fftc::mk(w.x, (T)sign * w.y);
  }
  if (threadIdx.x < total) interp_row_part<T, KWM>(r0, q0, idx0, lds, tv, wh, g, p, row0, h, m - 1, sign, vacc);
  for (int t = threadIdx.x + nt; t < total; t += nt) {
    const int idx = walk_record(wt, t);
    Taps<T, KWM> q;
    load_taps<T, KWM>(taps, idx, q);
    interp_row_part<T, KWM>(recs[idx], q, idx, lds, tv, wh, g, p, row0, h, m - 1, sign, vacc);
  }
}

// dirty2ms row pass for rows or row parts of 4096 (256 threads): as
// row_ifft_interp_part_kernel (S = 1: whole rows), but the workgroup keeps
// its (row, part) and walks the batch's nb planes, as the ms2dirty row kernel
// does: the twiddle table, the part's tap factors and the thread's stage
// twiddles are set up once, and while plane p is transformed the walk table
// of plane p + 1 is in flight and each thread's first record of plane p is
// loaded.
template <typename T, int KWM, int S>
__global__ void __launch_bounds__(256) row_ifft_interp_part_planes_kernel(
    const cpx<T> *__restrict__ H, const Rec<T> *__restrict__ recs, const T *__restrict__ taps,
    const int *__restrict__ off, double2 *__restrict__ vacc, const cpx<T> *__restrict__ tw, Geom g, int p0, int nb,
    int sign) {
  constexpr int E = 16, m = 4096;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt[2];
  __shared__ cpx<T> whs[KWM];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int nf1 = g.nf1, N1 = g.N1, nt = blockDim.x, j = threadIdx.x;
  const int item = xcd_interleave(blockIdx.x, gridDim.x, 4 * S);
  const int row0 = item / S, h = item % S;
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(m), tw, nf1);
  if (j < 64) walk_table_build(wt[0], off, g, p0, row0, 1);
  if (j < KWM) {
    // w^(sign h i) for the taps i, from the global table (the LDS copy is being written)
    const int lb = fftc::twiddle_split(nf1);
    const fftc::TwoLevel<T, const cpx<T> *> tg{tw, tw + (nf1 >> lb), lb};
    const cpx<T> w = tg[h * j];
    whs[j] = fftc::mk(w.x, (T)sign * w.y);
  }
  __syncthreads();
  PartTw<T> pt;
  part_twiddles(pt, tv, nf1, sign);
  // input e of this thread is X[S i + h], i = j + nt e; only the first and
  // last quarters can be image columns (as in row_ifft_interp_part_kernel)
  const int h1 = N1 / 2, q = (h + h1) / S, k0 = S * j + h + h1;
  const long estep = (long)(nt >> kTileLog2) * g.tstride, mstep = (long)(m >> kTileLog2) * g.tstride;
  const long a0off = (long)row0 * kTile + h_slot(g, (h << g.slot_lg) + j + q);
  int cur = 0;
  for (int pl = 0; pl < nb; ++pl, cur ^= 1) {
    const int p = p0 + pl;
    const bool more = pl + 1 < nb;
    cpx<T> v[E];
    {
      const cpx<T> *a0 = H + h_plane(g, pl) + a0off;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        v[e] = fftc::mk((T)0, (T)0);
        if (e < E / 4) {
          if (k0 + S * nt * e < N1) v[e] = a0[e * estep];
        } else if (e >= 3 * E / 4) {
          if (k0 + S * nt * e >= nf1) v[e] = a0[e * estep - mstep];
        }
      }
    }
    WalkAhead wa;
    if (j < 64 && more) walk_issue(wa, off, g, p + 1, row0);
    __syncthreads();  // the previous plane's interpolation is done with the LDS; this plane's walk table is complete
    const int total = wt[cur].total;
    if (total > 0) {
      // the first record and (up to 8 taps per axis, where the registers
      // allow) its taps are in flight during the transform
      constexpr bool early = KWM <= 8;
      int idx0 = 0;
      Rec<T> r0;
      Taps<T, KWM> q0;
      if (j < total) {
        idx0 = walk_record(wt[cur], j);
        r0 = recs[idx0];
        if (early) load_taps<T, KWM>(taps, idx0, q0);
      }
      fft4096_part_regs_to_lds(lds, v, sign, pt);
      if (j < total) {
        if (!early) load_taps<T, KWM>(taps, idx0, q0);
        interp_row_part<T, KWM>(r0, q0, idx0, lds, tv, whs, g, p, row0, h, m - 1, sign, vacc);
      }
      for (int t = j + nt; t < total; t += nt) {
        const int idx = walk_record(wt[cur], t);
        Taps<T, KWM> qq;
        load_taps<T, KWM>(taps, idx, qq);
        interp_row_part<T, KWM>(recs[idx], qq, idx, lds, tv, whs, g, p, row0, h, m - 1, sign, vacc);
      }
    }
    if (j < 64 && more) walk_finish(wt[cur ^ 1], wa);
  }
}

// z' (n - 1 - D) gamma h of image pixel (row, col), cuft.cu:53-92 pre-stage
__device__ __forceinline__ double pixel_zp(int row, int col, const Geom &g) {
  double nm1, n;
  pixel_n(row, col, g, nm1, n);
  return (nm1 - g.D) * g.gamma * g.h;
}

// ms2dirty column pass fused with the w-accumulation (w_term_dft, cuft.cu:520-572):
// a workgroup owns cpw grid columns (cpw * nf2 = E * blockDim.x) and walks the
// batch's nb planes: load the row-transformed columns, FFT over nf2, then for
// the N2 image rows accT[col][row] += F_p e^{+i z' (p - P/2)}: the phase
// argument and its reduction in FP64 (sin / cos of the reduced argument at the
// planes' precision) and the sum over planes in FP64 registers.  Image rows per thread:
// cpw * N2 <= 8 * blockDim.x (the plan requires 2 N2 <= nf2).
template <typename T, int E>
__global__ void __launch_bounds__(MaxThreads<T>::value) col_fwd_kernel(const cpx<T> *__restrict__ H,
                                                                      double2 *__restrict__ accT,
                                                                      const cpx<T> *__restrict__ tw, Geom g,
                                                                      int p0, int nb, int first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int cpw = g.cpw, nf2 = g.nf2, N2 = g.N2, nt = blockDim.x;
  const int s0 = xcd_group(blockIdx.x, gridDim.x) * cpw;  // first column slot
  if (!slots_used(s0, g, cpw)) return;
  const int lg = 31 - __clz(nf2);
  const int nout = cpw * N2;
  double ar[8], ai[8], zp[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    ar[e] = ai[e] = zp[e] = 0.0;
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) zp[e] = pixel_zp(row, col, g);
    }
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(cpw * nf2), tw, nf2);
  for (int k = 0; k < nb; ++k) {
    // opaque copy of the thread id: the per-plane addresses are formed inside
    // the loop rather than held in registers across it
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    const cpx<T> *hk = H + h_plane(g, k);
    cpx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * nt;
      const int ci = i >> lg, k2 = i & (nf2 - 1), sl = s0 + ci;
      v[e] = slot_column(sl, g) >= 0 ? hk[h_slot(g, sl) + (k2 << kTileLog2)]
                                     : fftc::mk((T)0, (T)0);
    }
    __syncthreads();  // the previous plane's LDS reads are done
#pragma unroll
    for (int e = 0; e < E; ++e) lds[fftc::pad(tid + e * nt)] = v[e];
    __syncthreads();
    lds_fft_rows<T, E, 16>(lds, lg, cpw, +1, tv);
    const double m = (double)(p0 + k - g.P / 2);
    int tid2 = threadIdx.x;
    __asm__ volatile("" : "+v"(tid2));
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int o = tid2 + e * nt;
      if (o < nout) {
        const int ci = o / N2, row = o - ci * N2;
        const cpx<T> f = lds[fftc::pad(ci * nf2 + grid_index(row, N2, nf2))];
        double sn, cs;
        plane_phase<T>(zp[e] * m, &sn, &cs);
        const double fr = (double)f.x, fi = (double)f.y;
        ar[e] += fr * cs - fi * sn;
        ai[e] += fr * sn + fi * cs;
      }
      // one element's phase at a time (bounds the live FP64 temporaries)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) {
        const size_t a = (size_t)col * N2 + row;
        double2 r = make_double2(ar[e], ai[e]);
        if (!first) {
          const double2 q = accT[a];
          r.x += q.x;
          r.y += q.y;
        }
        accT[a] = r;
      }
    }
  }
}

// dirty2ms column pass fused with the w-fill (w_term_idft, cuft.cu:575-612): a
// workgroup owns cpw grid columns and, for each of the batch's nb planes, places
// gT[col][row] e^{-i z' (p - P/2)} (FP64 phase argument) at the image rows of its columns
// (zeros elsewhere), runs the inverse FFT over nf2 and writes the full columns
// to the column-tiled H.
template <typename T, int E>
__global__ void __launch_bounds__(MaxThreads<T>::value) col_inv_kernel(const cpx<T> *__restrict__ gT,
                                                                      cpx<T> *__restrict__ H,
                                                                      const cpx<T> *__restrict__ tw, Geom g,
                                                                      int p0, int nb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int cpw = g.cpw_inv, nf2 = g.nf2, N2 = g.N2, nt = blockDim.x;
  const int s0 = xcd_group(blockIdx.x, gridDim.x) * cpw;  // first column slot
  if (!slots_used(s0, g, cpw)) return;
  const int lg = 31 - __clz(nf2);
  const int nout = cpw * N2, nzero = cpw * (nf2 - N2), h2 = N2 / 2;
  bool full = true;  // every column slot of the workgroup holds an image column
  for (int c = 0; c < cpw; ++c) full = full && slot_column(s0 + c, g) >= 0;
  cpx<T> gv[8];
  double zp[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    gv[e] = fftc::mk((T)0, (T)0);
    zp[e] = 0.0;
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) {
        gv[e] = gT[(size_t)col * N2 + row];
        zp[e] = pixel_zp(row, col, g);
      }
    }
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(cpw * nf2), tw, nf2);
  for (int k = 0; k < nb; ++k) {
    const double m = (double)(p0 + k - g.P / 2);
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    __syncthreads();  // the previous plane's LDS reads are done
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int o = tid + e * nt;
      if (o < nout) {
        const int ci = o / N2, row = o - ci * N2;
        double sn, cs;
        plane_phase<T>(-zp[e] * m, &sn, &cs);
        const double gr = (double)gv[e].x, gi = (double)gv[e].y;
        lds[fftc::pad(ci * nf2 + grid_index(row, N2, nf2))] =
            fftc::mk((T)(gr * cs - gi * sn), (T)(gr * sn + gi * cs));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int z = tid; z < nzero; z += nt) {
      const int ci = z / (nf2 - N2), j = z - ci * (nf2 - N2);
      lds[fftc::pad(ci * nf2 + N2 - h2 + j)] = fftc::mk((T)0, (T)0);
    }
    __syncthreads();
    lds_fft_rows<T, E, 16>(lds, lg, cpw, -1, tv);
    int tid2 = threadIdx.x;
    __asm__ volatile("" : "+v"(tid2));
    cpx<T> *hk = H + h_plane(g, k);
    if (sizeof(T) == 4 && kTile == 4 && cpw == 2) {
      // half tiles: thread k2 writes row k2's two columns as one 16-byte piece
#pragma unroll
      for (int q = 0; q < E / 2; ++q) {
        const int k2 = tid2 + q * nt;
        const cpx<T> a = lds[fftc::pad(k2)], b = lds[fftc::pad(nf2 + k2)];
        cpx<T> *d = hk + h_slot(g, s0) + (k2 << kTileLog2);
        if (full) {
          *reinterpret_cast<float4 *>(d) = make_float4(a.x, a.y, b.x, b.y);
        } else {
          if (slot_column(s0, g) >= 0) d[0] = a;
          if (slot_column(s0 + 1, g) >= 0) d[1] = b;
        }
      }
    } else if (sizeof(T) == 4 && kTile == 4 && cpw % kTile == 0) {
      // whole column tiles: 16-byte pieces, piece r of tile t holding row
      // r / 2, columns 2 (r & 1) and 2 (r & 1) + 1 of the tile, so that
      // consecutive lanes write consecutive 16 bytes of H
      const int tlg = lg + 1;  // 2 nf2 pieces per tile
      float4 w[E / 2];
#pragma unroll
      for (int q = 0; q < E / 2; ++q) {
        const int P = tid2 + q * nt, t = P >> tlg, r = P & ((1 << tlg) - 1), c = kTile * t + 2 * (r & 1);
        const cpx<T> a = lds[fftc::pad(c * nf2 + (r >> 1))], b = lds[fftc::pad((c + 1) * nf2 + (r >> 1))];
        w[q] = make_float4(a.x, a.y, b.x, b.y);
      }
      if (full && cpw == kTile) {
        // one tile: its pieces are contiguous
        float4 *d = reinterpret_cast<float4 *>(hk + h_slot(g, s0)) + tid2;
#pragma unroll
        for (int q = 0; q < E / 2; ++q) d[q * nt] = w[q];
      } else {
#pragma unroll
        for (int q = 0; q < E / 2; ++q) {
          const int P = tid2 + q * nt, t = P >> tlg, r = P & ((1 << tlg) - 1), c = kTile * t + 2 * (r & 1);
          cpx<T> *d = hk + h_slot(g, s0 + kTile * t) + 2 * r;  // 16-byte aligned: tiles start on 32 bytes
          if (full) {
            *reinterpret_cast<float4 *>(d) = w[q];
          } else {
            if (slot_column(s0 + c, g) >= 0) d[0] = fftc::mk((T)w[q].x, (T)w[q].y);
            if (slot_column(s0 + c + 1, g) >= 0) d[1] = fftc::mk((T)w[q].z, (T)w[q].w);
          }
        }
      }
    } else {
      cpx<T> v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[fftc::pad(tid2 + e * nt)];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = tid2 + e * nt;
        const int ci = i >> lg, k2 = i & (nf2 - 1), sl = s0 + ci;
        if (slot_column(sl, g) >= 0) hk[h_slot(g, sl) + (k2 << kTileLog2)] = v[e];
      }
    }
  }
}

// --------------------------------------------------------------------------
// FP64 columns of 16384 (256 KiB, more than a CU's LDS): one workgroup of 1024
// threads per column slot keeps the column in registers, 16 elements per
// thread, and runs the four Stockham stages (radices 16, 16, 16, 4) with the
// exchanges between stages through a one-component LDS buffer (16384 doubles
// plus pad), the real parts first, then the imaginary parts.  Thread j holds
// x[j + 1024 r] in v[r] before each radix-16 stage and, after the radix-4
// stage, X[j + 1024 b + 4096 r] in v[4 b + r] (b, r = 0..3).
// --------------------------------------------------------------------------
namespace colr {
constexpr int N = 16384, NT = 1024;
// elements of the exchange buffer (rounded up to 16 bytes: the twiddles
// follow): kBuf doubles in FP64, kBuf complex values in FP32
constexpr int kBuf = fftc::padded_size(N) + 1;
template <typename T> __device__ __forceinline__ constexpr int buf_scalars() { return sizeof(T) == 4 ? 2 * kBuf : kBuf; }

// v through LDS: stored at padded position st(r), reloaded from ld(r) (each a
// base plus a constant step, one address register); in FP32 as complex values,
// in FP64 one component at a time (the column of 16384 complex doubles does
// not fit the LDS)
template <typename T, typename St, typename Ld>
__device__ __forceinline__ void exchange(cpx<T> *v, T *buf, St st, Ld ld) {
  if constexpr (sizeof(T) == 4) {
    cpx<T> *b = reinterpret_cast<cpx<T> *>(buf);
#pragma unroll
    for (int r = 0; r < 16; ++r) b[st(r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = b[ld(r)];
    __syncthreads();
    return;
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) buf[st(r)] = v[r].x;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].x = buf[ld(r)];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) buf[st(r)] = v[r].y;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].y = buf[ld(r)];
  __syncthreads();
}

// u[r] *= w^r (sign applied to w) for r = 1..R-1, w = tv[base]: w, w^2, w^4, w^8
// from the table, each other power formed next to its use (few live registers)
template <typename T, int R, typename Tw>
__device__ __forceinline__ void apply_twiddles(cpx<T> *u, const Tw &tv, int base, int sign) {
  const int n = N - 1;
  cpx<T> p[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if ((1 << i) < R) {
      const cpx<T> w = tv[(base << i) & n];
      p[i] = fftc::mk(w.x, (T)sign * w.y);
    }
  }
#pragma unroll
  for (int r = 1; r < R; ++r) {
    // w^r as the product of the powers of two in r
    cpx<T> w = p[__builtin_ctz(r)];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (((r >> i) & 1) && i != __builtin_ctz(r)) w = fftc::mul(w, p[i]);
    u[r] = fftc::mul(u[r], w);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// the thread id through an empty asm: the stage addresses are formed where they
// are used (hoisted out of the plane loop they would not fit the registers)
__device__ __forceinline__ int opaque_tid() {
  int j = threadIdx.x;
  __asm__ volatile("" : "+v"(j));
  return j;
}

// 16384-point FFT of the column held as described above (sign +1 forward in the
// sense of the row pass, -1 inverse); tv: twiddles e^{2 pi i t / 16384}.
// Padded positions: pad(j + 1024 r) = pad(j) + 1088 r, pad(16 j + r) = 17 j + r,
// pad(d + 16 r) = pad(d) + 17 r and pad(d + 256 r) = pad(d) + 272 r for d mod 16 = d mod 16 < 16.
template <typename T, typename Tw>
__device__ __forceinline__ void fft(cpx<T> *v, T *buf, int sign, const Tw &tv) {
  // stage 1 (sub-transform size 1)
  fftc::dft_small<T, 16>(v, sign);
  __builtin_amdgcn_sched_barrier(0);
  {
    const int j = opaque_tid(), s0 = 17 * j, l0 = fftc::pad(j);
    exchange(v, buf, [&](int r) { return s0 + r; }, [&](int r) { return l0 + 1088 * r; });
  }
  // stage 2 (size 16): twiddles W256^(k r), k = j mod 16
  {
    const int j = opaque_tid(), k = j & 15;
    apply_twiddles<T, 16>(v, tv, 64 * k, sign);
    fftc::dft_small<T, 16>(v, sign);
    __builtin_amdgcn_sched_barrier(0);
    const int j2 = opaque_tid(), k2 = j2 & 15, s0 = fftc::pad(16 * (j2 - k2) + k2), l0 = fftc::pad(j2);
    exchange(v, buf, [&](int r) { return s0 + 17 * r; }, [&](int r) { return l0 + 1088 * r; });
  }
  // stage 3 (size 256): twiddles W4096^(k r), k = j mod 256; reloaded as the
  // radix-4 inputs j + 1024 b + 4096 r
  {
    const int j = opaque_tid(), k = j & 255;
    apply_twiddles<T, 16>(v, tv, 4 * k, sign);
    fftc::dft_small<T, 16>(v, sign);
    __builtin_amdgcn_sched_barrier(0);
    const int j2 = opaque_tid(), k2 = j2 & 255, s0 = fftc::pad(16 * (j2 - k2) + k2), l0 = fftc::pad(j2);
    exchange(v, buf, [&](int r) { return s0 + 272 * r; },
             [&](int e) { return l0 + 1088 * (e >> 2) + 4352 * (e & 3); });
  }
  // stage 4 (size 4096, radix 4): butterfly beta = j + 1024 b, twiddles W16384^(beta r)
  const int j = opaque_tid();
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    apply_twiddles<T, 4>(v + 4 * b, tv, j + 1024 * b, sign);
    fftc::dft_small<T, 4>(v + 4 * b, sign);
    __builtin_amdgcn_sched_barrier(0);
  }
}
}  // namespace colr

// ms2dirty column pass for FP64 columns of 16384 (col_fwd_kernel's work: column
// FFT, then the N2 image rows accumulated with their w phases in FP64).  Of the
// outputs X[j + 1024 b + 4096 r] only r = 0 and r = 3 can be image rows (grid
// indices within N2 / 2 <= nf2 / 4 of 0).
template <typename T>
__global__ void __launch_bounds__(colr::NT) col_fwd_reg_kernel(const cpx<T> *__restrict__ H,
                                                               double2 *__restrict__ accT,
                                                               const cpx<T> *__restrict__ tw, Geom g, int p0,
                                                               int nb, int first) {
  using colr::N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *buf = reinterpret_cast<T *>(smem_raw);
  const int j = threadIdx.x, N2 = g.N2;
  const int sl = xcd_group(blockIdx.x, gridDim.x);  // column slot
  const int col = slot_column(sl, g);
  if (col < 0) return;
  const auto tv = lds_twiddles<T>(reinterpret_cast<cpx<T> *>(buf + colr::buf_scalars<T>()), tw, N);
  // output q of this thread: X[j + 1024 (q mod 4) + 4096 r], r = 0 (q < 4) or 3;
  // its image row, or -1 (rows and w factors are recomputed per plane: held
  // across the plane loop they would not fit the registers beside the column)
  auto out_row = [&](int q) { return image_index(colr::opaque_tid() + 1024 * (q & 3) + (q < 4 ? 0 : 3 * 4096), N2, N); };
  double ar[8], ai[8], zq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    ar[q] = ai[q] = 0.0;
    // FP32 columns leave room to keep the w arguments across the planes
    const int row = out_row(q);
    zq[q] = (sizeof(T) == 4 && row >= 0) ? pixel_zp(row, col, g) : 0.0;
  }
  __syncthreads();
  const cpx<T> *hs = H + h_slot(g, sl);
  for (int k = 0; k < nb; ++k) {
    const cpx<T> *hk = hs + h_plane(g, k);
    cpx<T> v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = hk[(size_t)(j + 1024 * r) << kTileLog2];
    colr::fft(v, buf, +1, tv);
    const double m = (double)(p0 + k - g.P / 2);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = out_row(q);
      if (row >= 0) {
        const cpx<T> f = v[4 * (q & 3) + (q < 4 ? 0 : 3)];
        double sn, cs;
        plane_phase<T>((sizeof(T) == 4 ? zq[q] : pixel_zp(row, col, g)) * m, &sn, &cs);
        const double fr = (double)f.x, fi = (double)f.y;
        ar[q] += fr * cs - fi * sn;
        ai[q] += fr * sn + fi * cs;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int row = out_row(q);
    if (row >= 0) {
      const size_t a = (size_t)col * N2 + row;
      double2 r = make_double2(ar[q], ai[q]);
      if (!first) {
        const double2 o = accT[a];
        r.x += o.x;
        r.y += o.y;
      }
      accT[a] = r;
    }
  }
}

// dirty2ms column pass for FP64 columns of 16384 (col_inv_kernel's work: the
// image rows of the column with their w phases, zeros elsewhere, inverse FFT,
// the full column written to H).  The inputs x[j + 1024 r] are image rows for
// r = 0..3 and 12..15 only.
template <typename T>
__global__ void __launch_bounds__(colr::NT) col_inv_reg_kernel(const cpx<T> *__restrict__ gT,
                                                               cpx<T> *__restrict__ H,
                                                               const cpx<T> *__restrict__ tw, Geom g, int p0,
                                                               int nb) {
  using colr::N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *buf = reinterpret_cast<T *>(smem_raw);
  const int j = threadIdx.x, N2 = g.N2;
  const int sl = xcd_group(blockIdx.x, gridDim.x);  // column slot
  const int col = slot_column(sl, g);
  if (col < 0) return;
  const auto tv = lds_twiddles<T>(reinterpret_cast<cpx<T> *>(buf + colr::buf_scalars<T>()), tw, N);
  // input q of this thread: x[j + 1024 r], r = q (q < 4) or q + 8; its image
  // row, or -1 (recomputed per plane, as in col_fwd_reg_kernel)
  auto in_row = [&](int q) { return image_index(colr::opaque_tid() + 1024 * (q < 4 ? q : q + 8), N2, N); };
  // FP32 columns leave room to keep the w arguments and the image values
  // across the planes
  constexpr bool keep = sizeof(T) == 4;
  double zq[8];
  cpx<T> gq[8];
  const cpx<T> *gc = gT + (size_t)col * N2;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int row = in_row(q);
    zq[q] = (keep && row >= 0) ? pixel_zp(row, col, g) : 0.0;
    gq[q] = (keep && row >= 0) ? gc[row] : fftc::mk((T)0, (T)0);
  }
  __syncthreads();
  cpx<T> *hs = H + h_slot(g, sl);
  for (int k = 0; k < nb; ++k) {
    const double m = (double)(p0 + k - g.P / 2);
    cpx<T> v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = fftc::mk((T)0, (T)0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = in_row(q);
      if (row >= 0) {
        const cpx<T> gv = keep ? gq[q] : gc[row];
        double sn, cs;
        plane_phase<T>(-(keep ? zq[q] : pixel_zp(row, col, g)) * m, &sn, &cs);
        const double gr = (double)gv.x, gi = (double)gv.y;
        v[q < 4 ? q : q + 8] = fftc::mk((T)(gr * cs - gi * sn), (T)(gr * sn + gi * cs));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    colr::fft(v, buf, -1, tv);
    cpx<T> *hk = hs + h_plane(g, k);
#pragma unroll
    for (int e = 0; e < 16; ++e) hk[(size_t)(j + 1024 * (e >> 2) + 4096 * (e & 3)) << kTileLog2] = v[e];
  }
}

// --------------------------------------------------------------------------
// launchers (the LDS path takes kernel widths up to 16: the plan sends wider
// kernels to the full-plane path)
// --------------------------------------------------------------------------
template <typename K>
int allow_lds(K kernel, int bytes) {
  if (bytes > 65536) CURIG_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  return kOk;
}

inline int col_groups(const Geom &g, int cpw) { return (g.nslot + cpw - 1) / cpw; }

template <typename T> inline bool whole_row_planes(const Geom &g) {
  return g.rsplit == 0 && g.rpw == 1 && g.nf1 == 4096;
}

// ms2dirty, one batch on the LDS path: spread + row FFT -> H (column-tiled)
template <typename T>
void lds_spread_rows(const Rec<T> *recs, const T *taps, const int *off, const Geom &g, int p0, int nb, cpx<T> *H,
                     const cpx<T> *tw1, hipStream_t st) {
  if (g.rsplit == kRowSplit) {
    const int nt = g.nf1 / (16 * kRowSplit), grd = kRowSplit * g.nf2;
    const int lb = g.lds_rowp_bytes;
    if (g.kw <= 8)
      spread_row_part_planes_kernel<T, 16, 8, kRowSplit><<<grd, nt, lb, st>>>(recs, taps, off, H, tw1, g, p0, nb);
    else
      spread_row_part_planes_kernel<T, 16, 16, kRowSplit><<<grd, nt, lb, st>>>(recs, taps, off, H, tw1, g, p0, nb);
    return;
  }
  if (whole_row_planes<T>(g)) {
    // rows of 4096: one workgroup per row walks the batch's planes
    const int lb = g.lds_row_bytes;
    if (g.kw <= 8)
      spread_row_part_planes_kernel<T, 16, 8, 1><<<g.nf2, 256, lb, st>>>(recs, taps, off, H, tw1, g, p0, nb);
    else
      spread_row_part_planes_kernel<T, 16, 16, 1><<<g.nf2, 256, lb, st>>>(recs, taps, off, H, tw1, g, p0, nb);
    return;
  }
  const dim3 grd(g.nf2 / g.rpw, nb);
  const int nt = g.rpw * g.nf1 / 16;
  if (g.kw <= 8)
    spread_row_fft_kernel<T, 16, 8><<<grd, nt, g.lds_row_bytes, st>>>(recs, taps, off, H, tw1, g, p0);
  else
    spread_row_fft_kernel<T, 16, 16><<<grd, nt, g.lds_row_bytes, st>>>(recs, taps, off, H, tw1, g, p0);
}

template <typename T>
int allow_spread_rows(const Geom &g) {
  CURIG_TRY(allow_lds(spread_row_fft_kernel<T, 16, 8>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(spread_row_fft_kernel<T, 16, 16>, g.lds_row_bytes));
  if (whole_row_planes<T>(g)) {
    const int lb = g.lds_row_bytes;
    CURIG_TRY((allow_lds(spread_row_part_planes_kernel<T, 16, 8, 1>, lb)));
    CURIG_TRY((allow_lds(spread_row_part_planes_kernel<T, 16, 16, 1>, lb)));
  }
  if (g.rsplit == kRowSplit) {
    const int lb = g.lds_rowp_bytes;
    CURIG_TRY((allow_lds(spread_row_part_planes_kernel<T, 16, 8, kRowSplit>, lb)));
    CURIG_TRY((allow_lds(spread_row_part_planes_kernel<T, 16, 16, kRowSplit>, lb)));
  }
  return kOk;
}

// rows (rsplit 0) or row parts (rsplit kRowSplit) of 4096 points (FP64:
// 173 VGPRs at the two workgroups per CU that its LDS allows)
template <typename T> inline bool interp_planes(const Geom &g) {
  return (g.rsplit == 0 && g.rpw == 1 && g.nf1 == 4096) || (g.rsplit == kRowSplit && g.nf1 / kRowSplit == 4096);
}

// dirty2ms, one batch on the LDS path: inverse row FFT fused with the interpolation
template <typename T>
void lds_row_interp(const Rec<T> *recs, const T *taps, const int *off, const cpx<T> *H, double2 *vacc,
                    const cpx<T> *tw1, const Geom &g, int p0, int nb, hipStream_t st) {
  if (interp_planes<T>(g)) {
    // rows or row parts of 4096: one workgroup per (row, part) walks the batch's planes
    const int S = g.rsplit == kRowSplit ? kRowSplit : 1;
    const int lb = S == 1 ? g.lds_row_bytes : g.lds_rowp_bytes, grd = S * g.nf2;
    if (S == 1) {
      if (g.kw <= 8)
        row_ifft_interp_part_planes_kernel<T, 8, 1><<<grd, 256, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0, nb, -1);
      else
        row_ifft_interp_part_planes_kernel<T, 16, 1><<<grd, 256, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0, nb, -1);
    } else {
      if (g.kw <= 8)
        row_ifft_interp_part_planes_kernel<T, 8, kRowSplit><<<grd, 256, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0, nb,
                                                                                -1);
      else
        row_ifft_interp_part_planes_kernel<T, 16, kRowSplit><<<grd, 256, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0,
                                                                                 nb, -1);
    }
    return;
  }
  if (g.rsplit == kRowSplit) {
    const dim3 grd(kRowSplit * g.nf2, nb);
    const int nt = g.nf1 / (16 * kRowSplit), lb = g.lds_rowp_bytes;
    if (g.kw <= 8)
      row_ifft_interp_part_kernel<T, 16, 8, kRowSplit><<<grd, nt, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0, -1);
    else
      row_ifft_interp_part_kernel<T, 16, 16, kRowSplit><<<grd, nt, lb, st>>>(H, recs, taps, off, vacc, tw1, g, p0, -1);
    return;
  }
  const dim3 grd(g.nf2 / g.rpw, nb);
  const int nt = g.rpw * g.nf1 / 16;
  if (g.kw <= 8)
    row_ifft_interp_kernel<T, 16, 8><<<grd, nt, g.lds_row_bytes, st>>>(H, recs, taps, off, vacc, tw1, g, p0, -1);
  else
    row_ifft_interp_kernel<T, 16, 16><<<grd, nt, g.lds_row_bytes, st>>>(H, recs, taps, off, vacc, tw1, g, p0, -1);
}

template <typename T>
int allow_row_interp(const Geom &g) {
  CURIG_TRY(allow_lds(row_ifft_interp_kernel<T, 16, 8>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_kernel<T, 16, 16>, g.lds_row_bytes));
  if (g.rsplit == kRowSplit) {
    CURIG_TRY((allow_lds(row_ifft_interp_part_kernel<T, 16, 8, kRowSplit>, g.lds_rowp_bytes)));
    CURIG_TRY((allow_lds(row_ifft_interp_part_kernel<T, 16, 16, kRowSplit>, g.lds_rowp_bytes)));
  }
  if (interp_planes<T>(g)) {
    const bool s1 = g.rsplit != kRowSplit;
    const int lb = s1 ? g.lds_row_bytes : g.lds_rowp_bytes;
    CURIG_TRY((allow_lds(row_ifft_interp_part_planes_kernel<T, 8, 1>, lb)));
    CURIG_TRY((allow_lds(row_ifft_interp_part_planes_kernel<T, 16, 1>, lb)));
    CURIG_TRY((allow_lds(row_ifft_interp_part_planes_kernel<T, 8, kRowSplit>, lb)));
    CURIG_TRY((allow_lds(row_ifft_interp_part_planes_kernel<T, 16, kRowSplit>, lb)));
  }
  return kOk;
}

// ms2dirty, one batch on the LDS path: column FFT + w-accumulation into accT
template <typename T>
void lds_cols_forward(const Geom &g, int p0, int nb, const cpx<T> *H, double2 *accT, const cpx<T> *tw2, int first,
                      hipStream_t st) {
  if (g.colreg) {
    col_fwd_reg_kernel<T><<<col_groups(g, 1), colr::NT, g.lds_col_bytes, st>>>(H, accT, tw2, g, p0, nb, first);
    return;
  }
  col_fwd_kernel<T, 16><<<col_groups(g, g.cpw), g.cpw * g.nf2 / 16, g.lds_col_bytes, st>>>(H, accT, tw2, g, p0, nb, first);
}

// dirty2ms, one batch on the LDS path: w-fill + inverse column FFT -> H (column-tiled)
template <typename T>
void lds_cols_backward(const Geom &g, int p0, int nb, const cpx<T> *gT, cpx<T> *H, const cpx<T> *tw2,
                       hipStream_t st) {
  if (g.colreg) {
    col_inv_reg_kernel<T><<<col_groups(g, 1), colr::NT, g.lds_col_bytes, st>>>(gT, H, tw2, g, p0, nb);
    return;
  }
  col_inv_kernel<T, 16><<<col_groups(g, g.cpw_inv), g.cpw_inv * g.nf2 / 16, g.lds_coli_bytes, st>>>(gT, H, tw2, g, p0, nb);
}

template <typename T>
int allow_cols(const Geom &g) {
  // every column workgroup needs its LDS block (a launch with less would
  // address LDS it does not own)
  if (g.lds_col_bytes <= 0 || (!g.colreg && g.lds_coli_bytes <= 0)) {
    set_error("column pass without an LDS size (%d, %d)", g.lds_col_bytes, g.lds_coli_bytes);
    return kErrArg;
  }
  if (g.colreg) {
    CURIG_TRY(allow_lds(col_fwd_reg_kernel<T>, g.lds_col_bytes));
    CURIG_TRY(allow_lds(col_inv_reg_kernel<T>, g.lds_col_bytes));
    return kOk;
  }
  CURIG_TRY(allow_lds(col_fwd_kernel<T, 16>, g.lds_col_bytes));
  CURIG_TRY(allow_lds(col_inv_kernel<T, 16>, g.lds_coli_bytes));
  return kOk;
}

// one source per (precision, kernel group), so the heavy kernels compile in parallel
#define CURIG_LDS_SPREAD_INST(T)                                                                                   \
  template void lds_spread_rows<T>(const Rec<T> *, const T *, const int *, const Geom &, int, int, cpx<T> *,       \
                                   const cpx<T> *, hipStream_t);                                                   \
  template int allow_spread_rows<T>(const Geom &);
#define CURIG_LDS_INTERP_INST(T)                                                                                   \
  template void lds_row_interp<T>(const Rec<T> *, const T *, const int *, const cpx<T> *, double2 *,              \
                                  const cpx<T> *, const Geom &, int, int, hipStream_t);                            \
  template int allow_row_interp<T>(const Geom &);
#define CURIG_LDS_COLS_INST(T)                                                                                     \
  template void lds_cols_forward<T>(const Geom &, int, int, const cpx<T> *, double2 *, const cpx<T> *, int,        \
                                    hipStream_t);                                                                  \
  template void lds_cols_backward<T>(const Geom &, int, int, const cpx<T> *, cpx<T> *, const cpx<T> *,             \
                                     hipStream_t);                                                                 \
  template int allow_cols<T>(const Geom &);

}  // namespace curig
