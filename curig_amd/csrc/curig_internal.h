// Internal declarations shared by the gridder's host and device sources.
//
// Layout conventions (fixed by the reference, src/astro/cugridder.cu:188,399):
//   image  dirty[nx][ny], x-major; N1 = ny (fast axis, pairs with v), N2 = nx (pairs with u)
//   grid   plane[P][nf2][nf1]; the grid's fast axis (nf1) carries v, the slow axis (nf2) carries u
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace curig {

// Return codes of every entry point (0 = success).  1 and 2 keep the meaning
// setup_conv_opts gives them in the reference (conv_interp_invoker.cu:46-116).
enum Status : int {
  kOk = 0,
  kErrEpsTooSmall = 1,  // epsilon < 1.1e-16
  kErrSigma = 2,        // upsampling factor <= 1
  kErrSize = 3,         // image size <= 0 or grid too large
  kErrNrow = 4,         // nrow < 0
  kErrArg = 5,          // NULL pointer, sign not +-1, bad option
  kErrHip = 6,          // HIP runtime error
  kErrAlloc = 7,        // device allocation failed
  kErrFft = 8,          // rocFFT error
  kErrNonFinite = 9,    // non-finite uvw / frequency
  kErrScratch = 10,     // request exceeds a fixed-size buffer
};

void set_error(const char *fmt, ...);
const char *last_error();

#define CURIG_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::curig::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,                \
                         hipGetErrorString(e_));                                     \
      return (e_ == hipErrorOutOfMemory) ? ::curig::kErrAlloc : ::curig::kErrHip;     \
    }                                                                                \
  } while (0)

#define CURIG_TRY(expr)        \
  do {                         \
    int s_ = (expr);           \
    if (s_ != 0) return s_;    \
  } while (0)

constexpr double kSpeedOfLight = 299792458.0;  // include/utils.h:70
constexpr double kEpsilonMin = 1.1e-16;        // include/utils.h:66
constexpr int kMaxKernelWidth = 28;            // include/curafft_plan.h:14
constexpr int kMaxQuadNodes = 2 * (2 + 3 * kMaxKernelWidth / 2);  // 2p, deconv.cu:75

// ES kernel options (conv_opts, include/curafft_plan.h:28-48)
struct KernelOpts {
  int kw = 0;
  double beta = 0, c = 0, halfwidth = 0, upsampfac = 2.0;
};

// Host plan math (plan_math.cpp)
int setup_conv_opts(KernelOpts &o, double eps, double upsampfac);
int next235beven(int n, int b);
int get_num_cells(int ms, const KernelOpts &o);
void set_nhg_type3(double S, double X, const KernelOpts &o, int &nf, double &h, double &gam);
void gauss_legendre(int n, double *x, double *w);
// 2p scaled nodes x_n and weights g_n = alpha w_n phi_norm(x_n); returns 2p
int kernel_quadrature(const KernelOpts &o, double *x, double *g);
// phihat_norm(k) (-1)^k for k = 0..nf/2 (normalised kernel, peak 1)
void fseries_grid(const KernelOpts &o, int nf, double *out);
// piecewise-polynomial kernel: kw intervals x (deg + 1) coefficients in powers
// of z in [-1, 1], highest first; returns deg (-1 when tol is not reached)
constexpr int kMaxHornerDeg = 24;
int horner_fit(const KernelOpts &o, double tol, std::vector<double> &coef);
double horner_tol(double eps, bool dbl);

// Everything the device kernels need for one (channel, direction) problem.
struct Geom {
  int N1, N2;         // image: N1 = ny (fast), N2 = nx
  int nf1, nf2, P;    // grid sizes: nf1 (v), nf2 (u), w planes
  int kw;
  int kww;            // kernel width along w: kw, or 1 in the 2-D form (one plane, weight 1)
  int hdeg;           // >= 0: kernel values by Horner's rule on the piecewise polynomial below
  const double *hcoef_d;  // [hdeg + 1][kw] coefficients (device; one power of every tap contiguous), FP64 and FP32 copies
  const float *hcoef_f;
  double beta, c;     // normalised ES kernel exp(beta (sqrt(1 - c d^2) - 1))
  double dx, dy;      // pixel sizes (rad): dx pairs with u / image rows, dy with v / columns
  int divn;           // 1: image = sum / n (w-stacking); 0: w taken as 0 and no 1/n (2-D transform)
  const double *zout; // device [N2][N1]: each pixel's n - 1 given (type-3 outputs; n taken as 1), or null
  double C, D, gamma, h;  // type-3 w parameters (t3attrib): i_center, o_center, gamma, h
  double uscale, vscale, wscale;  // 2 pi f/c dx, 2 pi f/c dy, 2 pi f/c sign
  int T1, T2, ntx, nty, ntiles;   // tile size (cells) along nf1 / nf2, tile counts
  int Pb, nbatch;     // planes per batch, number of batches
  int Pbin, nlayer;   // planes per bin layer (rows are binned by first plane / Pbin), layers
  int nquad;          // number of quadrature nodes for phihat3
  // pruned LDS FFT path (power-of-two grids); fft_lds == 0 selects rocFFT
  int fft_lds;
  int rpw;            // grid rows per workgroup in the row pass (rpw * nf1 = 16 * threads)
  int cpw;            // grid columns per workgroup in the column pass (cpw * nf2 = 16 * threads)
  int cpw_inv;        // the same for the dirty2ms column pass (whole column tiles where they fit)
  int lds_coli_bytes; // its LDS
  int nct;            // column tiles of a plane's row-transformed layout (nslot / kTile, rounded up)
  int tstride;        // elements from one column tile to the next: nf2 * kTile + a pad
  int nslot;          // column slots of that layout: N1, or nf1 / 2 with split rows
  int slot_lg;        // split rows: log2 of the slots per row part (nf1 / (2 rsplit))
  int lds_row_bytes;  // dynamic LDS of the row-pass workgroup
  int lds_col_bytes;  // dynamic LDS of the column-pass workgroup
  int rsplit;         // > 1: rows run as rsplit workgroups of nf1 / rsplit (too long for two per CU)
  int lds_rowp_bytes; // dynamic LDS of a row-part workgroup
  int colreg;         // 1: FP64 columns of 16384, held in registers by 1024 threads (one column slot each)
};

// Column-tile width of the intermediate planes (LDS path, both directions):
// 2^kTileLog2 image columns per contiguous chunk.  4 columns (32 bytes FP32)
// measured fastest at 8192^2 on MI355X (2: row writes too narrow; 8 and 16:
// column reads too sparse; column-major planes for dirty2ms: strided row reads).
#ifndef CURIG_TILE_LOG2
#define CURIG_TILE_LOG2 2
#endif
constexpr int kTileLog2 = CURIG_TILE_LOG2;
// Bytes between consecutive column tiles beyond their nf2 * kTile elements.
// Without it the tiles of a plane sit 2^k bytes apart and the kTile-column
// chunks that one row writes (or reads) all fall into the same memory channel.
#ifndef CURIG_TILE_PAD_BYTES
#define CURIG_TILE_PAD_BYTES 4352
#endif
constexpr int kTilePadBytes = CURIG_TILE_PAD_BYTES;
// Parts a long row is split into (2, 4 or 8; see the row-part kernels).  4
// measured fastest at 8192^2 on MI355X (FP32 row pass 27.4 -> 25.1 ms and
// interpolation 24.5 -> 20.5 ms per 32-plane launch against 2).
#ifndef CURIG_ROW_SPLIT
#define CURIG_ROW_SPLIT 4
#endif
static_assert(CURIG_ROW_SPLIT == 2 || CURIG_ROW_SPLIT == 4 || CURIG_ROW_SPLIT == 8, "row split 2, 4 or 8");
constexpr int kRowSplit = CURIG_ROW_SPLIT;

// Tunables (autotuned per problem class; defaults are safe for any size)
struct Tuning {
  int T1 = 0, T2 = 0;   // 0 -> default (tile: 64 x 32 FP32, 32 x 32 FP64; row strip: 2 rows)
  int Pbin = 0;         // planes per bin layer (0 -> 4)
  int Pb = 0;           // 0 -> from the memory budget
  int fft = 0;          // 0 auto (LDS path when it applies), 1 force rocFFT
  int lds_bytes = 0;    // LDS per workgroup of the device (0: 160 KiB, gfx950)
  int kereval = 0;      // kernel evaluation: 0 exp(sqrt()), 1 piecewise polynomial (Horner)
  size_t grid_budget_bytes = size_t(48) << 30;
};

}  // namespace curig
