#pragma once
#include <rocfft/rocfft.h>

#include "curig_internal.h"

namespace curig {

struct FftPlan {
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  size_t work_bytes = 0;
};

// Batched in-place 2-D complex FFT of `batch` planes of nf2 rows x nf1 columns.
// sign = +1: sum e^{+2 pi i ...} (ms2dirty), -1: e^{-2 pi i ...} (dirty2ms).
int fft_plan(int device, int nf1, int nf2, int batch, bool dbl, int sign, FftPlan *out);
int fft_exec(const FftPlan &p, void *data, void *work, hipStream_t stream);
void fft_release_all();

}  // namespace curig
