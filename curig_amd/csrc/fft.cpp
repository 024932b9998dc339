// rocFFT plans for the batched in-place 2-D transforms of w planes.
// Plans are cached for the life of the process (the reference creates and
// destroys a cuFFT plan per call, src/astro/cugridder.cu:224-246).
#include <rocfft/rocfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "curig_internal.h"
#include "fft.h"

namespace curig {

namespace {
struct PlanEntry {
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  size_t work = 0;
};
using Key = std::tuple<int, int, int, int, int, int>;  // device, nf1, nf2, batch, double?, dir
std::map<Key, PlanEntry> g_plans;
std::mutex g_mu;
bool g_setup = false;
}  // namespace

int fft_plan(int device, int nf1, int nf2, int batch, bool dbl, int sign, FftPlan *out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_setup) {
    if (rocfft_setup() != rocfft_status_success) {
      set_error("rocfft_setup failed");
      return kErrFft;
    }
    g_setup = true;
  }
  Key k{device, nf1, nf2, batch, dbl ? 1 : 0, sign};
  auto it = g_plans.find(k);
  if (it == g_plans.end()) {
    PlanEntry e;
    size_t lengths[2] = {(size_t)nf1, (size_t)nf2};  // fastest first
    // rocFFT forward = exp(-i ...), inverse = exp(+i ...), both unnormalised
    rocfft_transform_type tt =
        sign > 0 ? rocfft_transform_type_complex_inverse : rocfft_transform_type_complex_forward;
    rocfft_status st = rocfft_plan_create(&e.plan, rocfft_placement_inplace, tt,
                                          dbl ? rocfft_precision_double : rocfft_precision_single, 2,
                                          lengths, (size_t)batch, nullptr);
    if (st != rocfft_status_success) {
      set_error("rocfft_plan_create(%d x %d, batch %d) failed: %d", nf1, nf2, batch, (int)st);
      return kErrFft;
    }
    rocfft_plan_get_work_buffer_size(e.plan, &e.work);
    if (rocfft_execution_info_create(&e.info) != rocfft_status_success) {
      set_error("rocfft_execution_info_create failed");
      return kErrFft;
    }
    it = g_plans.emplace(k, e).first;
  }
  out->plan = it->second.plan;
  out->info = it->second.info;
  out->work_bytes = it->second.work;
  return kOk;
}

int fft_exec(const FftPlan &p, void *data, void *work, hipStream_t stream) {
  if (rocfft_execution_info_set_stream(p.info, stream) != rocfft_status_success) {
    set_error("rocfft_execution_info_set_stream failed");
    return kErrFft;
  }
  if (p.work_bytes) {
    if (rocfft_execution_info_set_work_buffer(p.info, work, p.work_bytes) != rocfft_status_success) {
      set_error("rocfft_execution_info_set_work_buffer failed");
      return kErrFft;
    }
  }
  void *in[1] = {data};
  rocfft_status st = rocfft_execute(p.plan, in, nullptr, p.info);
  if (st != rocfft_status_success) {
    set_error("rocfft_execute failed: %d", (int)st);
    return kErrFft;
  }
  return kOk;
}

void fft_release_all() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_plans) {
    rocfft_execution_info_destroy(kv.second.info);
    rocfft_plan_destroy(kv.second.plan);
  }
  g_plans.clear();
}

}  // namespace curig
