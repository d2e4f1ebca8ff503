// Launch dispatch of the forward kernels (templates in wt_forward.hpp): the plain and
// batched kernels per texture mode (one translation unit each, wt_forward_t*.hip).
#include "wt_forward.hpp"

namespace wt {

hipError_t launch_forward_t0(const FwdParams& p, const LaunchShape& s, hipStream_t st);
hipError_t launch_forward_t1(const FwdParams& p, const LaunchShape& s, hipStream_t st);
hipError_t launch_forward_t2(const FwdParams& p, const LaunchShape& s, hipStream_t st);
hipError_t launch_forward_t3(const FwdParams& p, const LaunchShape& s, hipStream_t st);
hipError_t launch_forward_t4(const FwdParams& p, const LaunchShape& s, hipStream_t st);

hipError_t launch_forward(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  switch (tex_mode(p.ep.texels, p.ep.tex_fmt, p.ep.tex_wrap)) {
    case kTexR32Repeat: return launch_forward_t1(p, s, st);
    case kTexR32Mirror: return launch_forward_t2(p, s, st);
    case kTexR8Repeat: return launch_forward_t3(p, s, st);
    case kTexR8Mirror: return launch_forward_t4(p, s, st);
    default: return launch_forward_t0(p, s, st);
  }
}

long long forward_segment_groups(int N, const LaunchShape& s) {
  const int LC = s.lanes_per_row, K = s.steps, LR = 64 / LC;
  const long long segs = (long long)((N + LC * K - 1) / (LC * K)) * ((N + LR - 1) / LR);
  const int per_wg = kWavesPerWG / s.split;
  return (segs + per_wg - 1) / per_wg;
}

int batch_shape_for(int N) {
  // padded vertices per shape (the shapes share K, so the seeds per vertex are equal);
  // the first, widest shape wins ties
  long long best = 0;
  int pick = 0;
  for (int i = 0; i < kBatchShapes; ++i) {
    const LaunchShape& s = kBatchShape[i];
    const long long cols = s.lanes_per_row * s.steps, rows = 64 / s.lanes_per_row;
    const long long padded = (N + cols - 1) / cols * cols * ((N + rows - 1) / rows * rows);
    if (i == 0 || padded < best) {
      best = padded;
      pick = i;
    }
  }
  return pick;
}

hipError_t launch_forward_batched_t0(const BatchArgs& a, int shape, hipStream_t st);
hipError_t launch_forward_batched_t1(const BatchArgs& a, int shape, hipStream_t st);
hipError_t launch_forward_batched_t2(const BatchArgs& a, int shape, hipStream_t st);
hipError_t launch_forward_batched_t3(const BatchArgs& a, int shape, hipStream_t st);
hipError_t launch_forward_batched_t4(const BatchArgs& a, int shape, hipStream_t st);

hipError_t launch_forward_batched(const BatchArgs& a, int shape, hipStream_t st) {
  if (a.n_units <= 0) return hipSuccess;
  switch (tex_mode(a.texels, a.tex_fmt, a.tex_wrap)) {
    case kTexR32Repeat: return launch_forward_batched_t1(a, shape, st);
    case kTexR32Mirror: return launch_forward_batched_t2(a, shape, st);
    case kTexR8Repeat: return launch_forward_batched_t3(a, shape, st);
    case kTexR8Mirror: return launch_forward_batched_t4(a, shape, st);
    default: return launch_forward_batched_t0(a, shape, st);
  }
}

}  // namespace wt
