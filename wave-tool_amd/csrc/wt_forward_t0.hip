// Plain forward kernel (every launch shape) and batched kernel (every batch shape), texture mode kTexNone.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t0(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexNone>(p, s, st);
}
hipError_t launch_forward_batched_t0(const BatchArgs& a, int shape, hipStream_t st) {
  return launch_batched_mode<kTexNone>(a, shape, st);
}
}  // namespace wt
