// Plain forward kernel (every launch shape) and batched kernel (every batch shape), texture mode kTexR8Mirror.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t4(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR8Mirror>(p, s, st);
}
hipError_t launch_forward_batched_t4(const BatchArgs& a, int shape, hipStream_t st) {
  return launch_batched_mode<kTexR8Mirror>(a, shape, st);
}
}  // namespace wt
