// Plain forward kernel (every launch shape) and batched kernel (every batch shape), texture mode kTexR32Repeat.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t1(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR32Repeat>(p, s, st);
}
hipError_t launch_forward_batched_t1(const BatchArgs& a, int shape, hipStream_t st) {
  return launch_batched_mode<kTexR32Repeat>(a, shape, st);
}
}  // namespace wt
