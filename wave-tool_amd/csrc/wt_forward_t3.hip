// Plain forward kernel (every launch shape) and batched kernel (every batch shape), texture mode kTexR8Repeat.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t3(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR8Repeat>(p, s, st);
}
hipError_t launch_forward_batched_t3(const BatchArgs& a, int shape, hipStream_t st) {
  return launch_batched_mode<kTexR8Repeat>(a, shape, st);
}
}  // namespace wt
