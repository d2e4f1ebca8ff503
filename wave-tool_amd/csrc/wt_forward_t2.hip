// Plain forward kernel (every launch shape) and batched kernel (every batch shape), texture mode kTexR32Mirror.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t2(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR32Mirror>(p, s, st);
}
hipError_t launch_forward_batched_t2(const BatchArgs& a, int shape, hipStream_t st) {
  return launch_batched_mode<kTexR32Mirror>(a, shape, st);
}
}  // namespace wt
