// Plain forward kernel, every launch shape, texture mode kTexR8Mirror.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t4(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR8Mirror>(p, s, st);
}
}  // namespace wt
