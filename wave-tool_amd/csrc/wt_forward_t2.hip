// Plain forward kernel, every launch shape, texture mode kTexR32Mirror.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t2(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexR32Mirror>(p, s, st);
}
}  // namespace wt
