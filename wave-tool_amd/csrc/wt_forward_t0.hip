// Plain forward kernel, every launch shape, texture mode kTexNone.
#include "wt_forward.hpp"

namespace wt {
hipError_t launch_forward_t0(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  return launch_forward_mode<kTexNone>(p, s, st);
}
}  // namespace wt
