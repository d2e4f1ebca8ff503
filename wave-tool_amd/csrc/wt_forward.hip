// Launch dispatch of the forward kernels (templates in wt_forward.hpp): the plain
// kernel per texture mode (one translation unit each, wt_forward_t*.hip) and the
// persistent batched kernel.
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

// The batched kernel is built for two shapes: (K 8, LC 64, S 2) for grids wider than
// 256 and (K 4, LC 32, S 1) otherwise; the batch picks one for all its patches.
template <int K, int LC, int S>
static hipError_t launch_batched_t(const BatchArgs& a, int grid, hipStream_t st) {
  const dim3 g(grid), b(256);
  switch (tex_mode(a.texels, a.tex_fmt, a.tex_wrap)) {
    case kTexR32Repeat: hipLaunchKernelGGL((water_forward_batched_kernel<K, LC, S, kTexR32Repeat>), g, b, 0, st, a); break;
    case kTexR32Mirror: hipLaunchKernelGGL((water_forward_batched_kernel<K, LC, S, kTexR32Mirror>), g, b, 0, st, a); break;
    case kTexR8Repeat: hipLaunchKernelGGL((water_forward_batched_kernel<K, LC, S, kTexR8Repeat>), g, b, 0, st, a); break;
    case kTexR8Mirror: hipLaunchKernelGGL((water_forward_batched_kernel<K, LC, S, kTexR8Mirror>), g, b, 0, st, a); break;
    default: hipLaunchKernelGGL((water_forward_batched_kernel<K, LC, S, kTexNone>), g, b, 0, st, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_forward_batched(const BatchArgs& a, const LaunchShape& s, int grid, hipStream_t st) {
  if (s.steps == 8 && s.lanes_per_row == 64 && s.split == 2) return launch_batched_t<8, 64, 2>(a, grid, st);
  return launch_batched_t<4, 32, 1>(a, grid, st);
}

template <int K, int LC, int S, int kTex>
static int occupancy_kernel() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, water_forward_batched_kernel<K, LC, S, kTex>, 256, 0) !=
      hipSuccess) {
    (void)hipGetLastError();
    return 1;
  }
  return n > 0 ? n : 1;
}

template <int K, int LC, int S>
static int occupancy_t(int tex) {
  switch (tex) {
    case kTexR32Repeat: return occupancy_kernel<K, LC, S, kTexR32Repeat>();
    case kTexR32Mirror: return occupancy_kernel<K, LC, S, kTexR32Mirror>();
    case kTexR8Repeat: return occupancy_kernel<K, LC, S, kTexR8Repeat>();
    case kTexR8Mirror: return occupancy_kernel<K, LC, S, kTexR8Mirror>();
    default: return occupancy_kernel<K, LC, S, kTexNone>();
  }
}

int batched_blocks_per_cu(const LaunchShape& s, int tex) {
  if (s.steps == 8 && s.lanes_per_row == 64 && s.split == 2) return occupancy_t<8, 64, 2>(tex);
  return occupancy_t<4, 32, 1>(tex);
}

}  // namespace wt
