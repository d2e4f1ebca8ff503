// Kernel argument blocks and launchers shared between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "wt_device.hpp"

namespace wt {

// Plain forward: one frame of one grid. Corner algebra in FP64:
// P(u, v) = BL + u (BR - BL) + v (TL - BL) + u v (TR - TL - BR + BL), the expanded form of
// mix(mix(BL, BR, u), mix(TL, TR, u), v) (water-grid.vert:87-91).
struct FwdParams {
  double bl_x, bl_z;
  double du_x, du_z;    // BR - BL
  double dv_x, dv_z;    // TL - BL
  double duv_x, duv_z;  // TR - TL - BR + BL
  double inv_nm1;       // 1 / (gridLength - 1)
  double time;
  int N, W;
  const wt_gerstner_wave* waves;
  EpilogueArgs ep;
};

struct LaunchShape {
  int split;          // S: wavefronts sharing a block of vertices, 1 or 2
  int steps;          // K: 4, 8, 16
  int lanes_per_row;  // LC: 64, 32, 16, 8
  int wave_chunk;     // waves per LDS pass (64 in this build)
};

hipError_t launch_forward(const FwdParams& p, const LaunchShape& s, hipStream_t st);

}  // namespace wt
