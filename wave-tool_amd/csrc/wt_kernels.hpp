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

// One patch of a batch on the device: corner algebra precomputed on the host.
struct PatchDesc {
  double bl_x, bl_z, du_x, du_z, dv_x, dv_z, duv_x, duv_z, inv_nm1;
  double time;
  float cam[3];
  float s_h, s_d, bounce, scroll_x, scroll_z;
  int N, W, wave_offset, reserved;
  long long vertex_offset;  // first vertex of the patch in the output arrays
};

// One workgroup's segment group of one patch.
struct BatchUnit {
  int patch;
  int group;
};

struct BatchArgs {
  const PatchDesc* patches;
  const BatchUnit* units;
  int n_units;
  unsigned* counters;  // [0] next unit, [1] workgroups finished
  double time_offset;  // added to every patch's time (frame clock)
  const wt_gerstner_wave* waves;
  const void* texels;
  int tex_w, tex_h, tex_fmt, tex_wrap;
  float* pos;
  float* nrm;
  float* tan;
  float* bit;
};

// Segments a launch shape covers per workgroup for a grid of side N.
long long forward_segment_groups(int N, const LaunchShape& s);
// Launches the persistent batched kernel with `grid` workgroups.
hipError_t launch_forward_batched(const BatchArgs& a, const LaunchShape& s, int grid, hipStream_t st);
// Workgroups of the batched kernel resident per CU for this shape.
int batched_blocks_per_cu(const LaunchShape& s, int tex_mode);

}  // namespace wt
