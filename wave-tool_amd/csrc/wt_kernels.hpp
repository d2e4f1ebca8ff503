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

// One patch of a batch on the device: the corner algebra of FwdParams, set on the host
// at batch creation or by the device projector (launch_projector).
struct PatchDesc {
  double bl_x, bl_z, du_x, du_z, dv_x, dv_z, duv_x, duv_z, inv_nm1;
  double time;
  float cam[3];
  float s_h, s_d, bounce, scroll_x, scroll_z;
  int N, W, wave_offset;
  int shape;                // segment shape of this patch in the batched kernel (kBatchShapes)
  long long vertex_offset;  // first vertex of the patch in the output arrays
  int active;               // 0: the patch's camera misses the water; its units write nothing
  int reserved;
};

// Corner algebra of a grid from its corners [4][4] (BL, TL, BR, TR; x and z used).
template <class D>
__host__ __device__ inline void set_corner_algebra(D& d, const double (*c)[4]) {
  d.bl_x = c[0][0];
  d.bl_z = c[0][2];
  d.du_x = c[2][0] - c[0][0];
  d.du_z = c[2][2] - c[0][2];
  d.dv_x = c[1][0] - c[0][0];
  d.dv_z = c[1][2] - c[0][2];
  d.duv_x = c[3][0] - c[1][0] - c[2][0] + c[0][0];
  d.duv_z = c[3][2] - c[1][2] - c[2][2] + c[0][2];
}

// One workgroup's segment group of one patch.
struct BatchUnit {
  int patch;
  int group;
};

struct BatchArgs {
  const PatchDesc* patches;
  const BatchUnit* units;  // every unit of the batch, grouped by segment shape
  long long unit_begin;    // this launch's units: [unit_begin, unit_begin + n_units)
  int n_units;
  double time_offset;  // added to every patch's time (frame clock)
  const wt_gerstner_wave* waves;
  const void* texels;
  int tex_w, tex_h, tex_fmt, tex_wrap;
  float* pos;
  float* nrm;
  float* tan;
  float* bit;
};

// Varyings pass (wt_vertex_varyings): one thread per vertex.
struct VaryingArgs {
  float vp[16];    // viewProjection, column-major
  float vyaw[16];  // viewMatOnlyYaw, column-major
  float cam[3];
  float s_h;
  double hm_off_s, hm_off_t;  // scroll offsets reduced to one wrap period
  const void* texels;
  int tex_w, tex_h, tex_fmt, tex_wrap;
  const float* pos;
  const float* nrm;
  long long count;
  float* clip;
  float* ndc0;
  float* view;
  float* nview;
  float* sample;
};
hipError_t launch_varyings(const VaryingArgs& a, hipStream_t st);

// Segments a launch shape covers per workgroup for a grid of side N.
long long forward_segment_groups(int N, const LaunchShape& s);

// Segment shapes the batched kernel chooses from, per patch (all K = 8, S = 2): 512 x 1
// for wide grids, 256 x 2 and 128 x 4 for narrower ones.
constexpr int kBatchShapes = WT_BATCH_SHAPES;
constexpr LaunchShape kBatchShape[kBatchShapes] = {{2, 8, 64, 64}, {2, 8, 32, 64}, {2, 8, 16, 64}};
// The shape with the fewest padded vertices for a grid of side N (ties to the wider
// shape).
int batch_shape_for(int N);

// Device projector: one thread per camera runs the reference projector in FP64.
// corners (count x 16, row = corner BL, TL, BR, TR) and status (wt_status per camera)
// may be null; with patches set, patch i takes camera i's corner algebra and camera
// position and is switched off when its status is not WT_OK.
struct ProjectorArgs {
  const wt_camera* cameras;
  const float* amplitudes;
  int count;
  double* corners;
  int* status;
  PatchDesc* patches;
};
hipError_t launch_projector(const ProjectorArgs& a, hipStream_t st);

// Triangle index list of an n x n grid (n >= 2), 6 (n - 1)^2 entries, ids offset by base.
hipError_t launch_grid_indices(uint32_t n, uint32_t base, uint32_t* out, hipStream_t st);

// Launches the batched kernel of segment shape `shape` (kBatchShape), one workgroup per
// unit of a.
hipError_t launch_forward_batched(const BatchArgs& a, int shape, hipStream_t st);

// Fragment stage (wt_water_shade): the water program's fragment shader over a
// per-pixel fragment buffer. RGBA8 images are read as one 32-bit word per texel.
struct ShadeArgs {
  float fog[4];  // fogColourFarAtCurrentTime
  float fog_far, fog_near;
  float inv_fog_range;  // 1 / (fogDepthRadiusFar - fogDepthRadiusNear), used only between the radii
  int tint_off, soft_off;   // tintDeltaDepthThreshold / softEdgesDeltaDepthThreshold == 0
  float inv_tint, inv_soft;  // their reciprocals otherwise
  float sun[3];
  float shininess, strength, clarity, zfar, znear;
  float zfar_m_znear, inv_zfar;
  float inv_viewport_w, inv_viewport_h;  // 1 / viewportWidthHeight
  const float* depth;
  int depth_w, depth_h;
  const unsigned* refl;
  int refl_w, refl_h;
  const unsigned* refr;
  int refr_w, refr_h;
  const unsigned* sky;  // six faces, sky_len^2 texels each
  int sky_len;
  const float4* frags;  // three float4 per pixel
  const unsigned char* coverage;
  float4* colour;
  int width, count;  // count = width x height <= 2^30
  int blend;
};
hipError_t launch_shade(const ShadeArgs& a, hipStream_t st);

}  // namespace wt
