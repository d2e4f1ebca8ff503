// Device projector: the reference's per-frame projector (src/render-engine.cpp:797-1042)
// for many cameras at once, one thread per camera, in FP64. It writes corners, a
// status per camera, and/or re-aims the patches of a batch in place so a captured
// frame loop recomputes every patch's grid without a host round trip.
#include "wt_kernels.hpp"
#include "wt_projector.hpp"

namespace wt {

constexpr int kProjectorThreads = 64;

__global__ __launch_bounds__(kProjectorThreads) void projector_kernel(ProjectorArgs a) {
  const int i = (int)blockIdx.x * kProjectorThreads + (int)threadIdx.x;
  if (i >= a.count) return;
  const wt_camera cam = a.cameras[i];
  const float amp = a.amplitudes[i];
  double c[4][4];
  int st = WT_ERR_INVALID_ARG;
  if (proj::camera_valid(cam) && __builtin_isfinite(amp) && amp >= 0.0f)
    st = proj::projector_corners(proj::camera_from(cam), (double)amp, c);
  if (a.corners) {
    double* o = a.corners + 16 * (size_t)i;
    for (int k = 0; k < 16; ++k) o[k] = st == WT_OK ? c[k / 4][k % 4] : __builtin_nan("");
  }
  if (a.status) a.status[i] = st;
  if (a.patches) {
    PatchDesc& d = a.patches[i];
    if (st == WT_OK) {
      set_corner_algebra(d, c);
      d.cam[0] = cam.position[0];
      d.cam[1] = cam.position[1];
      d.cam[2] = cam.position[2];
    }
    d.active = st == WT_OK;
  }
}

hipError_t launch_projector(const ProjectorArgs& a, hipStream_t st) {
  if (a.count <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((a.count + kProjectorThreads - 1) / kProjectorThreads);
  hipLaunchKernelGGL(projector_kernel, dim3(blocks), dim3(kProjectorThreads), 0, st, a);
  return hipGetLastError();
}

}  // namespace wt
