// Water-grid index buffer on the device: the reference builds it once on the host
// (src/program.cpp:686-717); here it can be written straight into a caller's device
// buffer, e.g. once per patch of a batch with the patch's first vertex as the base.
#include "wt_kernels.hpp"

namespace wt {

constexpr int kIndexThreads = 256;

// One thread per cell (row r = blockIdx.y, column c): two counter-clockwise triangles
// (r,c) (r+1,c+1) (r+1,c) and (r,c) (r,c+1) (r+1,c+1), ids row * n + col + base.
__global__ __launch_bounds__(kIndexThreads) void grid_indices_kernel(uint32_t n, uint32_t base, uint32_t* out) {
  const uint32_t m = n - 1;
  const uint32_t c = blockIdx.x * kIndexThreads + threadIdx.x;
  if (c >= m) return;
  const uint32_t r = blockIdx.y;
  const uint32_t bl = base + r * n + c, br = bl + 1, tl = bl + n, tr = tl + 1;
  uint32_t* o = out + 6 * ((size_t)r * m + c);
  o[0] = bl;
  o[1] = tr;
  o[2] = tl;
  o[3] = bl;
  o[4] = br;
  o[5] = tr;
}

hipError_t launch_grid_indices(uint32_t n, uint32_t base, uint32_t* out, hipStream_t st) {
  const uint32_t m = n - 1;
  const dim3 grid((m + kIndexThreads - 1) / kIndexThreads, m);
  hipLaunchKernelGGL(grid_indices_kernel, grid, dim3(kIndexThreads), 0, st, n, base, out);
  return hipGetLastError();
}

}  // namespace wt
