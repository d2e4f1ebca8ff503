// The water program's varyings (assets/shaders/water-grid.vert:66-70, 141-158, 193) from
// the displaced positions and normals: clip-space position, NDC xy of the height-0
// point, the view vector, the normal in the yaw-only view space and the heightmap
// sample. Every input is read once and every output written once, so the pass is bound
// by HBM (24 B read + up to 52 B written per vertex).
#include <hip/hip_runtime.h>

#include "wt_device.hpp"
#include "wt_kernels.hpp"

namespace wt {

// M [4x4 column-major] * (x, y, z, w)
__device__ __forceinline__ float4 mat4_mul(const float* m, float x, float y, float z, float w) {
  float4 r;
  r.x = fmaf(m[0], x, fmaf(m[4], y, fmaf(m[8], z, m[12] * w)));
  r.y = fmaf(m[1], x, fmaf(m[5], y, fmaf(m[9], z, m[13] * w)));
  r.z = fmaf(m[2], x, fmaf(m[6], y, fmaf(m[10], z, m[14] * w)));
  r.w = fmaf(m[3], x, fmaf(m[7], y, fmaf(m[11], z, m[15] * w)));
  return r;
}

// Vertices per thread, kVaryingsPerThread apart by the grid size so every access
// instruction stays coalesced. A thread issues all its input loads, then all its texel
// loads, before it uses any of them, so each wavefront keeps several round trips to
// memory in flight.
constexpr int kVaryingsPerThread = 2;

template <int kTex>
__global__ __launch_bounds__(256) void vertex_varyings_kernel(VaryingArgs a) {
  constexpr int R = kVaryingsPerThread;
  const long long T = (long long)gridDim.x * blockDim.x;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float px[R], py[R], pz[R], qx[R], qy[R], qz[R];
  HmTap tap[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long long v = t + r * T;
    if (v < a.count) {
      const float* P = a.pos + 3 * v;
      px[r] = P[0];
      py[r] = P[1];
      pz[r] = P[2];
      if (a.nview) {
        const float* N = a.nrm + 3 * v;
        qx[r] = N[0];
        qy[r] = N[1];
        qz[r] = N[2];
      }
    }
  }
  if constexpr (kTex != kTexNone) {
    if (a.sample) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long long v = t + r * T;
        if (v < a.count) {  // water-grid.vert:141 at the displaced point (the bump only moves y)
          double sd = fma((double)a.s_h, (double)px[r], a.hm_off_s), td = fma(-(double)a.s_h, (double)pz[r], a.hm_off_t);
          if (tex_mirror(kTex)) {
            sd -= 2.0 * floor(0.5 * sd);
            td -= 2.0 * floor(0.5 * td);
          } else {
            sd -= floor(sd);
            td -= floor(td);
          }
          tap[r] = heightmap_fetch<kTex>(a.texels, a.tex_w, a.tex_h, (float)sd, (float)td);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long long v = t + r * T;
    if (v >= a.count) continue;
    // every output streams past the caches (non-temporal stores: 16% faster here)
    if (a.clip) {  // water-grid.vert:150
      const float4 c = mat4_mul(a.vp, px[r], py[r], pz[r], 1.0f);
      float* o = a.clip + 4 * v;
      store3_nt(o, c.x, c.y, c.z);
      __builtin_nontemporal_store(c.w, o + 3);
    }
    if (a.ndc0) {  // water-grid.vert:153-154
      const float4 c = mat4_mul(a.vp, px[r], 0.0f, pz[r], 1.0f);
      float* o = a.ndc0 + 2 * v;
      __builtin_nontemporal_store(c.x / c.w, o);
      __builtin_nontemporal_store(c.y / c.w, o + 1);
    }
    if (a.view)  // water-grid.vert:158
      store3_nt(a.view + 3 * v, a.cam[0] - px[r], a.cam[1] - py[r], a.cam[2] - pz[r]);
    if (a.nview) {  // water-grid.vert:193
      const float4 n = mat4_mul(a.vyaw, qx[r], qy[r], qz[r], 0.0f);
      const float inv = __builtin_amdgcn_rsqf(n.x * n.x + n.y * n.y + n.z * n.z);
      store3_nt(a.nview + 3 * v, n.x * inv, n.y * inv, n.z * inv);
    }
    if constexpr (kTex != kTexNone) {
      if (a.sample) {
        float dds, ddt;
        __builtin_nontemporal_store(heightmap_eval<kTex>(tap[r], a.tex_w, a.tex_h, &dds, &ddt), a.sample + v);
      }
    }
  }
}

hipError_t launch_varyings(const VaryingArgs& a, hipStream_t st) {
  if (a.count == 0) return hipSuccess;
  const long long threads = (a.count + kVaryingsPerThread - 1) / kVaryingsPerThread;
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  switch (tex_mode(a.texels, a.tex_fmt, a.tex_wrap)) {
    case kTexR32Repeat: hipLaunchKernelGGL((vertex_varyings_kernel<kTexR32Repeat>), grid, block, 0, st, a); break;
    case kTexR32Mirror: hipLaunchKernelGGL((vertex_varyings_kernel<kTexR32Mirror>), grid, block, 0, st, a); break;
    case kTexR8Repeat: hipLaunchKernelGGL((vertex_varyings_kernel<kTexR8Repeat>), grid, block, 0, st, a); break;
    case kTexR8Mirror: hipLaunchKernelGGL((vertex_varyings_kernel<kTexR8Mirror>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((vertex_varyings_kernel<kTexNone>), grid, block, 0, st, a); break;
  }
  return hipGetLastError();
}

}  // namespace wt
