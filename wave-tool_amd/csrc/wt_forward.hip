// Plain forward kernel: the water vertex program (assets/shaders/water-grid.vert:114-193)
// over a gridLength^2 projected grid, with analytic normal / tangent / bitangent.
//
// Work decomposition (gfx950, wave64):
//  * A wavefront owns a "segment": LR = 64/LC grid rows x (LC*K) columns. Lane l sits
//    on row l/LC and column c0 + (l % LC); its k-th vertex is LC*k columns further.
//    Consecutive lanes hold consecutive columns, so every store instruction writes a
//    contiguous run of float3 (coalesced).
//  * Along a row the grid point is affine in the column (bilinear corners), so each
//    wave's phase advances by a constant angle per LC columns. A lane seeds
//    (cos, sin) once per wave from an FP64-reduced phase (native v_sin/v_cos on an
//    argument in [-1/2, 1/2) revolution) and steps K vertices with an FP32 rotation
//    (4 VALU ops per vertex-wave instead of two transcendentals). The rotation angles
//    are computed per (row, wave) in FP64 and rounded once.
//  * Per-wave constants are staged in LDS per chunk of kChunk waves (WaveLds), the
//    per-row rotations in a per-wavefront LDS table.
//  * The 8 running sums per vertex (x/z displacement, y, and the 5 distinct entries
//    of the analytic Jacobian) are accumulated either with VALU FMAs (variant 1) or
//    on the matrix cores with v_mfma_f32_4x4x1_16b_f32 (variant 2): 16 independent
//    4x4 outer products per instruction = (4 vertices x 4 coefficients) per block,
//    cos-weighted sums in one accumulator set and sin-weighted sums in another.
#include <hip/hip_runtime.h>

#include "wt_device.hpp"
#include "wt_kernels.hpp"

namespace wt {

constexpr int kChunk = 64;      // waves staged per LDS pass
constexpr int kWavesPerWG = 4;  // 4 wavefronts of 64 = 256 threads

typedef float v4f __attribute__((ext_vector_type(4)));

// FP64 row geometry: x(col) = xa + col * xs, z(col) = za + col * zs.
struct RowGeom {
  double xa, xs, za, zs;
};
__device__ __forceinline__ RowGeom row_geom(const FwdParams& p, int row) {
  const double v = (double)row * p.inv_nm1;
  RowGeom g;
  g.xa = fma(v, p.dv_x, p.bl_x);
  g.za = fma(v, p.dv_z, p.bl_z);
  g.xs = fma(v, p.duv_x, p.du_x) * p.inv_nm1;
  g.zs = fma(v, p.duv_z, p.du_z) * p.inv_nm1;
  return g;
}

// Stage waves [w0, w0 + nw) of the frame into LDS (every thread of the block calls).
__device__ __forceinline__ void stage_chunk(const FwdParams& p, int w0, int nw, WaveLds* s_wave) {
  for (int t = threadIdx.x; t < nw; t += blockDim.x) stage_wave(p.waves[w0 + t], p.time, &s_wave[t]);
}

// Rotation per LC columns for each (row of this segment, wave of the chunk), in FP64.
template <int LC, int LR>
__device__ __forceinline__ void build_rotations(const FwdParams& p, int row_base, int nw, const WaveLds* s_wave,
                                                float2* s_rot, int lane) {
  for (int e = lane; e < LR * nw; e += 64) {
    const int r = e / nw, i = e - r * nw;
    int row = row_base + r;
    row = row < p.N ? row : p.N - 1;
    const RowGeom g = row_geom(p, row);
    const double step = (s_wave[i].kx * g.xs + s_wave[i].kz * g.zs) * (double)LC;
    double f = step - floor(step);
    const double ang = f * kTwoPi;
    double sn, cs;
    sincos(ang, &sn, &cs);
    s_rot[r * kChunk + i] = make_float2((float)cs, (float)sn);
  }
}

template <int K, int LC, int kVar, int kHm>
__global__ __launch_bounds__(256) void water_forward_kernel(FwdParams p) {
  constexpr int LR = 64 / LC;
  __shared__ WaveLds s_wave[kChunk];
  __shared__ float2 s_rot[kWavesPerWG][LR * kChunk];
  __shared__ float s_xpose[kVar == 2 ? kWavesPerWG : 1][kVar == 2 ? 64 * 9 : 1];

  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const long long gw = (long long)blockIdx.x * kWavesPerWG + wid;
  const int segs_per_rg = (p.N + LC * K - 1) / (LC * K);
  const int n_rg = (p.N + LR - 1) / LR;
  const bool active = gw < (long long)segs_per_rg * n_rg;
  const int seg_id = active ? (int)gw : 0;
  const int rg = seg_id / segs_per_rg, seg = seg_id - rg * segs_per_rg;
  const int row_base = rg * LR;
  const int lr = lane / LC, lc = lane - lr * LC;
  const int row = row_base + lr;
  const int col0 = seg * LC * K + lc;
  const RowGeom g = row_geom(p, row < p.N ? row : p.N - 1);
  const double x0 = fma((double)col0, g.xs, g.xa), z0 = fma((double)col0, g.zs, g.za);

  float acc[K][8];
  v4f accC[kVar == 2 ? K : 1], accS[kVar == 2 ? K : 1];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[j][q] = 0.0f;
  if (kVar == 2) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      accC[j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
      accS[j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }

  for (int w0 = 0; w0 < p.W; w0 += kChunk) {
    const int nw = p.W - w0 < kChunk ? p.W - w0 : kChunk;
    __syncthreads();
    stage_chunk(p, w0, nw, s_wave);
    __syncthreads();
    build_rotations<LC, LR>(p, row_base, nw, s_wave, s_rot[wid], lane);
    __syncthreads();
    const float2* rot_row = &s_rot[wid][lr * kChunk];
    for (int i = 0; i < nw; ++i) {
      const WaveLds& wv = s_wave[i];
      const float2 rot = rot_row[i];
      const float r = reduce_rev(fma(wv.kx, x0, fma(wv.kz, z0, wv.c0)));
      float c = __builtin_amdgcn_cosf(r), s = __builtin_amdgcn_sinf(r);
      if (kVar == 1) {
        const float c0 = wv.cc[0], c1 = wv.cc[1], c2 = wv.cc[2], c3 = wv.cc[3];
        const float s0 = wv.cs[0], s1 = wv.cs[1], s2 = wv.cs[2], s3 = wv.cs[3];
#pragma unroll
        for (int j = 0; j < K; ++j) {
          acc[j][0] = fmaf(c0, c, acc[j][0]);
          acc[j][1] = fmaf(c1, c, acc[j][1]);
          acc[j][2] = fmaf(c2, c, acc[j][2]);
          acc[j][3] = fmaf(c3, c, acc[j][3]);
          acc[j][4] = fmaf(s0, s, acc[j][4]);
          acc[j][5] = fmaf(s1, s, acc[j][5]);
          acc[j][6] = fmaf(s2, s, acc[j][6]);
          acc[j][7] = fmaf(s3, s, acc[j][7]);
          const float cn = fmaf(c, rot.x, -s * rot.y);
          s = fmaf(s, rot.x, c * rot.y);
          c = cn;
        }
      } else {
        const float bc = wv.cc[lane & 3], bs = wv.cs[lane & 3];
#pragma unroll
        for (int j = 0; j < K; ++j) {
          accC[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(c, bc, accC[j], 0, 0, 0);
          accS[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(s, bs, accS[j], 0, 0, 0);
          const float cn = fmaf(c, rot.x, -s * rot.y);
          s = fmaf(s, rot.x, c * rot.y);
          c = cn;
        }
      }
    }
  }

  // MFMA accumulators: register r of lane 4b+k holds output k of the vertex on lane
  // 4b+r. Transpose through a per-wavefront LDS tile (row stride 9 floats).
  float* xp = s_xpose[kVar == 2 ? wid : 0];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (kVar == 2) {
      const int base = lane & ~3, k = lane & 3;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        xp[(base + r) * 9 + k] = accC[j][r];
        xp[(base + r) * 9 + 4 + k] = accS[j][r];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[j][q] = xp[lane * 9 + q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int col = col0 + j * LC;
    if (active && row < p.N && col < p.N) {
      const double x = fma((double)col, g.xs, g.xa), z = fma((double)col, g.zs, g.za);
      vertex_epilogue<kHm>(p.ep, x, z, acc[j], (size_t)row * (size_t)p.N + (size_t)col);
    }
  }
}

// ------------------------------------------------------------------ dispatch
template <int K, int LC, int kVar, int kHm>
static hipError_t launch_t(const FwdParams& p, hipStream_t st) {
  constexpr int LR = 64 / LC;
  const long long segs = (long long)((p.N + LC * K - 1) / (LC * K)) * ((p.N + LR - 1) / LR);
  const long long blocks = (segs + kWavesPerWG - 1) / kWavesPerWG;
  hipLaunchKernelGGL((water_forward_kernel<K, LC, kVar, kHm>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

template <int K, int LC, int kVar>
static hipError_t launch_hm(const FwdParams& p, hipStream_t st) {
  return p.ep.texels ? launch_t<K, LC, kVar, 1>(p, st) : launch_t<K, LC, kVar, 0>(p, st);
}
template <int K, int LC>
static hipError_t launch_var(const FwdParams& p, int var, hipStream_t st) {
  return var == 2 ? launch_hm<K, LC, 2>(p, st) : launch_hm<K, LC, 1>(p, st);
}
template <int K>
static hipError_t launch_lc(const FwdParams& p, int lc, int var, hipStream_t st) {
  switch (lc) {
    case 64: return launch_var<K, 64>(p, var, st);
    case 32: return launch_var<K, 32>(p, var, st);
    case 16: return launch_var<K, 16>(p, var, st);
    default: return launch_var<K, 8>(p, var, st);
  }
}

hipError_t launch_forward(const FwdParams& p, const LaunchShape& s, hipStream_t st) {
  switch (s.steps) {
    case 4: return launch_lc<4>(p, s.lanes_per_row, s.variant, st);
    case 8: return launch_lc<8>(p, s.lanes_per_row, s.variant, st);
    default: return launch_lc<16>(p, s.lanes_per_row, s.variant, st);
  }
}

}  // namespace wt
