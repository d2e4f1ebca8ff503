// Device-side building blocks shared by the forward, batched and adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wt.h"

namespace wt {

// Non-temporal stores of a float3 (results written once and not read again by the
// writing kernel).
__device__ __forceinline__ void store3_nt(float* o, float x, float y, float z) {
  __builtin_nontemporal_store(x, o);
  __builtin_nontemporal_store(y, o + 1);
  __builtin_nontemporal_store(z, o + 2);
}


constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kInvTwoPi = 0.15915494309189533576888376337251;

// Per-wave constants staged in LDS, shared by every wavefront of a workgroup.
// Phases are carried in revolutions (1 rev = 2 pi rad), the unit of v_sin_f32 /
// v_cos_f32. kx, kz: revolutions per metre along world x and z (w Dx / 2pi,
// w Dz / 2pi); c0: fract((phi t + phi0) / 2pi) reduced in FP64, so the FP32 part of
// the kernel never sees a large phase.
struct WaveLds {
  double kx, kz, c0;
  float cc[4];  // multiply cos: Q A Dx (x disp), Q A Dz (z disp), A w Dx (dy/dx), A w Dz (dy/dz)
  float cs[4];  // multiply sin: A (y disp), Q A w Dx^2, Q A w Dx Dz, Q A w Dz^2
};

// fract of an FP64 phase in revolutions (v_fract_f64), rounded to FP32: the
// argument of v_sin_f32 / v_cos_f32 stays in [0, 1] revolution.
__device__ __forceinline__ float reduce_rev(double r) { return (float)__builtin_amdgcn_fract(r); }

// cos and sin of 2 pi f for an FP64 fraction f in [0, 1): v_cos_f32 / v_sin_f32 of the
// FP32-rounded fraction plus the first-order term of the rounding residual d (radians,
// |d| < 4e-7), so the result has the accuracy of the hardware functions (about 1e-7
// absolute) instead of that of an FP32 argument.
__device__ __forceinline__ void sincos_rev(double f, float* c, float* s) {
  const float a = (float)f;
  const float d = (float)((f - (double)a) * kTwoPi);
  const float ca = __builtin_amdgcn_cosf(a), sa = __builtin_amdgcn_sinf(a);
  *c = fmaf(-d, sa, ca);
  *s = fmaf(d, ca, sa);
}

__device__ __forceinline__ void stage_wave(const wt_gerstner_wave& g, double t, WaveLds* o) {
  const double w = (double)g.frequency_w, dx = (double)g.xzDirection_D[0], dz = (double)g.xzDirection_D[1];
  o->kx = w * dx * kInvTwoPi;
  o->kz = w * dz * kInvTwoPi;
  const double ph = ((double)g.phaseConstant_phi * t + (double)g.phaseOffset_phi0) * kInvTwoPi;
  o->c0 = __builtin_amdgcn_fract(ph);
  const float A = g.amplitude_A, Q = g.steepness_Q_i, wf = g.frequency_w;
  const float Dx = g.xzDirection_D[0], Dz = g.xzDirection_D[1];
  const float qa = Q * A, aw = A * wf, qaw = qa * wf;
  o->cc[0] = qa * Dx;
  o->cc[1] = qa * Dz;
  o->cc[2] = aw * Dx;
  o->cc[3] = aw * Dz;
  o->cs[0] = A;
  o->cs[1] = qaw * Dx * Dx;
  o->cs[2] = qaw * Dx * Dz;
  o->cs[3] = qaw * Dz * Dz;
}

// Entry i of a table that no thread writes during the launch, read through the constant
// address space: a wave-uniform index then loads it with scalar loads into SGPRs.
template <class T>
__device__ __forceinline__ T load_uniform(const T* table, long long i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T ConstT;
  return ((ConstT*)table)[i];
#else
  return table[i];  // the host pass only parses device code
#endif
}

// Heightmap sampling mode, a compile-time parameter of the kernels: none, or texel
// format x wrap mode (src/texture.cpp:54-72 sets the reference's to 8-bit, MIRRORED_REPEAT).
enum TexMode : int { kTexNone = 0, kTexR32Repeat = 1, kTexR32Mirror = 2, kTexR8Repeat = 3, kTexR8Mirror = 4 };
constexpr bool tex_r8(int m) { return m == kTexR8Repeat || m == kTexR8Mirror; }
constexpr bool tex_mirror(int m) { return m == kTexR32Mirror || m == kTexR8Mirror; }
__host__ __device__ inline int tex_mode(const void* texels, int fmt, int wrap) {
  if (!texels) return kTexNone;
  const bool r8 = fmt == WT_TEXEL_R8_UNORM, mirror = wrap == WT_WRAP_MIRRORED_REPEAT;
  return r8 ? (mirror ? kTexR8Mirror : kTexR8Repeat) : (mirror ? kTexR32Mirror : kTexR32Repeat);
}

// Bilinear heightmap tap with GL texel-centre and wrap rules, split in two so a lane can
// issue the texel loads of all its vertices before it needs any of them.
// heightmap_fetch: (s, t) are texture coordinates already reduced to one wrap period
// ([0, 1) for REPEAT, [0, 2) for MIRRORED_REPEAT); index arithmetic is clamped so any
// input stays in bounds. heightmap_eval: value and dH/ds, dH/dt of the same cell.
struct HmTap {
  float t00, t10, t01, t11, a, b;
};

// Texel index wrap for the two taps of a footprint (GL REPEAT / MIRRORED_REPEAT over
// n texels; one period is n, or 2n for the mirror). wrap_lo takes i in [-1, period - 1]
// (only -1 is out of range), wrap_hi takes i in [0, period] (only period is): an unsigned
// min picks the in-range one of i and i -/+ period. The mirror then folds [n, 2n) back
// onto [0, n) as 2n - 1 - i.
template <int M>
__device__ __forceinline__ uint32_t wrap_fold(uint32_t i, uint32_t n) {
  return tex_mirror(M) ? min(i, 2u * n - 1u - i) : i;
}
template <int M>
__device__ __forceinline__ uint32_t wrap_lo(int i, uint32_t n) {
  const uint32_t period = tex_mirror(M) ? 2u * n : n;
  return wrap_fold<M>(min((uint32_t)i, (uint32_t)i + period), n);
}
template <int M>
__device__ __forceinline__ uint32_t wrap_hi(int i, uint32_t n) {
  const uint32_t period = tex_mirror(M) ? 2u * n : n;
  return wrap_fold<M>(min((uint32_t)i, (uint32_t)i - period), n);
}

// One texel at a 32-bit element offset from the (wave-uniform) texel base: the address
// is the base plus a zero-extended 32-bit byte offset (a heightmap holds at most 2^28
// texels, 1 GiB in R32F), so no 64-bit address arithmetic per tap.
template <int M>
__device__ __forceinline__ float load_texel(const void* __restrict__ texels, uint32_t idx) {
  const char* base = (const char*)texels;
  if (tex_r8(M)) return (float)*(const uint8_t*)(base + idx);
  return *(const float*)(base + (idx << 2));
}

template <int M>
__device__ __forceinline__ HmTap heightmap_fetch(const void* __restrict__ texels, int W, int H, float s, float t) {
  const int period_w = tex_mirror(M) ? 2 * W : W;
  const int period_h = tex_mirror(M) ? 2 * H : H;
  const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
  const float fu = floorf(u), fv = floorf(v);
  HmTap h;
  h.a = u - fu;
  h.b = v - fv;
  // clamp to [-1, period - 1] with one v_med3_f32; with a NaN operand it returns the
  // min3 of the three, which is -1 here, so a NaN coordinate lands on texel -1
  const int i0 = (int)__builtin_amdgcn_fmed3f(fu, -1.0f, (float)(period_w - 1));
  const int j0 = (int)__builtin_amdgcn_fmed3f(fv, -1.0f, (float)(period_h - 1));
  const uint32_t c0 = wrap_lo<M>(i0, (uint32_t)W), c1 = wrap_hi<M>(i0 + 1, (uint32_t)W);
  // rows times the width in 24-bit multiplies (both factors below 2^14)
  const uint32_t r0 = __umul24(wrap_lo<M>(j0, (uint32_t)H), (uint32_t)W);
  const uint32_t r1 = __umul24(wrap_hi<M>(j0 + 1, (uint32_t)H), (uint32_t)W);
  h.t00 = load_texel<M>(texels, r0 + c0);
  h.t10 = load_texel<M>(texels, r0 + c1);
  h.t01 = load_texel<M>(texels, r1 + c0);
  h.t11 = load_texel<M>(texels, r1 + c1);
  return h;
}

template <int M>
__device__ __forceinline__ float heightmap_eval(HmTap h, int W, int H, float* dds, float* ddt) {
  if (tex_r8(M)) {  // unorm8: c / 255
    const float k = 1.0f / 255.0f;
    h.t00 *= k; h.t10 *= k; h.t01 *= k; h.t11 *= k;
  }
  const float lo = fmaf(h.a, h.t10 - h.t00, h.t00), hi = fmaf(h.a, h.t11 - h.t01, h.t01);
  *dds = (float)W * fmaf(h.b, (h.t11 - h.t01) - (h.t10 - h.t00), h.t10 - h.t00);
  *ddt = (float)H * fmaf(h.a, (h.t11 - h.t10) - (h.t01 - h.t00), h.t01 - h.t00);
  return fmaf(h.b, hi - lo, lo);
}

// Per-vertex epilogue, in two phases so the texel loads of all of a lane's vertices are
// in flight together: vertex_fetch forms the displaced point and issues the heightmap
// loads; vertex_finish adds the bump at the displaced point and the bounce, forms
// N / T / B and stores.
struct EpilogueArgs {
  float cam[3];
  float s_h, s_d, bounce;
  float two_sd, bump_k;       // 2 s_d and 2 s_d s_h, formed once per launch (2.0f * s_d * s_h)
  double hm_off_s, hm_off_t;  // texture-space scroll offsets, already reduced to one period
  const void* texels;
  int tex_w, tex_h, tex_fmt, tex_wrap;
  float* pos;
  float* nrm;
  float* tan;
  float* bit;
};

struct VertexPending {
  float px, pz;
  HmTap tap;
};

// acc: X, Z, Cx, Cz (cos sums) and Y, Sxx, Sxz, Szz (sin sums).
template <int M>
__device__ __forceinline__ VertexPending vertex_fetch(const EpilogueArgs& e, double x, double z, const float acc[8]) {
  VertexPending v;
  const double pxd = x + (double)acc[0], pzd = z + (double)acc[1];
  v.px = (float)pxd;
  v.pz = (float)pzd;
  if constexpr (M != kTexNone) {
    // water-grid.vert:103-110: sample at s_h * (x, -z) of the displaced point. The
    // texture coordinate is formed from the FP64 displaced point and reduced to one
    // wrap period in FP64, so its FP32 rounding is relative to a number below 2.
    double sd = fma((double)e.s_h, pxd, e.hm_off_s), td = fma(-(double)e.s_h, pzd, e.hm_off_t);
    if (tex_mirror(M)) {
      sd -= 2.0 * floor(0.5 * sd);
      td -= 2.0 * floor(0.5 * td);
    } else {
      sd = __builtin_amdgcn_fract(sd);
      td = __builtin_amdgcn_fract(td);
    }
    v.tap = heightmap_fetch<M>(e.texels, e.tex_w, e.tex_h, (float)sd, (float)td);
  }
  return v;
}

// The output rows of one grid row: the four arrays advanced to the row's first vertex
// (64-bit arithmetic once per row; each vertex then adds a 32-bit column offset).
struct RowOut {
  float* pos;
  float* nrm;
  float* tan;
  float* bit;
};
__device__ __forceinline__ float* row_at(float* row, uint32_t byte_offset) {
  return (float*)((char*)row + byte_offset);
}
__device__ __forceinline__ RowOut row_out(const EpilogueArgs& e, size_t row_first_vertex) {
  const size_t o = 3 * row_first_vertex;
  return RowOut{e.pos + o, e.nrm + o, e.tan ? e.tan + o : nullptr, e.bit ? e.bit + o : nullptr};
}

template <int M>
__device__ __forceinline__ void vertex_finish(const EpilogueArgs& e, const VertexPending& v, const float acc[8],
                                              const RowOut& ro, uint32_t col) {
  const float Cx = acc[2], Cz = acc[3];
  const float Y = acc[4], Sxx = acc[5], Sxz = acc[6], Szz = acc[7];
  const float px = v.px, pz = v.pz;
  float py = Y;
  // dP/dx and dP/dz of the Gerstner surface (no reference counterpart: the
  // reference differences neighbours, water-grid.vert:164-187)
  float tx = 1.0f - Sxx, ty = Cx, tz = -Sxz;
  float bx = -Sxz, by = Cz, bz = 1.0f - Szz;
  if constexpr (M != kTexNone) {
    float dds, ddt;
    const float r = heightmap_eval<M>(v.tap, e.tex_w, e.tex_h, &dds, &ddt);
    py += e.two_sd * (r - 0.5f);  // water-grid.vert:108-110: s_d * 2 (r - 0.5)
    const float k = e.bump_k;
    const float hpx = k * dds, hpz = -k * ddt;  // d(bump)/dPx, d(bump)/dPz
    ty += hpx * tx + hpz * tz;
    by += hpx * bx + hpz * bz;
  }
  py += e.bounce;  // water-grid.vert:147
  // N = normalize(B x T): y-up for the undisplaced plane
  float nx = by * tz - bz * ty, ny = bz * tx - bx * tz, nz = bx * ty - by * tx;
  const float inv_n = __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
  nx *= inv_n; ny *= inv_n; nz *= inv_n;
  // flip toward the camera (water-grid.vert:190)
  const float vx = e.cam[0] - px, vy = e.cam[1] - py, vz = e.cam[2] - pz;
  if (nx * vx + ny * vy + nz * vz < 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
  // the outputs are written once per frame and not read back by this kernel: stream
  // them past the caches (4% faster on the 1024^2 x 64-wave frame)
  // byte offset within the row as a 32-bit value (a row is at most 32768 x 12 bytes), so
  // each store adds it to the row pointer without 64-bit shifts; col < 2^24, so the
  // full-rate 24-bit multiply (a 32-bit one issues at quarter rate)
  const uint32_t b = __umul24(12u, col);
  store3_nt(row_at(ro.pos, b), px, py, pz);
  store3_nt(row_at(ro.nrm, b), nx, ny, nz);
  if (ro.tan) {
    const float it = __builtin_amdgcn_rsqf(tx * tx + ty * ty + tz * tz);
    store3_nt(row_at(ro.tan, b), tx * it, ty * it, tz * it);
  }
  if (ro.bit) {
    const float ib = __builtin_amdgcn_rsqf(bx * bx + by * by + bz * bz);
    store3_nt(row_at(ro.bit, b), bx * ib, by * ib, bz * ib);
  }
}

}  // namespace wt
