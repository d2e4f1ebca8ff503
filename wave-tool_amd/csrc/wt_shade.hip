// The water program's fragment stage (assets/shaders/water-grid.frag:70-162) as a
// deferred pass: one thread per covered pixel of a fragment buffer (the interpolated
// varyings and gl_FragCoord.z the rasteriser would hand the shader) reads the depth,
// local reflection / refraction and skybox textures with GL's filtering and writes the
// fragment colour, optionally blended over the colour buffer as the reference's blend
// state does (src/render-engine.cpp:302-303).
//
// Per covered pixel the pass moves 48 B of fragment record, a depth texel, three
// filtered RGBA8 lookups whose texels neighbouring pixels share, and a 16 B colour.
#include <hip/hip_runtime.h>

#include "wt_kernels.hpp"

namespace wt {

// One GL_LINEAR + CLAMP_TO_EDGE lookup of an RGBA8 image: the four texels and the
// weights. u = s * width - 1/2, i0 = floor(u), alpha = frac(u), indices clamped to the
// image (GL 4.6 core spec 8.14.2).
struct Rgba8Tap {
  unsigned t00, t10, t01, t11;
  float a, b;
};

__device__ __forceinline__ void rgba8_fetch(const unsigned* __restrict__ tex, int w, int h, float s, float t,
                                            Rgba8Tap& o) {
  const float u = fmaf(s, (float)w, -0.5f), v = fmaf(t, (float)h, -0.5f);
  const float fu = floorf(u), fv = floorf(v);
  o.a = u - fu;
  o.b = v - fv;
  // clamped as floats first, so a coordinate far outside the image (or NaN, from a
  // zero view vector) converts to an index inside it
  const int iu = (int)fminf(fmaxf(fu, -1.0f), (float)w), iv = (int)fminf(fmaxf(fv, -1.0f), (float)h);
  const int i0 = min(max(iu, 0), w - 1), i1 = min(max(iu + 1, 0), w - 1);
  const int j0 = min(max(iv, 0), h - 1), j1 = min(max(iv + 1, 0), h - 1);
  // sides are at most 32768, so the texel index fits 32 bits
  o.t00 = tex[j0 * w + i0];
  o.t10 = tex[j0 * w + i1];
  o.t01 = tex[j1 * w + i0];
  o.t11 = tex[j1 * w + i1];
}

__device__ __forceinline__ float4 unorm8x4(unsigned v) {
  return make_float4((float)(v & 255u), (float)((v >> 8) & 255u), (float)((v >> 16) & 255u), (float)(v >> 24));
}

// The filtered value, unorm c / 255 per channel.
__device__ __forceinline__ float4 rgba8_eval(const Rgba8Tap& o) {
  const float4 c00 = unorm8x4(o.t00), c10 = unorm8x4(o.t10), c01 = unorm8x4(o.t01), c11 = unorm8x4(o.t11);
  const float w00 = (1.0f - o.a) * (1.0f - o.b), w10 = o.a * (1.0f - o.b), w01 = (1.0f - o.a) * o.b, w11 = o.a * o.b;
  constexpr float k = 1.0f / 255.0f;
  return make_float4(fmaf(w00, c00.x, fmaf(w10, c10.x, fmaf(w01, c01.x, w11 * c11.x))) * k,
                     fmaf(w00, c00.y, fmaf(w10, c10.y, fmaf(w01, c01.y, w11 * c11.y))) * k,
                     fmaf(w00, c00.z, fmaf(w10, c10.z, fmaf(w01, c01.z, w11 * c11.z))) * k,
                     fmaf(w00, c00.w, fmaf(w10, c10.w, fmaf(w01, c01.w, w11 * c11.w))) * k);
}

// The filtered rgb only (the skybox lookup's alpha is never read).
__device__ __forceinline__ float3 rgba8_eval_rgb(const Rgba8Tap& o) {
  const float4 c00 = unorm8x4(o.t00), c10 = unorm8x4(o.t10), c01 = unorm8x4(o.t01), c11 = unorm8x4(o.t11);
  const float w00 = (1.0f - o.a) * (1.0f - o.b), w10 = o.a * (1.0f - o.b), w01 = (1.0f - o.a) * o.b, w11 = o.a * o.b;
  constexpr float k = 1.0f / 255.0f;
  return make_float3(fmaf(w00, c00.x, fmaf(w10, c10.x, fmaf(w01, c01.x, w11 * c11.x))) * k,
                     fmaf(w00, c00.y, fmaf(w10, c10.y, fmaf(w01, c01.y, w11 * c11.y))) * k,
                     fmaf(w00, c00.z, fmaf(w10, c10.z, fmaf(w01, c01.z, w11 * c11.z))) * k);
}

// GLSL requires 2.5 ulp for a / b, 3 ulp for exp2 / log2 and inherits pow's and
// sqrt's precision from them: the hardware reciprocal, reciprocal square root, exp2
// and log2 (about 1 ulp each) are within what the shader itself may compute.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// pow(x, y) for x >= 0 as exp2(y log2 x), with pow(x, 0) = 1.
__device__ __forceinline__ float pow_nonneg(float x, float y) {
  return y == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// Cube-map face selection (GL 4.6 core spec 8.13, table 8.19), faces in GL target order
// +X, -X, +Y, -Y, +Z, -Z; on equal magnitudes x wins over y and y over z. Returns the
// face and its (s, t).
__device__ __forceinline__ int cube_face(float rx, float ry, float rz, float& s, float& t) {
  const float ax = fabsf(rx), ay = fabsf(ry), az = fabsf(rz);
  int face;
  float sc, tc, ma;
  if (ax >= ay && ax >= az) {
    face = rx >= 0.0f ? 0 : 1;
    sc = rx >= 0.0f ? -rz : rz;
    tc = -ry;
    ma = ax;
  } else if (ay >= az) {
    face = ry >= 0.0f ? 2 : 3;
    sc = rx;
    tc = ry >= 0.0f ? rz : -rz;
    ma = ay;
  } else {
    face = rz >= 0.0f ? 4 : 5;
    sc = rz >= 0.0f ? rx : -rx;
    tc = -ry;
    ma = az;
  }
  const float inv = rcp(ma);
  s = 0.5f * fmaf(sc, inv, 1.0f);
  t = 0.5f * fmaf(tc, inv, 1.0f);
  return face;
}

// linearizeDepth, water-grid.frag:66-68 (zfar_m_znear = zFar - zNear in FP32).
__device__ __forceinline__ float linearize_depth(const ShadeArgs& a, float d) {
  return (a.znear * d) * rcp(a.zfar - d * a.zfar_m_znear);
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// GLSL mix(x, y, t) = x (1 - t) + y t
__device__ __forceinline__ float mixf(float x, float y, float t) { return fmaf(y, t, x * (1.0f - t)); }

constexpr int kShadeBlock = 256;

// Workgroups go to the eight XCDs round robin (workgroup b to XCD b mod 8), each with
// its own L2. The texels a row of pixels reads are shared with the rows above and
// below it (bilinear footprints, smooth distortion), so each XCD takes whole chunks of
// kShadeChunk consecutive spans of kShadeBlock pixels (a few rows of the frame), chunks
// dealt to the XCDs in turn: the k-th workgroup of XCD x shades span
// ((k / kShadeChunk) * 8 + x) * kShadeChunk + k mod kShadeChunk. Within a chunk the
// rows' shared texel lines stay in one L2; dealing chunks in turn keeps the XCDs evenly
// loaded when only part of the frame is covered.
constexpr int kXcds = 8;
constexpr int kShadeChunk = 64;

// One thread per pixel; the threads of a wavefront take 64 consecutive pixels of a row,
// so the record loads and colour stores coalesce and the lookups of the local targets
// (sampled near the pixel's own position) touch a few texel lines each. The loads come
// in three rounds, each issued whole before any of its results is used: the coverage
// byte; the fragment record and depth texel; the twelve texels of the three filtered
// lookups.
__global__ __launch_bounds__(kShadeBlock) void water_shade_kernel(ShadeArgs a) {
  const int k = (int)blockIdx.x / kXcds;  // the grid is a multiple of kXcds * kShadeChunk
  const int span = ((k / kShadeChunk) * kXcds + (int)blockIdx.x % kXcds) * kShadeChunk + k % kShadeChunk;
  const int p = span * kShadeBlock + (int)threadIdx.x;
  if (p >= a.count || (a.coverage && !a.coverage[p])) return;

  // the fragment record and the depth texel (water-grid.frag:77, :82)
  const float4* fr = a.frags + 3 * (size_t)p;
  const float4 f0 = fr[0], f1 = fr[1], f2 = fr[2];  // normal.xyz view.x | view.yz yaw.xy | yaw.z ndc0.xy fragZ
  const int row = (int)((unsigned)p / (unsigned)a.width), col = p - row * a.width;
  const float uvx = ((float)col + 0.5f) * a.inv_viewport_w, uvy = ((float)row + 0.5f) * a.inv_viewport_h;  // :77
  // GL_NEAREST + CLAMP_TO_EDGE (clamped as floats before the conversion)
  const int di = (int)fminf(fmaxf(floorf(uvx * (float)a.depth_w), 0.0f), (float)(a.depth_w - 1));
  const int dj = (int)fminf(fmaxf(floorf(uvy * (float)a.depth_h), 0.0f), (float)(a.depth_h - 1));
  const float depth_texel = a.depth[dj * a.depth_w + di];
  float4 dst = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (a.blend) dst = a.colour[p];

  // texture coordinates and the lookups' texel loads
  const float vx = f0.w, vy = f1.x, vz = f1.y;
  const float vv = fmaf(vx, vx, fmaf(vy, vy, vz * vz));
  const float rlen = __builtin_amdgcn_rsqf(vv);
  const float len = vv * rlen;                              // length(viewVecRaw) (:71)
  const float depth_clamped = clamp01(len * a.inv_zfar);   // :72
  const float uv0x = (f2.y + 1.0f) * 0.5f, uv0y = (f2.z + 1.0f) * 0.5f;  // :79
  const float dist = 0.1f * (1.0f - depth_clamped);         // :120-121
  const float yx = mixf(uvx, uv0x, dist), yy = mixf(uvy, uv0y, dist);  // :124-125
  const float nyx = f1.z, nyz = f2.x;  // normalVecInViewSpaceOnlyYaw.x, .z
  Rgba8Tap refl, refr, skyt;
  rgba8_fetch(a.refl, a.refl_w, a.refl_h, clamp01(fmaf(dist, nyx, yx)), clamp01(fmaf(dist, -nyz, yy)), refl);  // :128
  rgba8_fetch(a.refr, a.refr_w, a.refr_h, clamp01(fmaf(dist, -nyx, yx)), clamp01(fmaf(dist, nyz, yy)), refr);  // :129
  // R = reflect(-viewVec, normal) = -v + 2 (n . v) n (:95)
  const float wx = vx * rlen, wy = vy * rlen, wz = vz * rlen;  // viewVec (:73)
  const float nx = f0.x, ny = f0.y, nz = f0.z;
  const float nd = fmaf(nx, wx, fmaf(ny, wy, nz * wz));  // fresnel_cos_theta (:110)
  const float rx = fmaf(2.0f * nd, nx, -wx), ry = fmaf(2.0f * nd, ny, -wy), rz = fmaf(2.0f * nd, nz, -wz);
  {
    float s, t;
    const int face = cube_face(rx, ry, rz, s, t);
    rgba8_fetch(a.sky + (size_t)face * (size_t)a.sky_len * (size_t)a.sky_len, a.sky_len, a.sky_len, s, t, skyt);
  }

  // the colour (water-grid.frag:84-162)
  const float back = linearize_depth(a, depth_texel);  // :82
  const float in_front = back > 0.999f ? 0.0f : 1.0f;  // hack_skybox_in_back (:84)
  const float delta = clamp01(back - linearize_depth(a, f2.w));  // :85-89
  const float sun = a.strength * pow_nonneg(clamp01(fmaf(rx, a.sun[0], fmaf(ry, a.sun[1], rz * a.sun[2]))),
                                            a.shininess);  // :100
  const float x = 1.0f - nd;
  const float x2 = x * x;
  const float fresnel = fmaf(1.0f - 0.02f, x2 * x2 * x, 0.02f);  // :111
  const float3 sky = rgba8_eval_rgb(skyt);
  const float4 lrefl = rgba8_eval(refl), lrefr = rgba8_eval(refr);
  const float tint_f = a.tint_off ? 0.0f : in_front * (1.0f - clamp01(delta * a.inv_tint));  // :138
  const float tint_scale = clamp01(a.sun[1]);                                                // :141
  const float tr = mixf(0.0f, 0.3f, tint_f) * tint_scale, tg = mixf(0.341f, 0.941f, tint_f) * tint_scale,
              tb = mixf(0.482f, 0.903f, tint_f) * tint_scale;  // :136-139
  const float kt = a.clarity * lrefr.w * (1.0f - delta);       // :143
  const float trans_r = mixf(tr, lrefr.x, kt), trans_g = mixf(tg, lrefr.y, kt), trans_b = mixf(tb, lrefr.z, kt);
  const float refl_r = mixf(sky.x + sun, lrefl.x, lrefl.w), refl_g = mixf(sky.y + sun, lrefl.y, lrefl.w),
              refl_b = mixf(sky.z + sun, lrefl.z, lrefl.w);  // :145
  float cr = mixf(trans_r, refl_r, fresnel), cg = mixf(trans_g, refl_g, fresnel),
        cb = mixf(trans_b, refl_b, fresnel);  // :151
  const float alpha = a.soft_off ? 1.0f : 1.0f - in_front * (1.0f - clamp01(delta * a.inv_soft));  // :154-156
  const float fog_a = depth_clamped >= a.fog_far
                          ? a.fog[3]
                          : (depth_clamped <= a.fog_near ? 0.0f
                                                         : a.fog[3] * ((depth_clamped - a.fog_near) * a.inv_fog_range));  // :161
  cr = mixf(cr, a.fog[0], fog_a);  // :162
  cg = mixf(cg, a.fog[1], fog_a);
  cb = mixf(cb, a.fog[2], fog_a);
  float4 out = make_float4(cr, cg, cb, alpha);
  if (a.blend)  // GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA on every channel
    out = make_float4(fmaf(cr, alpha, dst.x * (1.0f - alpha)), fmaf(cg, alpha, dst.y * (1.0f - alpha)),
                      fmaf(cb, alpha, dst.z * (1.0f - alpha)), fmaf(alpha, alpha, dst.w * (1.0f - alpha)));
  // the colour buffer is not read again by this pass: stream it past the caches
  float* o = &a.colour[p].x;
  __builtin_nontemporal_store(out.x, o);
  __builtin_nontemporal_store(out.y, o + 1);
  __builtin_nontemporal_store(out.z, o + 2);
  __builtin_nontemporal_store(out.w, o + 3);
}

hipError_t launch_shade(const ShadeArgs& a, hipStream_t st) {
  if (a.count == 0) return hipSuccess;
  // spans of kShadeBlock pixels, rounded up to whole rounds of chunks; the spans past
  // the frame have no pixels and return at once
  const int per_round = kXcds * kShadeChunk;
  const int spans = (a.count + kShadeBlock - 1) / kShadeBlock;
  const int groups = (spans + per_round - 1) / per_round * per_round;
  hipLaunchKernelGGL(water_shade_kernel, dim3((unsigned)groups), dim3(kShadeBlock), 0, st, a);
  return hipGetLastError();
}

}  // namespace wt
