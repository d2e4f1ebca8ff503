// wave-tool water fragment stage ORACLE — test infrastructure only.
//
// A scalar CPU restatement of the water program's fragment shader
// (assets/shaders/water-grid.frag:66-162) and of the GL texture lookups it makes,
// templated on the scalar type: `double` is the parity oracle, `float` the
// multithreaded CPU baseline timed beside the GPU. Nothing in the shipped library
// links, includes or calls this file; only tests/ and bench.py's cpu_baseline leg
// use it, as the checker or the timed CPU baseline.
//
// The reference has no tests of this shader. The pins are closed forms exercised in
// tests/test_oracle_shade.py: GL texel selection and bilinear weights (affine texel
// images are reproduced exactly inside, clamped at the edges), the cube-map face table
// (axis directions land on face centres, every face and orientation checked against
// brute-force texel-centre directions), linearizeDepth's fixed points 0 -> 0 and
// 1 -> 1, the Fresnel limits (normal incidence 0.02, grazing 1), the fog ramp, the
// soft-edge and tint thresholds, and whole-pixel cases that collapse to one
// textbook expression (constant textures, flat normal, no fog).
#pragma once

#include <cmath>
#include <cstdint>

#include "wt_oracle.hpp"

namespace wto {

// RGBA8 image (GL_RGBA / GL_UNSIGNED_BYTE, unorm c / 255), rows bottom to top as GL
// stores them (row j is t index j).
struct Rgba8Image {
  const uint8_t* texels;
  long long width, height;
};

template <class T> inline V4<T> rgba8Texel(const Rgba8Image& im, long long i, long long j) {
  // CLAMP_TO_EDGE: indices clamped to [0, size - 1]
  i = i < 0 ? 0 : (i >= im.width ? im.width - 1 : i);
  j = j < 0 ? 0 : (j >= im.height ? im.height - 1 : j);
  const uint8_t* p = im.texels + 4 * ((size_t)j * (size_t)im.width + (size_t)i);
  return {T(p[0]) / T(255), T(p[1]) / T(255), T(p[2]) / T(255), T(p[3]) / T(255)};
}

// GL_LINEAR with CLAMP_TO_EDGE (GL 4.6 core spec 8.14.2): u = s * width - 1/2,
// i0 = floor(u), alpha = frac(u), texels i0 and i0 + 1 clamped to the image; the same
// along t. The reflection / refraction targets are set up this way at
// src/render-engine.cpp:146-149, 175-178, the skybox cube map at :93-97.
template <class T> inline V4<T> sampleRgba8Linear(const Rgba8Image& im, T s, T t) {
  const T u = s * T(im.width) - T(0.5), v = t * T(im.height) - T(0.5);
  const T fu = std::floor(u), fv = std::floor(v);
  const long long i0 = (long long)fu, j0 = (long long)fv;
  const T a = u - fu, b = v - fv;
  const V4<T> t00 = rgba8Texel<T>(im, i0, j0), t10 = rgba8Texel<T>(im, i0 + 1, j0);
  const V4<T> t01 = rgba8Texel<T>(im, i0, j0 + 1), t11 = rgba8Texel<T>(im, i0 + 1, j0 + 1);
  return t00 * ((T(1) - a) * (T(1) - b)) + t10 * (a * (T(1) - b)) + t01 * ((T(1) - a) * b) + t11 * (a * b);
}

// GL_NEAREST with CLAMP_TO_EDGE on a one-channel float image: texel floor(s * width),
// clamped (the depth texture, src/render-engine.cpp:131-134).
template <class T> inline T sampleFloatNearest(const float* texels, long long width, long long height, T s, T t) {
  long long i = (long long)std::floor(s * T(width)), j = (long long)std::floor(t * T(height));
  i = i < 0 ? 0 : (i >= width ? width - 1 : i);
  j = j < 0 ? 0 : (j >= height ? height - 1 : j);
  return T(texels[(size_t)j * (size_t)width + (size_t)i]);
}

// Cube-map face selection (GL 4.6 core spec 8.13, table 8.19). Faces in GL target
// order +X, -X, +Y, -Y, +Z, -Z. The major axis is the component of largest magnitude;
// on ties (which the spec leaves to the implementation) x wins over y and y over z.
// (s, t) = ((sc / |ma| + 1) / 2, (tc / |ma| + 1) / 2).
template <class T> inline int cubeFace(const V3<T>& r, T* s, T* t) {
  const T ax = std::fabs(r.x), ay = std::fabs(r.y), az = std::fabs(r.z);
  int face;
  T sc, tc, ma;
  if (ax >= ay && ax >= az) {
    face = r.x >= T(0) ? 0 : 1;
    sc = r.x >= T(0) ? -r.z : r.z;
    tc = -r.y;
    ma = ax;
  } else if (ay >= az) {
    face = r.y >= T(0) ? 2 : 3;
    sc = r.x;
    tc = r.y >= T(0) ? r.z : -r.z;
    ma = ay;
  } else {
    face = r.z >= T(0) ? 4 : 5;
    sc = r.z >= T(0) ? r.x : -r.x;
    tc = -r.y;
    ma = az;
  }
  *s = T(0.5) * (sc / ma + T(1));
  *t = T(0.5) * (tc / ma + T(1));
  return face;
}

// texture(samplerCube, r): face selection, then GL_LINEAR + CLAMP_TO_EDGE inside that
// face (GL_TEXTURE_CUBE_MAP_SEAMLESS is never enabled by the reference). faces holds
// the six length x length RGBA8 images back to back in GL target order.
template <class T> inline V4<T> sampleCube(const uint8_t* faces, long long length, const V3<T>& r) {
  T s, t;
  const int f = cubeFace(r, &s, &t);
  const Rgba8Image im{faces + (size_t)f * (size_t)length * (size_t)length * 4, length, length};
  return sampleRgba8Linear(im, s, t);
}

// The fragment shader's uniforms (water-grid.frag:34-55), set at
// src/render-engine.cpp:1060-1098.
template <class T> struct ShadeUniforms {
  V4<T> fogColourFarAtCurrentTime;
  T fogDepthRadiusFar, fogDepthRadiusNear;
  T softEdgesDeltaDepthThreshold;
  V3<T> sunPosition;
  T sunShininess, sunStrength;
  T tintDeltaDepthThreshold;
  T viewportWidth, viewportHeight;
  T waterClarity;
  T zFar, zNear;
};

struct ShadeTextures {
  const float* depth;  // depthTexture2D
  long long depthWidth, depthHeight;
  Rgba8Image reflections, refractions;  // localReflectionsTexture2D, localRefractionsTexture2D
  const uint8_t* skybox;                // samplerCube skybox, six faces
  long long skyboxLength;
};

// The interpolated inputs of one fragment (water-grid.frag:57-61) and gl_FragCoord.
template <class T> struct Fragment {
  V3<T> normal, viewVecRaw, normalVecInViewSpaceOnlyYaw;
  V2<T> xyPositionNDCSpaceHeight0;
  T fragCoordX, fragCoordY, fragCoordZ;
};

// linearizeDepth, water-grid.frag:66-68.
template <class T> inline T linearizeDepth(const ShadeUniforms<T>& u, T depth) {
  return (u.zNear * depth) / (u.zFar - depth * (u.zFar - u.zNear));
}

// pow(x, 5) of water-grid.frag:111 as a product (x >= 0 for an interpolated unit
// normal, where it equals pow).
template <class T> inline T pow5(T x) { return x * x * x * x * x; }

// main(), water-grid.frag:70-162: the fragment colour before blending.
template <class T> inline V4<T> waterFragment(const ShadeUniforms<T>& u, const ShadeTextures& tx, const Fragment<T>& f) {
  const T viewVecLength = length(f.viewVecRaw);                                // :71
  const T viewVecDepthClamped = clamp(viewVecLength / u.zFar, T(0), T(1));     // :72
  const V3<T> viewVec{f.viewVecRaw.x / viewVecLength, f.viewVecRaw.y / viewVecLength,
                      f.viewVecRaw.z / viewVecLength};  // :73
  const T uvx = f.fragCoordX / u.viewportWidth, uvy = f.fragCoordY / u.viewportHeight;  // :77
  const T uv0x = (f.xyPositionNDCSpaceHeight0.x + T(1)) / T(2);               // :79
  const T uv0y = (f.xyPositionNDCSpaceHeight0.y + T(1)) / T(2);
  const T depthFragBack =
      linearizeDepth(u, sampleFloatNearest<T>(tx.depth, tx.depthWidth, tx.depthHeight, uvx, uvy));  // :82
  const T hackSkyboxInBack = depthFragBack > T(0.999f) ? T(0) : T(1);         // :84
  const T depthFrag = linearizeDepth(u, f.fragCoordZ);                        // :85
  const T deltaDepthClamped = clamp(depthFragBack - depthFrag, T(0), T(1));   // :87-89
  const V3<T> R = reflect(-viewVec, f.normal);                                // :95
  const V4<T> sky = sampleCube<T>(tx.skybox, tx.skyboxLength, R);             // :96
  const T sunDot = clamp(dot(R, u.sunPosition), T(0), T(1));
  const T sun = u.sunStrength * std::pow(sunDot, u.sunShininess);            // :100 (white sun)
  const T fresnelCos = dot(f.normal, viewVec);                                // :110
  const T fresnel = T(0.02f) + (T(1) - T(0.02f)) * pow5(T(1) - fresnelCos);    // :111
  const T distortion = T(0.1f) * (T(1) - viewVecDepthClamped);                 // :118-121 (both 0.1)
  const T yx = uvx * (T(1) - distortion) + uv0x * distortion;                 // :124-125 mix
  const T yy = uvy * (T(1) - distortion) + uv0y * distortion;
  const V3<T>& ny = f.normalVecInViewSpaceOnlyYaw;
  const T reflU = clamp(yx + distortion * ny.x, T(0), T(1));                  // :126, :128
  const T reflV = clamp(yy + distortion * -ny.z, T(0), T(1));
  const T refrU = clamp(yx + distortion * -ny.x, T(0), T(1));                 // :127, :129
  const T refrV = clamp(yy + distortion * ny.z, T(0), T(1));
  const V4<T> localRefl = sampleRgba8Linear<T>(tx.reflections, reflU, reflV);  // :132
  const V4<T> localRefr = sampleRgba8Linear<T>(tx.refractions, refrU, refrV);  // :133
  const V3<T> deep{T(0.0f), T(0.341f), T(0.482f)}, shallow{T(0.3f), T(0.941f), T(0.903f)};  // :136-137
  const T tintFactor = T(0) == u.tintDeltaDepthThreshold
                           ? T(0)
                           : hackSkyboxInBack * (T(1) - clamp(deltaDepthClamped / u.tintDeltaDepthThreshold, T(0), T(1)));
  const V3<T> tintNoon = mix(deep, shallow, tintFactor);                       // :138-139
  const V3<T> tintNow = tintNoon * clamp(u.sunPosition.y, T(0), T(1));         // :141
  const V3<T> transmission =
      mix(tintNow, localRefr.xyz(), u.waterClarity * localRefr.w * (T(1) - deltaDepthClamped));  // :143
  const V3<T> reflection = mix(sky.xyz() + V3<T>{sun, sun, sun}, localRefl.xyz(), localRefl.w);  // :145
  V3<T> rgb = mix(transmission, reflection, fresnel);                          // :151
  const T edgeHardness = T(0) == u.softEdgesDeltaDepthThreshold
                             ? T(1)
                             : T(1) - hackSkyboxInBack * (T(1) - clamp(deltaDepthClamped / u.softEdgesDeltaDepthThreshold,
                                                                        T(0), T(1)));  // :154-156
  const V4<T>& fog = u.fogColourFarAtCurrentTime;                              // :160-161
  const T fogA = viewVecDepthClamped >= u.fogDepthRadiusFar
                     ? fog.w
                     : (viewVecDepthClamped <= u.fogDepthRadiusNear
                            ? T(0)
                            : fog.w * ((viewVecDepthClamped - u.fogDepthRadiusNear) /
                                       (u.fogDepthRadiusFar - u.fogDepthRadiusNear)));
  rgb = mix(rgb, fog.xyz(), fogA);                                             // :162
  return {rgb.x, rgb.y, rgb.z, edgeHardness};
}

// glBlendFunc(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA) (src/render-engine.cpp:302-303) on
// every channel of a floating-point colour buffer (no clamp before blending).
template <class T> inline V4<T> blendOver(const V4<T>& src, const V4<T>& dst) {
  return src * src.w + dst * (T(1) - src.w);
}

}  // namespace wto
