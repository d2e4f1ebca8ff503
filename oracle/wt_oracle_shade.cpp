// C entry points of the fragment-stage ORACLE (test infrastructure only; see
// wt_oracle_shade.hpp). Loaded through ctypes by tests/ and bench.py's cpu_baseline
// leg. The double instantiation is the parity oracle; the float instantiation run
// over std::threads is the CPU baseline.
#include "wt_oracle_shade.hpp"

#include <algorithm>
#include <thread>
#include <vector>

extern "C" {

typedef struct {
  double fog[4];  // fogColourFarAtCurrentTime
  double fog_far, fog_near, soft_edges, sun[3], shininess, strength, tint, viewport[2], clarity, zfar, znear;
  const float* depth;
  int32_t depth_w, depth_h;
  const uint8_t* refl;
  int32_t refl_w, refl_h;
  const uint8_t* refr;
  int32_t refr_w, refr_h;
  const uint8_t* sky;
  int32_t sky_len;
} wto_shade;

}  // extern "C"

namespace {

template <class T> wto::ShadeUniforms<T> uniforms_of(const wto_shade* s) {
  wto::ShadeUniforms<T> u;
  u.fogColourFarAtCurrentTime = {T(s->fog[0]), T(s->fog[1]), T(s->fog[2]), T(s->fog[3])};
  u.fogDepthRadiusFar = T(s->fog_far);
  u.fogDepthRadiusNear = T(s->fog_near);
  u.softEdgesDeltaDepthThreshold = T(s->soft_edges);
  u.sunPosition = {T(s->sun[0]), T(s->sun[1]), T(s->sun[2])};
  u.sunShininess = T(s->shininess);
  u.sunStrength = T(s->strength);
  u.tintDeltaDepthThreshold = T(s->tint);
  u.viewportWidth = T(s->viewport[0]);
  u.viewportHeight = T(s->viewport[1]);
  u.waterClarity = T(s->clarity);
  u.zFar = T(s->zfar);
  u.zNear = T(s->znear);
  return u;
}

wto::ShadeTextures textures_of(const wto_shade* s) {
  wto::ShadeTextures t;
  t.depth = s->depth;
  t.depthWidth = s->depth_w;
  t.depthHeight = s->depth_h;
  t.reflections = {s->refl, s->refl_w, s->refl_h};
  t.refractions = {s->refr, s->refr_w, s->refr_h};
  t.skybox = s->sky;
  t.skyboxLength = s->sky_len;
  return t;
}

// Pixels [first, first + count) of a width-wide frame; fragments hold 12 floats per
// pixel {normal, viewVecRaw, normalVecInViewSpaceOnlyYaw, xyPositionNDCSpaceHeight0,
// gl_FragCoord.z}; gl_FragCoord.xy is the pixel centre. Uncovered pixels are left as
// they are; covered ones get the fragment colour, or with dst the blend over dst.
template <class T, class O>
int run_shade(const wto_shade* s, const float* frags, const uint8_t* coverage, long long width, long long first,
              long long count, O* out, const O* dst, int nthreads) {
  if (!s || !frags || !out || width <= 0) return 1;
  const wto::ShadeUniforms<T> u = uniforms_of<T>(s);
  const wto::ShadeTextures tx = textures_of(s);
  auto work = [&](long long lo, long long hi) {
    for (long long p = lo; p < hi; ++p) {
      if (coverage && !coverage[p]) continue;
      const float* r = frags + 12 * p;
      wto::Fragment<T> f;
      f.normal = {T(r[0]), T(r[1]), T(r[2])};
      f.viewVecRaw = {T(r[3]), T(r[4]), T(r[5])};
      f.normalVecInViewSpaceOnlyYaw = {T(r[6]), T(r[7]), T(r[8])};
      f.xyPositionNDCSpaceHeight0 = {T(r[9]), T(r[10])};
      f.fragCoordX = T(p % width) + T(0.5);
      f.fragCoordY = T(p / width) + T(0.5);
      f.fragCoordZ = T(r[11]);
      wto::V4<T> c = wto::waterFragment(u, tx, f);
      if (dst) c = wto::blendOver(c, wto::V4<T>{T(dst[4 * p]), T(dst[4 * p + 1]), T(dst[4 * p + 2]), T(dst[4 * p + 3])});
      O* o = out + 4 * p;
      o[0] = O(c.x);
      o[1] = O(c.y);
      o[2] = O(c.z);
      o[3] = O(c.w);
    }
  };
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  const long long chunk = (count + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    const long long lo = first + t * chunk, hi = std::min(first + count, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
  return 0;
}

}  // namespace

extern "C" {

// out / dst index pixels of the whole frame (4 values per pixel); only
// [first, first + count) are written.
int wto_water_shade_f64(const wto_shade* s, const float* frags, const uint8_t* coverage, long long width,
                        long long first, long long count, double* out, const double* dst, int nthreads) {
  return run_shade<double, double>(s, frags, coverage, width, first, count, out, dst, nthreads);
}
int wto_water_shade_f32(const wto_shade* s, const float* frags, const uint8_t* coverage, long long width,
                        long long first, long long count, float* out, const float* dst, int nthreads) {
  return run_shade<float, float>(s, frags, coverage, width, first, count, out, dst, nthreads);
}

// Cube-map face and (s, t) of direction r (GL 4.6 core spec table 8.19).
int wto_cube_face(const double r[3], double* s, double* t) {
  return wto::cubeFace(wto::V3<double>{r[0], r[1], r[2]}, s, t);
}

// GL_LINEAR + CLAMP_TO_EDGE lookup of an RGBA8 image at (s, t).
void wto_sample_rgba8(const uint8_t* texels, int w, int h, double s, double t, double out[4]) {
  const wto::V4<double> c = wto::sampleRgba8Linear<double>(wto::Rgba8Image{texels, w, h}, s, t);
  out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w;
}

// texture(samplerCube, r) on six length x length RGBA8 faces.
void wto_sample_cube(const uint8_t* faces, int length, const double r[3], double out[4]) {
  const wto::V4<double> c = wto::sampleCube<double>(faces, length, wto::V3<double>{r[0], r[1], r[2]});
  out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w;
}

// GL_NEAREST + CLAMP_TO_EDGE lookup of a float image at (s, t).
double wto_sample_depth(const float* texels, int w, int h, double s, double t) {
  return wto::sampleFloatNearest<double>(texels, w, h, s, t);
}

}  // extern "C"
