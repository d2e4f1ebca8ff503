// Host side of the water-displacement stage (see water_stage.hpp).
#include "water_stage.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace wavetool {

// ------------------------------------------------------------------ WaveSet
wt_status WaveSet::add(const GerstnerWave& w) {
  // src/render-engine.h:146-154
  const float len = std::sqrt(w.xzDirection_D[0] * w.xzDirection_D[0] + w.xzDirection_D[1] * w.xzDirection_D[1]);
  const bool ok = w.amplitude_A >= 0.0f && w.frequency_w >= 0.0f && w.phaseConstant_phi >= 0.0f &&
                  w.steepness_Q >= 0.0f && w.steepness_Q <= 1.0f && len >= 1.0f - 0.001f && len <= 1.0f + 0.001f &&
                  std::isfinite(w.phaseOffset_phi0);
  if (!ok || waves_.size() >= WT_MAX_WAVES) return WT_ERR_INVALID_ARG;
  waves_.push_back(w);
  total_amplitude_ += w.amplitude_A;
  return WT_OK;
}

void WaveSet::clear() {
  waves_.clear();
  total_amplitude_ = 0.0f;
}

std::vector<wt_gerstner_wave> WaveSet::uniform_block(uint32_t normaliser) const {
  std::vector<wt_gerstner_wave> out(waves_.size());
  std::vector<float> q(waves_.size());
  for (size_t i = 0; i < waves_.size(); ++i) {
    const GerstnerWave& w = waves_[i];
    wt_gerstner_wave& o = out[i];
    o.amplitude_A = w.amplitude_A;
    o.frequency_w = w.frequency_w;
    o.phaseConstant_phi = w.phaseConstant_phi;
    o.steepness_Q_i = 0.0f;
    o.xzDirection_D[0] = w.xzDirection_D[0];
    o.xzDirection_D[1] = w.xzDirection_D[1];
    o.phaseOffset_phi0 = w.phaseOffset_phi0;
    o.reserved = 0.0f;
    q[i] = w.steepness_Q;
  }
  if (!out.empty())
    wt_normalize_steepness(out.data(), q.data(), (uint32_t)out.size(), normaliser ? normaliser : Count());
  return out;
}

// ------------------------------------------------------------------ WaterGrid
WaterGrid::~WaterGrid() { release(); }

void WaterGrid::release() {
  for (float** p : {&pos_, &nrm_, &tan_, &bit_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  n_ = 0;
}

wt_status WaterGrid::allocate(uint32_t gridLength, bool tangents) {
  release();
  if (gridLength < 2 || gridLength > 16384) return WT_ERR_INVALID_ARG;
  const size_t bytes = (size_t)gridLength * gridLength * 3 * sizeof(float);
  float** want[4] = {&pos_, &nrm_, tangents ? &tan_ : nullptr, tangents ? &bit_ : nullptr};
  for (float** p : want) {
    if (!p) continue;
    if (hipMalloc((void**)p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      release();
      return WT_ERR_OOM;
    }
  }
  n_ = gridLength;
  return WT_OK;
}

std::vector<uint32_t> WaterGrid::triangleIndices(uint32_t n) {
  // src/program.cpp:686-717: vertex id = row * n + col, two CCW triangles per cell
  std::vector<uint32_t> idx;
  if (n < 2) return idx;
  idx.reserve((size_t)6 * (n - 1) * (n - 1));
  for (uint32_t r = 0; r + 1 < n; ++r) {
    for (uint32_t c = 0; c + 1 < n; ++c) {
      const uint32_t bl = r * n + c, br = bl + 1, tl = bl + n, tr = tl + 1;
      idx.insert(idx.end(), {bl, tr, tl, bl, br, tr});
    }
  }
  return idx;
}

// ------------------------------------------------------------------ WaterStage
WaterStage::~WaterStage() {
  if (copy_done_) {
    (void)hipEventSynchronize((hipEvent_t)copy_done_);
    (void)hipEventDestroy((hipEvent_t)copy_done_);
  }
  if (h_waves_) (void)hipHostFree(h_waves_);
  if (d_waves_) (void)hipFree(d_waves_);
  if (ctx_) wt_destroy(ctx_);
}

wt_status WaterStage::fail(wt_status st, const std::string& msg) {
  err_ = msg;
  return st;
}

wt_status WaterStage::create(uint32_t gridLength, bool tangents) {
  if (ctx_) return fail(WT_ERR_INVALID_ARG, "WaterStage::create called twice");
  wt_status st = wt_create(&ctx_);
  if (st != WT_OK) return fail(st, "wt_create failed");
  st = grid_.allocate(gridLength, tangents);
  if (st != WT_OK) return fail(st, "vertex buffer allocation failed");
  const size_t bytes = sizeof(wt_gerstner_wave) * WT_MAX_WAVES;
  if (hipMalloc((void**)&d_waves_, bytes) != hipSuccess || hipHostMalloc((void**)&h_waves_, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(WT_ERR_OOM, "wave block allocation failed");
  }
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(WT_ERR_HIP, "hipEventCreate failed");
  copy_done_ = ev;
  return WT_OK;
}

wt_status WaterStage::update(const wt_camera& camera, const WaveSet& waves, double time, const WaterSettings& s,
                             const wt_heightmap* heightmap, void* stream) {
  if (!ctx_) return fail(WT_ERR_INVALID_ARG, "WaterStage::update before create");
  if (waves.Count() > WT_MAX_WAVES) return fail(WT_ERR_INVALID_ARG, "too many waves");
  // 1. projector corners of the displaceable slab (render-engine.cpp:803)
  const float amp = waves.TotalAmplitude() + s.heightmapDisplacementScale + s.verticalBounceWaveAmplitude;
  double corners[4][4];
  wt_status st = wt_projector_corners(&camera, amp, corners);
  if (st != WT_OK) return fail(st, st == WT_ERR_NO_INTERSECTION ? "frustum misses the water slab" : "projector failed");
  std::memcpy(corners_, corners, sizeof corners_);
  // 2. uniform block -> device (wait until the previous frame's copy has read the staging)
  const std::vector<wt_gerstner_wave> block = waves.uniform_block();
  hipStream_t hs = (hipStream_t)stream;
  if (!block.empty()) {
    if (hipEventSynchronize((hipEvent_t)copy_done_) != hipSuccess) return fail(WT_ERR_HIP, "hipEventSynchronize failed");
    std::memcpy(h_waves_, block.data(), sizeof(wt_gerstner_wave) * block.size());
    if (hipMemcpyAsync(d_waves_, h_waves_, sizeof(wt_gerstner_wave) * block.size(), hipMemcpyHostToDevice, hs) !=
            hipSuccess ||
        hipEventRecord((hipEvent_t)copy_done_, hs) != hipSuccess)
      return fail(WT_ERR_HIP, "wave block upload failed");
  }
  // 3. uniforms of the water program (render-engine.cpp:1057-1110), bounce of :299-300
  wt_water_params p;
  std::memset(&p, 0, sizeof p);
  std::memcpy(p.corners, corners, sizeof p.corners);
  for (int k = 0; k < 3; ++k) p.cameraPosition[k] = camera.position[k];
  p.gridLength = grid_.gridLength();
  p.heightmapSampleScale = s.heightmapSampleScale;
  p.heightmapDisplacementScale = s.heightmapDisplacementScale;
  p.verticalBounceWaveDisplacement =
      s.verticalBounceWaveAmplitude * std::sin(6.283185307179586f * s.verticalBounceWavePhase);
  p.heightmapScrollVelocity[0] = s.heightmapScrollVelocity[0];
  p.heightmapScrollVelocity[1] = s.heightmapScrollVelocity[1];
  p.waveAnimationTimeInSeconds = time;
  // 4. the displacement kernel into the grid's buffers
  const wt_vertex_buffers vb = grid_.buffers();
  st = wt_water_displace(ctx_, &p, d_waves_, waves.Count(), heightmap, &vb, stream);
  if (st != WT_OK) return fail(st, wt_last_error(ctx_));
  return WT_OK;
}

wt_status WaterStage::varyings(const wt_camera& camera, double time, const WaterSettings& s,
                               const wt_heightmap* heightmap, const wt_varying_buffers& out, void* stream) {
  if (!ctx_) return fail(WT_ERR_INVALID_ARG, "WaterStage::varyings before create");
  double view[16], proj[16], fwd[3], view_yaw[16], proj_yaw[16], fwd_yaw[3];
  wt_camera yaw_only = camera;
  yaw_only.pitch_deg = 0.0f;  // Camera cameraOnlyYaw; setRotation(yaw, 0) (render-engine.cpp:279-281)
  if (wt_camera_matrices(&camera, view, proj, fwd) != WT_OK ||
      wt_camera_matrices(&yaw_only, view_yaw, proj_yaw, fwd_yaw) != WT_OK)
    return fail(WT_ERR_INVALID_ARG, "camera matrices");
  float vp[16], vy[16];
  for (int c = 0; c < 4; ++c)  // column-major: element (row r, column c) at [4c + r]
    for (int r = 0; r < 4; ++r) {
      double acc = 0.0;
      for (int k = 0; k < 4; ++k) acc += proj[4 * k + r] * view[4 * c + k];
      vp[4 * c + r] = (float)acc;
      vy[4 * c + r] = (float)view_yaw[4 * c + r];
    }
  wt_water_params p;
  std::memset(&p, 0, sizeof p);
  for (int k = 0; k < 3; ++k) p.cameraPosition[k] = camera.position[k];
  p.gridLength = grid_.gridLength();
  p.heightmapSampleScale = s.heightmapSampleScale;
  p.heightmapDisplacementScale = s.heightmapDisplacementScale;
  p.heightmapScrollVelocity[0] = s.heightmapScrollVelocity[0];
  p.heightmapScrollVelocity[1] = s.heightmapScrollVelocity[1];
  p.waveAnimationTimeInSeconds = time;
  const wt_status st = wt_vertex_varyings(ctx_, &p, heightmap, vp, vy, grid_.positions(), grid_.normals(),
                                          grid_.vertexCount(), &out, stream);
  return st == WT_OK ? WT_OK : fail(st, wt_last_error(ctx_));
}

wt_shade_params shadeUniforms(const ShadeSettings& s, uint32_t windowWidth, uint32_t windowHeight) {
  wt_shade_params u;
  std::memset(&u, 0, sizeof u);
  // sun position on the unit circle in the x-y plane, at -90 degrees at midnight
  const float timeOfDayInDays = s.timeOfDayInHours / 24.0f;
  const float theta = timeOfDayInDays * 6.28318530717958647692f - 1.57079632679489661923f;
  u.sunPosition[0] = std::cos(theta);
  u.sunPosition[1] = std::sin(theta);
  u.sunPosition[2] = 0.0f;
  const float dim = std::min(std::max(u.sunPosition[1], 0.0f), 1.0f);
  for (int k = 0; k < 3; ++k) u.fogColourFarAtCurrentTime[k] = dim * s.fogColourFarAtNoon[k];
  u.fogColourFarAtCurrentTime[3] = s.fogColourFarAtNoon[3];
  u.fogDepthRadiusFar = s.fogDepthRadiusFar;
  u.fogDepthRadiusNear = s.fogDepthRadiusNear;
  u.softEdgesDeltaDepthThreshold = s.softEdgesDeltaDepthThreshold;
  u.sunShininess = s.sunShininess;
  u.sunStrength = s.sunStrength;
  u.tintDeltaDepthThreshold = s.tintDeltaDepthThreshold;
  u.viewportWidthHeight[0] = (float)windowWidth;
  u.viewportWidthHeight[1] = (float)windowHeight;
  u.waterClarity = s.waterClarity;
  u.zFar = s.zFar;
  u.zNear = s.zNear;
  return u;
}

wt_status WaterStage::shade(const ShadeSettings& settings, const wt_shade_textures& textures,
                            const wt_water_fragment* fragments, const uint8_t* coverage, uint32_t width,
                            uint32_t height, float* colour, bool blend, void* stream) {
  if (!ctx_) return fail(WT_ERR_INVALID_ARG, "WaterStage::shade before create");
  const wt_shade_params u = shadeUniforms(settings, width, height);
  const wt_status st =
      wt_water_shade(ctx_, &u, &textures, fragments, coverage, width, height, colour, blend ? 1 : 0, stream);
  return st == WT_OK ? WT_OK : fail(st, wt_last_error(ctx_));
}

}  // namespace wavetool
