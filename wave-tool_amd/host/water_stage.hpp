// Host side of the water-displacement stage, over the C ABI of include/wt.h.
//
// The reference drives its water vertex program from RenderEngine (src/render-engine.h,
// src/render-engine.cpp): a set of geometry::GerstnerWave objects whose parameters are
// uploaded as the `gerstnerWaves[]` uniform block, the per-frame water block of
// RenderEngine::render (projector corners, uniforms, draw; src/render-engine.cpp:796-1115)
// and the water grid's MeshObject buffers (src/mesh-object.h:74-75, index list of
// src/program.cpp:686-717). This header provides the same three pieces for a renderer or
// a simulation loop that wants the displaced grid in device memory:
//   WaveSet     - the wave parameters and the uniform-block normalisation;
//   WaterGrid   - device vertex buffers (float3 SoA) and the triangle index list;
//   WaterStage  - the per-frame update: camera -> corners -> wave block -> kernel.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wt.h"

namespace wavetool {

// geometry::GerstnerWave (src/render-engine.h:130-167): the user parameters of one wave.
// phaseOffset_phi0 has no reference counterpart (the reference's phase has no offset).
struct GerstnerWave {
  float amplitude_A = 0.0f;        // crest height above the rest plane
  float frequency_w = 0.0f;        // angular wavenumber, 2 pi / L
  float phaseConstant_phi = 0.0f;  // angular speed (speed x w), rad/s
  float steepness_Q = 0.0f;        // in [0, 1] before normalisation
  float xzDirection_D[2] = {1.0f, 0.0f};
  float phaseOffset_phi0 = 0.0f;
};

// The reference keeps Count() and TotalAmplitude() in statics of GerstnerWave
// (src/render-engine.h:135-136, 156-165); a WaveSet keeps them per set so several sets
// can coexist.
class WaveSet {
 public:
  // Rejects (WT_ERR_INVALID_ARG) what the reference's constructor asserts against
  // (src/render-engine.h:146-154): A, w, phi >= 0, Q in [0, 1], |D| within 0.001 of 1.
  wt_status add(const GerstnerWave& wave);
  void clear();
  uint32_t Count() const { return (uint32_t)waves_.size(); }
  float TotalAmplitude() const { return total_amplitude_; }
  const std::vector<GerstnerWave>& waves() const { return waves_; }
  // The uniform block as src/render-engine.cpp:1067-1081 uploads it: Q_i = Q / (w A Count())
  // (0 where w A = 0), member order of the GLSL struct plus phi0. `normaliser` replaces
  // Count() in that formula when non-zero.
  std::vector<wt_gerstner_wave> uniform_block(uint32_t normaliser = 0) const;

 private:
  std::vector<GerstnerWave> waves_;
  float total_amplitude_ = 0.0f;
};

// The water grid's vertex buffers (the reference's MeshObject attribute arrays): device
// float3 arrays of gridLength^2 vertices, vertex id = row * gridLength + column, row 0 at
// the grid's bottom edge.
class WaterGrid {
 public:
  WaterGrid() = default;
  WaterGrid(const WaterGrid&) = delete;
  WaterGrid& operator=(const WaterGrid&) = delete;
  ~WaterGrid();
  // Allocates position and normal (and tangent / bitangent when `tangents`) arrays.
  wt_status allocate(uint32_t gridLength, bool tangents);
  uint32_t gridLength() const { return n_; }
  uint64_t vertexCount() const { return (uint64_t)n_ * n_; }
  float* positions() const { return pos_; }
  float* normals() const { return nrm_; }
  float* tangents() const { return tan_; }
  float* bitangents() const { return bit_; }
  wt_vertex_buffers buffers() const { return wt_vertex_buffers{pos_, nrm_, tan_, bit_}; }
  // The triangle list of src/program.cpp:686-717: per cell (row, col) the
  // counter-clockwise triangles (r,c) (r+1,c+1) (r+1,c) and (r,c) (r,c+1) (r+1,c+1).
  static std::vector<uint32_t> triangleIndices(uint32_t gridLength);

 private:
  void release();
  uint32_t n_ = 0;
  float* pos_ = nullptr;
  float* nrm_ = nullptr;
  float* tan_ = nullptr;
  float* bit_ = nullptr;
};

// RenderEngine's water parameters (src/render-engine.h:224-237) that the update reads.
struct WaterSettings {
  float heightmapDisplacementScale = 1.0f;
  float heightmapSampleScale = 0.02f;
  float verticalBounceWaveAmplitude = 0.1f;
  float verticalBounceWavePhase = 0.0f;  // in [0, 1]; bounce = A_b sin(2 pi phase)
  float heightmapScrollVelocity[2] = {0.0f, 0.0f};
};

// RenderEngine's scene parameters the water fragment shader reads
// (src/render-engine.h:213-235, the defaults here are the reference's).
struct ShadeSettings {
  float timeOfDayInHours = 9.0f;                           // in [0, 24]
  float fogColourFarAtNoon[4] = {1.0f, 1.0f, 1.0f, 0.1f};  // rgb tint, a = density
  float fogDepthRadiusFar = 1.0f;
  float fogDepthRadiusNear = 0.0f;
  float softEdgesDeltaDepthThreshold = 0.05f;
  float sunShininess = 50.0f;
  float sunStrength = 1.0f;
  float tintDeltaDepthThreshold = 0.1f;
  float waterClarity = 0.3f;
  float zFar = 100.0f;  // Z_FAR
  float zNear = 0.1f;   // Z_NEAR
};

// The fragment uniforms RenderEngine::render forms each frame: the sun direction from
// the time of day (src/render-engine.cpp:289-291), the fog colour dimmed by the sun's
// height (:295), the viewport from the window size (:1093), the rest copied
// (:1060-1098). FP32 throughout, as the reference computes them.
wt_shade_params shadeUniforms(const ShadeSettings& s, uint32_t windowWidth, uint32_t windowHeight);

// The per-frame water update of RenderEngine::render (src/render-engine.cpp:796-1115)
// with the draw replaced by the displacement kernel writing the grid's buffers.
class WaterStage {
 public:
  WaterStage() = default;
  WaterStage(const WaterStage&) = delete;
  WaterStage& operator=(const WaterStage&) = delete;
  ~WaterStage();
  wt_status create(uint32_t gridLength, bool tangents);
  // One frame on `stream` (a hipStream_t, NULL = default stream):
  //  1. DISPLACEABLE_AMPLITUDE = TotalAmplitude + s_d + A_b (render-engine.cpp:803) and
  //     the projector corners for `camera` (render-engine.cpp:797-1042); a frustum that
  //     misses the slab returns WT_ERR_NO_INTERSECTION and leaves the buffers untouched,
  //     as the reference skips the draw (render-engine.cpp:894);
  //  2. the uniform block of `waves` (render-engine.cpp:1067-1081), copied to the device;
  //  3. the vertical bounce A_b sin(2 pi phase) (render-engine.cpp:299-300);
  //  4. wt_water_displace into grid().
  // `heightmap` may be NULL (no bump). Host staging is reused, so the call waits for
  // the previous frame's wave-block copy before overwriting it.
  wt_status update(const wt_camera& camera, const WaveSet& waves, double waveAnimationTimeInSeconds,
                   const WaterSettings& settings, const wt_heightmap* heightmap, void* stream);
  // The water program's varyings (water-grid.vert:141-158, 193) for the grid's current
  // positions and normals: viewProjection and the yaw-only view matrix of `camera`
  // (src/render-engine.cpp:278-284), then wt_vertex_varyings into `out` (any field NULL
  // skips it). Call after update() on the same stream.
  wt_status varyings(const wt_camera& camera, double waveAnimationTimeInSeconds, const WaterSettings& settings,
                     const wt_heightmap* heightmap, const wt_varying_buffers& out, void* stream);
  // The fragment half of the water draw (water-grid.frag:70-162) as a deferred pass:
  // shadeUniforms(settings, width, height), then wt_water_shade over a fragment buffer
  // the renderer rasterised, blended over `colour` when `blend` (the reference's blend
  // state, src/render-engine.cpp:302-303).
  wt_status shade(const ShadeSettings& settings, const wt_shade_textures& textures,
                  const wt_water_fragment* fragments, const uint8_t* coverage, uint32_t width, uint32_t height,
                  float* colour, bool blend, void* stream);
  const WaterGrid& grid() const { return grid_; }
  const double* corners() const { return &corners_[0][0]; }
  const std::string& lastError() const { return err_; }

 private:
  wt_status fail(wt_status st, const std::string& msg);
  wt_ctx* ctx_ = nullptr;
  WaterGrid grid_;
  wt_gerstner_wave* d_waves_ = nullptr;  // device wave block, WT_MAX_WAVES entries
  wt_gerstner_wave* h_waves_ = nullptr;  // pinned staging
  void* copy_done_ = nullptr;            // hipEvent_t after the last staging copy
  double corners_[4][4] = {};
  std::string err_;
};

}  // namespace wavetool
