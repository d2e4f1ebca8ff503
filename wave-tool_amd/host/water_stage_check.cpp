// Drives the host water stage the way the reference's render loop does, on the
// reference's default scene, and writes what it produced for the tests to compare.
//
//   water_stage_check <out.bin>      default scene (src/render-engine.cpp:46-52: four waves,
//                                    camera fov 72, aspect 1, near 0.1, far 100 at (0, 4, 70)),
//                                    17 x 17 grid with tangents, a 16 x 16 8-bit MIRRORED_REPEAT
//                                    heightmap, s_h 0.02, s_d 1.0, A_b 0.1 at bounce phase 0,
//                                    frames at t = 0, 10, 3600 s. Writes the corners (4x4 f64)
//                                    then per frame position, normal, tangent, bitangent (f32)
//                                    and the clip-space position (float4, f32).
//   water_stage_check --indices <n>  prints the triangle index list of an n x n grid (no GPU)
//   water_stage_check --png <in.png> <out.bin>
//                                    decodes a heightmap PNG the way the reference loads it and
//                                    writes width, height (u32) and the R8 texels (no GPU)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "heightmap_png.hpp"
#include "water_stage.hpp"

using wavetool::GerstnerWave;
using wavetool::WaterGrid;
using wavetool::ShadeSettings;
using wavetool::WaterSettings;
using wavetool::shadeUniforms;
using wavetool::WaterStage;
using wavetool::WaveSet;

static int fail(const char* what, const std::string& detail = "") {
  std::fprintf(stderr, "water_stage_check: %s %s\n", what, detail.c_str());
  return 1;
}

int main(int argc, char** argv) {
  if (argc == 3 && std::strcmp(argv[1], "--indices") == 0) {
    const uint32_t n = (uint32_t)std::atoi(argv[2]);
    for (uint32_t v : WaterGrid::triangleIndices(n)) std::printf("%u\n", v);
    return 0;
  }
  if (argc == 5 && std::strcmp(argv[1], "--shade-uniforms") == 0) {
    // host-only: the fragment uniforms of the reference's defaults at a time of day
    ShadeSettings s;
    s.timeOfDayInHours = (float)std::atof(argv[2]);
    const wt_shade_params u = shadeUniforms(s, (uint32_t)std::atoi(argv[3]), (uint32_t)std::atoi(argv[4]));
    std::printf("%.9g %.9g %.9g %.9g\n", u.fogColourFarAtCurrentTime[0], u.fogColourFarAtCurrentTime[1],
                u.fogColourFarAtCurrentTime[2], u.fogColourFarAtCurrentTime[3]);
    std::printf("%.9g %.9g %.9g\n", u.sunPosition[0], u.sunPosition[1], u.sunPosition[2]);
    std::printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", u.fogDepthRadiusFar,
                u.fogDepthRadiusNear, u.softEdgesDeltaDepthThreshold, u.sunShininess, u.sunStrength,
                u.tintDeltaDepthThreshold, u.viewportWidthHeight[0], u.viewportWidthHeight[1], u.waterClarity, u.zFar,
                u.zNear, 0.0);
    return 0;
  }
  if (argc == 4 && std::strcmp(argv[1], "--png") == 0) {
    std::vector<uint8_t> texels;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!wavetool::load_heightmap_png(argv[2], &texels, &w, &h, &err)) return fail("png:", err);
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return fail("cannot open", argv[3]);
    std::fwrite(&w, 4, 1, f);
    std::fwrite(&h, 4, 1, f);
    std::fwrite(texels.data(), 1, texels.size(), f);
    std::fclose(f);
    return 0;
  }
  if (argc != 2)
    return fail("usage: water_stage_check <out.bin> | --indices <n> | --png <in.png> <out.bin> | --shade-uniforms "
                "<hours> <w> <h>");

  WaveSet waves;
  const float d1 = 1.0f / std::sqrt(2.0f);
  const GerstnerWave defaults[4] = {{0.06f, 1.0f, 2.0f, 1.0f, {1.0f, 0.0f}, 0.0f},
                                    {0.1f, 1.0f, 0.2f, 0.0f, {d1, d1}, 0.0f},
                                    {0.0f, 0.0f, 0.0f, 0.0f, {0.0f, 1.0f}, 0.0f},
                                    {0.0f, 0.0f, 0.0f, 0.0f, {0.0f, 1.0f}, 0.0f}};
  for (const GerstnerWave& w : defaults)
    if (waves.add(w) != WT_OK) return fail("WaveSet::add rejected a default wave");
  GerstnerWave bad = defaults[0];
  bad.steepness_Q = 1.5f;
  if (waves.add(bad) != WT_ERR_INVALID_ARG) return fail("WaveSet::add accepted Q > 1");

  const uint32_t N = 17;
  WaterStage stage;
  if (stage.create(N, true) != WT_OK) return fail("WaterStage::create", stage.lastError());

  std::vector<uint8_t> tex(16 * 16);
  for (int i = 0; i < 256; ++i) tex[i] = (uint8_t)((i * 37 + 11) % 256);
  uint8_t* d_tex = nullptr;
  if (hipMalloc((void**)&d_tex, tex.size()) != hipSuccess ||
      hipMemcpy(d_tex, tex.data(), tex.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail("heightmap upload");
  const wt_heightmap hm{d_tex, 16, 16, WT_TEXEL_R8_UNORM, WT_WRAP_MIRRORED_REPEAT};

  const wt_camera cam{72.0f, 1.0f, 0.1f, 100.0f, {0.0f, 4.0f, 70.0f}, -90.0f, 0.0f};
  WaterSettings settings;  // reference defaults: s_d 1.0, s_h 0.02, A_b 0.1, phase 0

  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return fail("cannot open", argv[1]);
  bool corners_written = false;
  const size_t nf = (size_t)N * N * 3;
  std::vector<float> host(nf), host_clip((size_t)N * N * 4);
  float* d_clip = nullptr;
  if (hipMalloc((void**)&d_clip, host_clip.size() * sizeof(float)) != hipSuccess) return fail("clip buffer");
  wt_varying_buffers vary{};
  vary.clip_position = d_clip;
  for (double t : {0.0, 10.0, 3600.0}) {
    if (stage.update(cam, waves, t, settings, &hm, nullptr) != WT_OK) return fail("update", stage.lastError());
    if (hipDeviceSynchronize() != hipSuccess) return fail("hipDeviceSynchronize");
    if (!corners_written) {
      std::fwrite(stage.corners(), sizeof(double), 16, f);
      corners_written = true;
    }
    const WaterGrid& g = stage.grid();
    for (const float* d : {g.positions(), g.normals(), g.tangents(), g.bitangents()}) {
      if (hipMemcpy(host.data(), d, nf * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return fail("readback");
      std::fwrite(host.data(), sizeof(float), nf, f);
    }
    if (stage.varyings(cam, t, settings, &hm, vary, nullptr) != WT_OK) return fail("varyings", stage.lastError());
    if (hipMemcpy(host_clip.data(), d_clip, host_clip.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("clip readback");
    std::fwrite(host_clip.data(), sizeof(float), host_clip.size(), f);
  }
  std::fclose(f);

  // the fragment stage on a 4 x 2 frame through WaterStage::shade (default settings):
  // inputs and colours go to <out>.shade for the oracle to repeat
  {
    const uint32_t W = 4, H = 2, P = W * H, L = 2;
    std::vector<wt_water_fragment> frags(P);
    std::vector<float> depth(P);
    std::vector<uint8_t> refl(P * 4), refr(P * 4), sky(6 * L * L * 4);
    for (uint32_t i = 0; i < P; ++i) {
      const float a = 0.05f * (float)i;
      wt_water_fragment& q = frags[i];
      q.normal[0] = 0.1f - a * 0.3f; q.normal[1] = 0.98f; q.normal[2] = 0.05f + a * 0.2f;
      q.view_vector[0] = 1.0f + 3.0f * a; q.view_vector[1] = 2.0f + a; q.view_vector[2] = -6.0f - 20.0f * a;
      q.normal_view_only_yaw[0] = 0.2f - a; q.normal_view_only_yaw[1] = 0.97f; q.normal_view_only_yaw[2] = -0.1f + a;
      q.ndc_xy_height0[0] = -0.6f + 0.3f * (float)(i % W); q.ndc_xy_height0[1] = -0.5f + 0.8f * (float)(i / W);
      q.frag_depth = 0.9f + 0.01f * (float)i;
      depth[i] = i == 5 ? 1.0f : 0.97f + 0.003f * (float)i;
      for (int c = 0; c < 4; ++c) {
        refl[4 * i + c] = (uint8_t)((i * 53 + c * 91 + 7) % 256);
        refr[4 * i + c] = (uint8_t)((i * 29 + c * 67 + 200) % 256);
      }
    }
    for (size_t k = 0; k < sky.size(); ++k) sky[k] = (uint8_t)((k * 13 + 40) % 256);
    void *d_frags = nullptr, *d_depth = nullptr, *d_refl = nullptr, *d_refr = nullptr, *d_sky = nullptr,
         *d_col = nullptr;
    if (hipMalloc(&d_frags, P * sizeof(wt_water_fragment)) != hipSuccess ||
        hipMalloc(&d_depth, P * sizeof(float)) != hipSuccess || hipMalloc(&d_refl, refl.size()) != hipSuccess ||
        hipMalloc(&d_refr, refr.size()) != hipSuccess || hipMalloc(&d_sky, sky.size()) != hipSuccess ||
        hipMalloc(&d_col, P * 4 * sizeof(float)) != hipSuccess)
      return fail("shade buffers");
    if (hipMemcpy(d_frags, frags.data(), P * sizeof(wt_water_fragment), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_depth, depth.data(), P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_refl, refl.data(), refl.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_refr, refr.data(), refr.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_sky, sky.data(), sky.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail("shade upload");
    const wt_shade_textures tx{(const float*)d_depth, W, H, {(const uint8_t*)d_refl, W, H},
                               {(const uint8_t*)d_refr, W, H}, (const uint8_t*)d_sky, L};
    const wavetool::ShadeSettings shade_settings;
    if (stage.shade(shade_settings, tx, (const wt_water_fragment*)d_frags, nullptr, W, H, (float*)d_col, false,
                    nullptr) != WT_OK)
      return fail("shade", stage.lastError());
    std::vector<float> col(P * 4);
    if (hipMemcpy(col.data(), d_col, col.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("shade readback");
    const std::string path = std::string(argv[1]) + ".shade";
    FILE* g = std::fopen(path.c_str(), "wb");
    if (!g) return fail("cannot open", path.c_str());
    std::fwrite(frags.data(), sizeof(wt_water_fragment), P, g);
    std::fwrite(depth.data(), sizeof(float), P, g);
    std::fwrite(refl.data(), 1, refl.size(), g);
    std::fwrite(refr.data(), 1, refr.size(), g);
    std::fwrite(sky.data(), 1, sky.size(), g);
    std::fwrite(col.data(), sizeof(float), col.size(), g);
    std::fclose(g);
    for (void* d : {d_frags, d_depth, d_refl, d_refr, d_sky, d_col}) (void)hipFree(d);
  }
  // a camera looking up from high above misses the slab: the frame is skipped
  const wt_camera up{72.0f, 1.0f, 0.1f, 50.0f, {0.0f, 100.0f, 0.0f}, -90.0f, 80.0f};
  if (stage.update(up, waves, 0.0, settings, &hm, nullptr) != WT_ERR_NO_INTERSECTION)
    return fail("expected WT_ERR_NO_INTERSECTION");
  (void)hipFree(d_tex);
  (void)hipFree(d_clip);
  std::printf("water_stage_check: ok\n");
  return 0;
}
