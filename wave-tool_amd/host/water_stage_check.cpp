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
#include <vector>

#include "heightmap_png.hpp"
#include "water_stage.hpp"

using wavetool::GerstnerWave;
using wavetool::WaterGrid;
using wavetool::WaterSettings;
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
  if (argc != 2) return fail("usage: water_stage_check <out.bin> | --indices <n> | --png <in.png> <out.bin>");

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
  // a camera looking up from high above misses the slab: the frame is skipped
  const wt_camera up{72.0f, 1.0f, 0.1f, 50.0f, {0.0f, 100.0f, 0.0f}, -90.0f, 80.0f};
  if (stage.update(up, waves, 0.0, settings, &hm, nullptr) != WT_ERR_NO_INTERSECTION)
    return fail("expected WT_ERR_NO_INTERSECTION");
  (void)hipFree(d_tex);
  (void)hipFree(d_clip);
  std::printf("water_stage_check: ok\n");
  return 0;
}
