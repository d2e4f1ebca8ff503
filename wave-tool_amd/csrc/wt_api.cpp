// C-ABI of the water-displacement stage: context, validation, dispatch, graphs and
// the forward autotuner. The reference counterpart of each entry point is named in
// include/wt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wt.h"
#include "wt_fit.hpp"
#include "wt_kernels.hpp"
#include "wt_projector.hpp"

struct wt_ctx {
  int device = 0;
  std::string err;
  bool forced = false;
  wt::LaunchShape forced_shape{};
  std::map<std::pair<int, int>, wt::LaunchShape> table;  // (N bucket, W bucket) -> shape
  std::map<std::pair<int, int>, int> batch_table;        // (N bucket, W bucket) -> batch shape
};

struct wt_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

wt_status fail(wt_ctx* ctx, wt_status s, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return s;
}

wt_status hip_fail(wt_ctx* ctx, hipError_t e, const char* where) {
  return fail(ctx, WT_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

bool is_finite(double v) { return std::isfinite(v); }

int bucket_log2(uint32_t v) {
  int b = 0;
  while ((1u << b) < v && b < 31) ++b;
  return b;
}

// Wave chunks a launch shape can stage: 32, 64 or 128 waves for the one-row shapes
// (two wavefronts per segment, 64 lanes per row), 64 for the others.
bool chunk_ok(const wt::LaunchShape& s) {
  if (s.wave_chunk == 64 || s.wave_chunk == 0) return true;
  return (s.wave_chunk == 32 || s.wave_chunk == 128) && s.split == 2 && s.lanes_per_row == 64;
}

bool valid_shape(const wt::LaunchShape& s) {
  const bool k_ok = s.steps == 4 || s.steps == 8 || s.steps == 16;
  const bool lc_ok = s.lanes_per_row == 64 || s.lanes_per_row == 32 || s.lanes_per_row == 16 || s.lanes_per_row == 8;
  return (s.split == 1 || s.split == 2) && k_ok && lc_ok && chunk_ok(s) && s.steps >= 2 * (s.split - 1);
}

wt::LaunchShape default_shape(uint32_t N) {
  // lanes per row no wider than the row, steps so one segment spans about a row
  wt::LaunchShape s{2, 8, 64, 64};
  if (N <= 64) s.lanes_per_row = 16;
  else if (N <= 256) s.lanes_per_row = 32;
  if (N <= 64) s.steps = 4;
  return s;
}

wt::LaunchShape pick_shape(wt_ctx* ctx, uint32_t N, uint32_t W) {
  if (ctx->forced) return ctx->forced_shape;
  auto it = ctx->table.find({bucket_log2(N), bucket_log2(W ? W : 1)});
  if (it != ctx->table.end()) return it->second;
  return default_shape(N);
}

bool is_device_pointer(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeUnified;
}

bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return cs != hipStreamCaptureStatusNone;
}

// texture-space scroll offset -s_h * v * t reduced to one wrap period in FP64
double reduce_offset(double v, int wrap) {
  const double period = wrap == WT_WRAP_REPEAT ? 1.0 : 2.0;
  return v - period * std::floor(v / period);
}

}  // namespace

namespace wt {
wt_status ctx_fail(wt_ctx* ctx, wt_status s, const std::string& msg) { return fail(ctx, s, msg); }

// shared with the batched / adjoint translation units
wt_status validate_heightmap(wt_ctx* ctx, const wt_heightmap* hm, bool check_ptr) {
  if (!hm) return WT_OK;
  if (!hm->texels) return fail(ctx, WT_ERR_INVALID_ARG, "heightmap.texels is NULL");
  if (hm->width < 1 || hm->height < 1 || hm->width > 16384 || hm->height > 16384)
    return fail(ctx, WT_ERR_INVALID_ARG, "heightmap size must be in [1, 16384]");
  if (hm->format != WT_TEXEL_R8_UNORM && hm->format != WT_TEXEL_R32F)
    return fail(ctx, WT_ERR_INVALID_ARG, "heightmap.format must be R8_UNORM or R32F");
  if (hm->wrap != WT_WRAP_REPEAT && hm->wrap != WT_WRAP_MIRRORED_REPEAT)
    return fail(ctx, WT_ERR_INVALID_ARG, "heightmap.wrap must be REPEAT or MIRRORED_REPEAT");
  if (hm->format == WT_TEXEL_R32F && (reinterpret_cast<uintptr_t>(hm->texels) & 3))
    return fail(ctx, WT_ERR_INVALID_ARG, "R32F heightmap texels must be 4-byte aligned");
  if (check_ptr && !is_device_pointer(hm->texels))
    return fail(ctx, WT_ERR_INVALID_ARG, "heightmap.texels is not device memory");
  return WT_OK;
}
}  // namespace wt

extern "C" {

int wt_abi_version(void) { return WT_ABI_VERSION; }

const char* wt_status_string(wt_status s) {
  switch (s) {
    case WT_OK: return "ok";
    case WT_ERR_INVALID_ARG: return "invalid argument";
    case WT_ERR_NO_INTERSECTION: return "camera frustum does not intersect the displaceable volume";
    case WT_ERR_HIP: return "HIP runtime error";
    case WT_ERR_OOM: return "out of memory";
    case WT_ERR_UNSUPPORTED: return "unsupported";
    case WT_ERR_DEGENERATE: return "degenerate projector geometry";
  }
  return "unknown status";
}

wt_status wt_create(wt_ctx** out) {
  if (!out) return WT_ERR_INVALID_ARG;
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return WT_ERR_HIP;
  wt_ctx* c = new (std::nothrow) wt_ctx();
  if (!c) return WT_ERR_OOM;
  c->device = dev;
  // resolve this device's fixed fit scratch now, with the device current; a failure is
  // kept and reported by wt_fit_loss_grad_static, the only entry point that uses it
  std::string why;
  (void)wt::static_fit_scratch(dev, &why);
  *out = c;
  return WT_OK;
}

void wt_destroy(wt_ctx* ctx) { delete ctx; }

const char* wt_last_error(const wt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static bool camera_ok(const wt_camera* c) { return c && wt::proj::camera_valid(*c); }

wt_status wt_camera_matrices(const wt_camera* cam, double view[16], double proj[16], double forward[3]) {
  if (!camera_ok(cam) || !view || !proj || !forward) return WT_ERR_INVALID_ARG;
  const wt::proj::Camera c = wt::proj::camera_from(*cam);
  for (int i = 0; i < 16; ++i) {
    view[i] = c.view.c[i / 4][i % 4];
    proj[i] = c.proj.c[i / 4][i % 4];
  }
  forward[0] = c.forward.x;
  forward[1] = c.forward.y;
  forward[2] = c.forward.z;
  return WT_OK;
}

wt_status wt_projector_corners(const wt_camera* cam, float amp, double corners[4][4]) {
  if (!camera_ok(cam) || !corners || !std::isfinite(amp) || amp < 0.0f) return WT_ERR_INVALID_ARG;
  return (wt_status)wt::proj::projector_corners(wt::proj::camera_from(*cam), (double)amp, corners);
}

wt_status wt_normalize_steepness(wt_gerstner_wave* waves, const float* q, uint32_t count, uint32_t norm_count) {
  if ((!waves || !q) && count) return WT_ERR_INVALID_ARG;
  if (norm_count == 0 && count) return WT_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < count; ++i) {
    // render-engine.cpp:1073, in FP32 like the reference
    const float wa = waves[i].frequency_w * waves[i].amplitude_A;
    waves[i].steepness_Q_i = wa != 0.0f ? q[i] / (wa * (float)norm_count) : 0.0f;
  }
  return WT_OK;
}

wt_status wt_water_displace(wt_ctx* ctx, const wt_water_params* p, const wt_gerstner_wave* d_waves,
                            uint32_t W, const wt_heightmap* hm, const wt_vertex_buffers* out, void* stream) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (!p || !out) return fail(ctx, WT_ERR_INVALID_ARG, "params and out must be non-NULL");
  const uint32_t N = p->gridLength;
  if (N < 2 || N > 32768) return fail(ctx, WT_ERR_INVALID_ARG, "gridLength must be in [2, 32768]");
  if (W > WT_MAX_WAVES) return fail(ctx, WT_ERR_INVALID_ARG, "wave_count exceeds WT_MAX_WAVES");
  if (W && !d_waves) return fail(ctx, WT_ERR_INVALID_ARG, "d_waves is NULL");
  if (!out->position || !out->normal) return fail(ctx, WT_ERR_INVALID_ARG, "position and normal buffers are required");
  const void* bufs[4] = {out->position, out->normal, out->tangent, out->bitangent};
  for (const void* b : bufs)
    if (b && (reinterpret_cast<uintptr_t>(b) & 3)) return fail(ctx, WT_ERR_INVALID_ARG, "output buffers must be 4-byte aligned");
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k)
      if (!is_finite(p->corners[i][k])) return fail(ctx, WT_ERR_INVALID_ARG, "corners must be finite");
  const float fs[8] = {p->cameraPosition[0], p->cameraPosition[1], p->cameraPosition[2], p->heightmapSampleScale,
                       p->heightmapDisplacementScale, p->verticalBounceWaveDisplacement,
                       p->heightmapScrollVelocity[0], p->heightmapScrollVelocity[1]};
  for (float v : fs)
    if (!std::isfinite(v)) return fail(ctx, WT_ERR_INVALID_ARG, "scalar parameters must be finite");
  if (!is_finite(p->waveAnimationTimeInSeconds)) return fail(ctx, WT_ERR_INVALID_ARG, "time must be finite");
  hipStream_t st = (hipStream_t)stream;
  const bool capturing = stream_capturing(st);
  wt_status hs = wt::validate_heightmap(ctx, hm, !capturing);
  if (hs != WT_OK) return hs;
  if (!capturing) {
    for (const void* b : bufs)
      if (b && !is_device_pointer(b)) return fail(ctx, WT_ERR_INVALID_ARG, "output buffers must be device memory");
    if (W && !is_device_pointer(d_waves)) return fail(ctx, WT_ERR_INVALID_ARG, "d_waves must be device memory");
  }

  wt::FwdParams fp{};
  wt::set_corner_algebra(fp, p->corners);  // [0] BL, [1] TL, [2] BR, [3] TR
  fp.inv_nm1 = 1.0 / (double)(N - 1);
  fp.time = p->waveAnimationTimeInSeconds;
  fp.N = (int)N;
  fp.W = (int)W;
  fp.waves = d_waves;
  wt::EpilogueArgs& e = fp.ep;
  e.cam[0] = p->cameraPosition[0];
  e.cam[1] = p->cameraPosition[1];
  e.cam[2] = p->cameraPosition[2];
  e.s_h = p->heightmapSampleScale;
  e.s_d = p->heightmapDisplacementScale;
  e.bounce = p->verticalBounceWaveDisplacement;
  e.two_sd = 2.0f * e.s_d;
  e.bump_k = 2.0f * e.s_d * e.s_h;
  e.pos = out->position;
  e.nrm = out->normal;
  e.tan = out->tangent;
  e.bit = out->bitangent;
  if (hm) {
    e.texels = hm->texels;
    e.tex_w = (int)hm->width;
    e.tex_h = (int)hm->height;
    e.tex_fmt = (int)hm->format;
    e.tex_wrap = (int)hm->wrap;
    const double t = p->waveAnimationTimeInSeconds, sh = p->heightmapSampleScale;
    e.hm_off_s = reduce_offset(-sh * (double)p->heightmapScrollVelocity[0] * t, e.tex_wrap);
    e.hm_off_t = reduce_offset(sh * (double)p->heightmapScrollVelocity[1] * t, e.tex_wrap);
  }
  const wt::LaunchShape shape = pick_shape(ctx, N, W);
  hipError_t he = wt::launch_forward(fp, shape, st);
  if (he != hipSuccess) return hip_fail(ctx, he, "water_forward_kernel launch");
  return WT_OK;
}

wt_status wt_vertex_varyings(wt_ctx* ctx, const wt_water_params* p, const wt_heightmap* hm, const float vp[16],
                             const float vyaw[16], const float* positions, const float* normals, uint64_t count,
                             const wt_varying_buffers* out, void* stream) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (!p || !out || !vp || !vyaw) return fail(ctx, WT_ERR_INVALID_ARG, "params, matrices and out must be non-NULL");
  if (count > WT_MAX_VARYING_VERTICES)
    return fail(ctx, WT_ERR_INVALID_ARG, "vertex_count too large (limit 2 (2^32 - 256), one launch)");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(vp[i]) || !std::isfinite(vyaw[i])) return fail(ctx, WT_ERR_INVALID_ARG, "matrices must be finite");
  const float fs[6] = {p->cameraPosition[0], p->cameraPosition[1], p->cameraPosition[2], p->heightmapSampleScale,
                       p->heightmapScrollVelocity[0], p->heightmapScrollVelocity[1]};
  for (float v : fs)
    if (!std::isfinite(v)) return fail(ctx, WT_ERR_INVALID_ARG, "scalar parameters must be finite");
  if (!is_finite(p->waveAnimationTimeInSeconds)) return fail(ctx, WT_ERR_INVALID_ARG, "time must be finite");
  const bool needs_pos = out->clip_position || out->ndc_xy_height0 || out->view_vector || out->heightmap_sample;
  if (needs_pos && !positions) return fail(ctx, WT_ERR_INVALID_ARG, "positions are required");
  if (out->normal_view_only_yaw && !normals) return fail(ctx, WT_ERR_INVALID_ARG, "normals are required");
  if (out->heightmap_sample && !hm) return fail(ctx, WT_ERR_INVALID_ARG, "heightmap_sample needs a heightmap");
  // float4 / float2 stores need 16 / 8-byte alignment, the rest 4
  if ((reinterpret_cast<uintptr_t>(out->clip_position) & 15) || (reinterpret_cast<uintptr_t>(out->ndc_xy_height0) & 7))
    return fail(ctx, WT_ERR_INVALID_ARG, "clip_position must be 16-byte and ndc_xy_height0 8-byte aligned");
  const void* bufs[7] = {positions, normals, out->clip_position, out->ndc_xy_height0, out->view_vector,
                         out->normal_view_only_yaw, out->heightmap_sample};
  for (const void* b : bufs)
    if (b && (reinterpret_cast<uintptr_t>(b) & 3)) return fail(ctx, WT_ERR_INVALID_ARG, "buffers must be 4-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  const bool capturing = stream_capturing(st);
  wt_status hs = wt::validate_heightmap(ctx, out->heightmap_sample ? hm : nullptr, !capturing);
  if (hs != WT_OK) return hs;
  if (!capturing)
    for (const void* b : bufs)
      if (b && !is_device_pointer(b)) return fail(ctx, WT_ERR_INVALID_ARG, "buffers must be device memory");
  wt::VaryingArgs a{};
  for (int i = 0; i < 16; ++i) {
    a.vp[i] = vp[i];
    a.vyaw[i] = vyaw[i];
  }
  for (int k = 0; k < 3; ++k) a.cam[k] = p->cameraPosition[k];
  a.s_h = p->heightmapSampleScale;
  if (out->heightmap_sample) {
    a.texels = hm->texels;
    a.tex_w = (int)hm->width;
    a.tex_h = (int)hm->height;
    a.tex_fmt = (int)hm->format;
    a.tex_wrap = (int)hm->wrap;
    const double t = p->waveAnimationTimeInSeconds, sh = p->heightmapSampleScale;
    a.hm_off_s = reduce_offset(-sh * (double)p->heightmapScrollVelocity[0] * t, a.tex_wrap);
    a.hm_off_t = reduce_offset(sh * (double)p->heightmapScrollVelocity[1] * t, a.tex_wrap);
  }
  a.pos = positions;
  a.nrm = normals;
  a.count = (long long)count;
  a.clip = out->clip_position;
  a.ndc0 = out->ndc_xy_height0;
  a.view = out->view_vector;
  a.nview = out->normal_view_only_yaw;
  a.sample = out->heightmap_sample;
  hipError_t he = wt::launch_varyings(a, st);
  if (he != hipSuccess) return hip_fail(ctx, he, "vertex_varyings_kernel launch");
  return WT_OK;
}

wt_status wt_set_launch_config(wt_ctx* ctx, const wt_launch_config* cfg) {
  if (!ctx || !cfg) return WT_ERR_INVALID_ARG;
  if (!cfg->wave_split && !cfg->steps_per_lane && !cfg->lanes_per_row && !cfg->wave_chunk) {
    ctx->forced = false;
    return WT_OK;
  }
  wt::LaunchShape s{(int)cfg->wave_split, (int)cfg->steps_per_lane, (int)cfg->lanes_per_row,
                    cfg->wave_chunk ? (int)cfg->wave_chunk : 64};
  if (!valid_shape(s)) return fail(ctx, WT_ERR_INVALID_ARG, "unsupported launch configuration");
  ctx->forced = true;
  ctx->forced_shape = s;
  return WT_OK;
}

wt_status wt_get_launch_config(wt_ctx* ctx, uint32_t N, uint32_t W, wt_launch_config* cfg) {
  if (!ctx || !cfg) return WT_ERR_INVALID_ARG;
  const wt::LaunchShape s = pick_shape(ctx, N, W);
  cfg->wave_split = (uint32_t)s.split;
  cfg->steps_per_lane = (uint32_t)s.steps;
  cfg->lanes_per_row = (uint32_t)s.lanes_per_row;
  cfg->wave_chunk = (uint32_t)(s.wave_chunk ? s.wave_chunk : 64);
  return WT_OK;
}

// Times every supported launch shape (and, for the one-row shapes, every wave chunk) on
// synthetic inputs of this size with `nfl` frames in flight (frame r on stream r mod nfl,
// writing buffer set r mod nfl) and stores the fastest per frame in the context's table.
static wt_status autotune_forward(wt_ctx* ctx, uint32_t N, uint32_t W, int with_hm, uint32_t nfl,
                                  wt_launch_config* best, float* best_ms) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (N < 2 || N > 8192 || W > WT_MAX_WAVES) return fail(ctx, WT_ERR_INVALID_ARG, "autotune size out of range");
  if (nfl < 1 || nfl > 4) return fail(ctx, WT_ERR_INVALID_ARG, "frames_in_flight must be in [1, 4]");
  const size_t V = (size_t)N * N;
  float* pos[4] = {};
  float* nrm[4] = {};
  float* tex = nullptr;
  wt_gerstner_wave* dw = nullptr;
  hipStream_t st[4] = {};
  hipEvent_t e0 = nullptr, e1 = nullptr, join[4] = {};
  wt_status rs = WT_OK;
  const uint32_t TW = 256;
  std::vector<wt_gerstner_wave> hw(W ? W : 1);
  for (uint32_t i = 0; i < W; ++i) {  // deterministic synthetic wave set
    const float L = 1.0f + 59.0f * (float)((i * 37u) % 101u) / 100.0f;
    const float ang = 0.7f * std::sin(1.7f * (float)i);
    hw[i] = wt_gerstner_wave{L / 90.0f, 6.2831853f / L, std::sqrt(9.81f * 6.2831853f / L), 0.5f / (float)W,
                             {std::cos(ang), std::sin(ang)}, 0.3f * (float)i, 0.0f};
  }
  std::vector<float> htex(TW * TW);
  for (uint32_t i = 0; i < TW * TW; ++i) htex[i] = 0.5f + 0.5f * std::sin(0.01f * (float)i);
  auto cleanup = [&]() {
    for (uint32_t j = 0; j < 4; ++j) {
      if (pos[j]) (void)hipFree(pos[j]);
      if (nrm[j]) (void)hipFree(nrm[j]);
      if (st[j]) (void)hipStreamDestroy(st[j]);
      if (join[j]) (void)hipEventDestroy(join[j]);
    }
    if (tex) (void)hipFree(tex);
    if (dw) (void)hipFree(dw);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
#define WT_TRY(x)                                \
  do {                                           \
    hipError_t _e = (x);                         \
    if (_e != hipSuccess) {                      \
      rs = hip_fail(ctx, _e, #x);                \
      cleanup();                                 \
      return rs;                                 \
    }                                            \
  } while (0)
  for (uint32_t j = 0; j < nfl; ++j) {
    WT_TRY(hipMalloc(&pos[j], V * 12));
    WT_TRY(hipMalloc(&nrm[j], V * 12));
    WT_TRY(hipStreamCreateWithFlags(&st[j], hipStreamNonBlocking));
    WT_TRY(hipEventCreateWithFlags(&join[j], hipEventDisableTiming));
  }
  WT_TRY(hipMalloc(&dw, sizeof(wt_gerstner_wave) * hw.size()));
  WT_TRY(hipMalloc(&tex, sizeof(float) * htex.size()));
  WT_TRY(hipMemcpy(dw, hw.data(), sizeof(wt_gerstner_wave) * hw.size(), hipMemcpyHostToDevice));
  WT_TRY(hipMemcpy(tex, htex.data(), sizeof(float) * htex.size(), hipMemcpyHostToDevice));
  WT_TRY(hipEventCreate(&e0));
  WT_TRY(hipEventCreate(&e1));
  wt_water_params p{};
  const double cc[4][4] = {{-60, 0, -10, 1}, {-900, 0, -1000, 1}, {60, 0, -10, 1}, {900, 0, -1000, 1}};
  std::memcpy(p.corners, cc, sizeof(cc));
  p.cameraPosition[1] = 10.0f;
  p.gridLength = N;
  p.heightmapSampleScale = 0.02f;
  p.heightmapDisplacementScale = 0.05f;
  p.waveAnimationTimeInSeconds = 10.0;
  wt_heightmap hm{tex, TW, TW, WT_TEXEL_R32F, WT_WRAP_REPEAT};
  const bool saved_forced = ctx->forced;
  const wt::LaunchShape saved_shape = ctx->forced_shape;
  float best_t = 1e30f, default_t = 1e30f;
  wt::LaunchShape best_s = default_shape(N);
  const wt::LaunchShape dflt = default_shape(N);
  const int reps = nfl > 1 ? 32 : 8;
  // frames [0, n) of one timing: frame r on stream r mod nfl, forked from and joined back
  // to stream 0
  auto frames = [&](int n) -> wt_status {
    if (nfl > 1) {
      if (hipEventRecord(join[0], st[0]) != hipSuccess) return WT_ERR_HIP;
      for (uint32_t j = 1; j < nfl; ++j)
        if (hipStreamWaitEvent(st[j], join[0], 0) != hipSuccess) return WT_ERR_HIP;
    }
    for (int r = 0; r < n; ++r) {
      p.waveAnimationTimeInSeconds = 10.0 + r / 60.0;
      wt_vertex_buffers vb{pos[r % nfl], nrm[r % nfl], nullptr, nullptr};
      const wt_status s0 = wt_water_displace(ctx, &p, dw, W, with_hm ? &hm : nullptr, &vb, st[r % nfl]);
      if (s0 != WT_OK) return s0;
    }
    for (uint32_t j = 1; j < nfl; ++j) {
      if (hipEventRecord(join[j], st[j]) != hipSuccess) return WT_ERR_HIP;
      if (hipStreamWaitEvent(st[0], join[j], 0) != hipSuccess) return WT_ERR_HIP;
    }
    return WT_OK;
  };
  const int splits[2] = {1, 2}, ks[3] = {4, 8, 16}, lcs[4] = {64, 32, 16, 8}, chunks[3] = {64, 32, 128};
  for (int var : splits)
    for (int k : ks)
      for (int lc : lcs)
        for (int ch : chunks) {
          if ((uint32_t)lc > N && lc != 8) continue;
          const wt::LaunchShape cand{var, k, lc, ch};
          if (!valid_shape(cand)) continue;
          ctx->forced = true;
          ctx->forced_shape = cand;
          wt_status s0 = frames((int)nfl);  // first launches of this shape
          if (s0 == WT_OK && hipStreamSynchronize(st[0]) != hipSuccess) s0 = WT_ERR_HIP;
          if (s0 != WT_OK) {
            ctx->forced = saved_forced;
            ctx->forced_shape = saved_shape;
            cleanup();
            return s0 == WT_ERR_HIP ? fail(ctx, s0, "autotune launch") : s0;
          }
          WT_TRY(hipEventRecord(e0, st[0]));
          s0 = frames(reps);
          WT_TRY(hipEventRecord(e1, st[0]));
          WT_TRY(hipEventSynchronize(e1));
          if (s0 != WT_OK) {
            ctx->forced = saved_forced;
            ctx->forced_shape = saved_shape;
            cleanup();
            return s0 == WT_ERR_HIP ? fail(ctx, s0, "autotune launch") : s0;
          }
          float ms = 0.0f;
          WT_TRY(hipEventElapsedTime(&ms, e0, e1));
          ms /= reps;
          if (cand.split == dflt.split && cand.steps == dflt.steps && cand.lanes_per_row == dflt.lanes_per_row &&
              cand.wave_chunk == dflt.wave_chunk)
            default_t = ms;
          if (ms < best_t) {
            best_t = ms;
            best_s = cand;
          }
        }
#undef WT_TRY
  // shapes within timing noise of each other trade places from run to run: keep the
  // default shape unless another is more than 3% faster
  if (default_t <= 1.03f * best_t) {
    best_s = dflt;
    best_t = default_t;
  }
  ctx->forced = saved_forced;
  ctx->forced_shape = saved_shape;
  ctx->table[{bucket_log2(N), bucket_log2(W ? W : 1)}] = best_s;
  if (best) {
    best->wave_split = (uint32_t)best_s.split;
    best->steps_per_lane = (uint32_t)best_s.steps;
    best->lanes_per_row = (uint32_t)best_s.lanes_per_row;
    best->wave_chunk = (uint32_t)best_s.wave_chunk;
  }
  if (best_ms) *best_ms = best_t;
  cleanup();
  return WT_OK;
}

wt_status wt_autotune_forward(wt_ctx* ctx, uint32_t N, uint32_t W, int with_hm, wt_launch_config* best,
                              float* best_ms) {
  return autotune_forward(ctx, N, W, with_hm, 1, best, best_ms);
}

wt_status wt_autotune_forward_overlapped(wt_ctx* ctx, uint32_t N, uint32_t W, int with_hm, uint32_t frames_in_flight,
                                         wt_launch_config* best, float* best_ms) {
  return autotune_forward(ctx, N, W, with_hm, frames_in_flight, best, best_ms);
}

wt_status wt_autotune_save(wt_ctx* ctx, const char* path) {
  if (!ctx || !path) return WT_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(ctx, WT_ERR_INVALID_ARG, std::string("cannot open ") + path);
  std::fprintf(f, "# wt forward autotune: n_bucket w_bucket wave_split steps lanes_per_row wave_chunk\n");
  for (const auto& kv : ctx->table)
    std::fprintf(f, "%d %d %d %d %d %d\n", kv.first.first, kv.first.second, kv.second.split, kv.second.steps,
                 kv.second.lanes_per_row, kv.second.wave_chunk);
  std::fprintf(f, "# batched: b n_bucket w_bucket batch_shape\n");
  for (const auto& kv : ctx->batch_table) std::fprintf(f, "b %d %d %d\n", kv.first.first, kv.first.second, kv.second);
  std::fclose(f);
  return WT_OK;
}

wt_status wt_autotune_load(wt_ctx* ctx, const char* path) {
  if (!ctx || !path) return WT_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(ctx, WT_ERR_INVALID_ARG, std::string("cannot open ") + path);
  char line[256];
  std::map<std::pair<int, int>, wt::LaunchShape> t;
  std::map<std::pair<int, int>, int> bt;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int nb, wb;
    if (line[0] == 'b') {
      int shape;
      if (std::sscanf(line + 1, "%d %d %d", &nb, &wb, &shape) != 3 || shape < 0 || shape >= wt::kBatchShapes) {
        std::fclose(f);
        return fail(ctx, WT_ERR_INVALID_ARG, "malformed autotune table");
      }
      bt[{nb, wb}] = shape;
      continue;
    }
    wt::LaunchShape s{};
    if (std::sscanf(line, "%d %d %d %d %d %d", &nb, &wb, &s.split, &s.steps, &s.lanes_per_row, &s.wave_chunk) != 6 ||
        !valid_shape(s)) {
      std::fclose(f);
      return fail(ctx, WT_ERR_INVALID_ARG, "malformed autotune table");
    }
    t[{nb, wb}] = s;
  }
  std::fclose(f);
  ctx->table = t;
  ctx->batch_table = bt;
  return WT_OK;
}

}  // extern "C"

struct wt_batch {
  wt_ctx* ctx = nullptr;
  std::vector<wt::PatchDesc> patches;
  wt::PatchDesc* d_patches = nullptr;
  wt::BatchUnit* d_units = nullptr;
  long long shape_begin[wt::kBatchShapes] = {};
  int shape_units[wt::kBatchShapes] = {};
  uint64_t vertices = 0;
  uint32_t max_wave_end = 0;
};

// Largest patch vertex_offset a batch accepts (2^48 float3 = 3 PB, beyond any device).
static constexpr uint64_t kMaxBatchVertexOffset = 1ull << 48;

// forced_shape < 0: each patch's shape from the context's batch autotune table, else
// the padded-vertex model (batch_shape_for)
static wt_status batch_create(wt_ctx* ctx, const wt_patch* patches, uint32_t count, int forced_shape,
                              wt_batch** out) {
  if (!ctx || !out) return WT_ERR_INVALID_ARG;
  *out = nullptr;
  if (!patches || count == 0) return fail(ctx, WT_ERR_INVALID_ARG, "need at least one patch");
  wt_batch* b = new (std::nothrow) wt_batch();
  if (!b) return WT_ERR_OOM;
  b->ctx = ctx;
  uint32_t max_n = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const wt_patch& p = patches[i];
    const std::string tag = "patch " + std::to_string(i) + ": ";
    // wave and vertex ranges checked in 64 bits: the kernels index with int wave offsets
    // and 64-bit vertex offsets, so a sum that wraps must not pass
    bool ok = p.gridLength >= 2 && p.gridLength <= 32768 && p.wave_count <= WT_MAX_WAVES &&
              (uint64_t)p.wave_offset + p.wave_count <= (uint64_t)INT32_MAX &&
              p.vertex_offset <= kMaxBatchVertexOffset && std::isfinite(p.waveAnimationTimeInSeconds);
    for (int r = 0; r < 4 && ok; ++r)
      for (int k = 0; k < 4; ++k) ok = ok && std::isfinite(p.corners[r][k]);
    const float fs[8] = {p.cameraPosition[0], p.cameraPosition[1], p.cameraPosition[2], p.heightmapSampleScale,
                         p.heightmapDisplacementScale, p.verticalBounceWaveDisplacement,
                         p.heightmapScrollVelocity[0], p.heightmapScrollVelocity[1]};
    for (float f : fs) ok = ok && std::isfinite(f);
    if (!ok) {
      delete b;
      return fail(ctx, WT_ERR_INVALID_ARG, tag + "gridLength in [2, 32768], wave_count <= WT_MAX_WAVES, "
                                                 "wave_offset + wave_count <= 2^31 - 1, vertex_offset <= 2^48, "
                                                 "finite values");
    }
    wt::PatchDesc d{};
    wt::set_corner_algebra(d, p.corners);
    d.active = 1;
    d.inv_nm1 = 1.0 / (double)(p.gridLength - 1);
    d.time = p.waveAnimationTimeInSeconds;
    d.cam[0] = p.cameraPosition[0]; d.cam[1] = p.cameraPosition[1]; d.cam[2] = p.cameraPosition[2];
    d.s_h = p.heightmapSampleScale;
    d.s_d = p.heightmapDisplacementScale;
    d.bounce = p.verticalBounceWaveDisplacement;
    d.scroll_x = p.heightmapScrollVelocity[0];
    d.scroll_z = p.heightmapScrollVelocity[1];
    d.N = (int)p.gridLength;
    d.W = (int)p.wave_count;
    d.wave_offset = (int)p.wave_offset;
    d.vertex_offset = (long long)p.vertex_offset;
    if (forced_shape >= 0) {
      d.shape = forced_shape;
    } else {
      auto it = ctx->batch_table.find({bucket_log2(p.gridLength), bucket_log2(p.wave_count ? p.wave_count : 1)});
      d.shape = it != ctx->batch_table.end() ? it->second : wt::batch_shape_for(d.N);
    }
    b->patches.push_back(d);
    b->vertices = std::max<uint64_t>(b->vertices, p.vertex_offset + (uint64_t)p.gridLength * p.gridLength);
    b->max_wave_end = std::max<uint32_t>(b->max_wave_end, p.wave_offset + p.wave_count);
    max_n = std::max(max_n, p.gridLength);
  }
  // units: one workgroup's segments of one patch, in that patch's segment shape
  std::vector<wt::BatchUnit> units;
  for (uint32_t i = 0; i < count; ++i) {
    const wt::PatchDesc& d = b->patches[i];
    const wt::LaunchShape& s = wt::kBatchShape[d.shape];
    const long long groups = wt::forward_segment_groups(d.N, s);
    for (long long gidx = 0; gidx < groups; ++gidx) units.push_back(wt::BatchUnit{(int)i, (int)gidx});
  }
  // grouped by shape (one launch each), most expensive units first within a shape: a
  // unit costs about its patch's wave count
  auto shape_of = [&](const wt::BatchUnit& x) { return b->patches[x.patch].shape; };
  std::stable_sort(units.begin(), units.end(), [&](const wt::BatchUnit& x, const wt::BatchUnit& y) {
    if (shape_of(x) != shape_of(y)) return shape_of(x) < shape_of(y);
    return b->patches[x.patch].W > b->patches[y.patch].W;
  });
  if (units.size() >= (1ull << 31)) {
    delete b;
    return fail(ctx, WT_ERR_UNSUPPORTED, "batch too large: more than 2^31 workgroup units");
  }
  for (size_t k = 0; k < units.size(); ++k) {
    const int sh = shape_of(units[k]);
    if (b->shape_units[sh]++ == 0) b->shape_begin[sh] = (long long)k;
  }
  hipError_t e = hipMalloc(&b->d_patches, sizeof(wt::PatchDesc) * count);
  if (e == hipSuccess) e = hipMalloc(&b->d_units, sizeof(wt::BatchUnit) * units.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->d_patches, b->patches.data(), sizeof(wt::PatchDesc) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_units, units.data(), sizeof(wt::BatchUnit) * units.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    wt_status s = hip_fail(ctx, e, "wt_batch_create");
    wt_batch_destroy(b);
    return s;
  }
  *out = b;
  return WT_OK;
}

extern "C" {

wt_status wt_batch_create(wt_ctx* ctx, const wt_patch* patches, uint32_t count, wt_batch** out) {
  return batch_create(ctx, patches, count, -1, out);
}

void wt_batch_destroy(wt_batch* b) {
  if (!b) return;
  if (b->d_patches) (void)hipFree(b->d_patches);
  if (b->d_units) (void)hipFree(b->d_units);
  delete b;
}

wt_status wt_batch_vertex_count(const wt_batch* b, uint64_t* vertices) {
  if (!b || !vertices) return WT_ERR_INVALID_ARG;
  *vertices = b->vertices;
  return WT_OK;
}

wt_status wt_batch_patch_launch_config(const wt_batch* b, uint32_t patch, wt_launch_config* cfg) {
  if (!b || !cfg || patch >= b->patches.size()) return WT_ERR_INVALID_ARG;
  const wt::LaunchShape& s = wt::kBatchShape[b->patches[patch].shape];
  cfg->wave_split = (uint32_t)s.split;
  cfg->steps_per_lane = (uint32_t)s.steps;
  cfg->lanes_per_row = (uint32_t)s.lanes_per_row;
  cfg->wave_chunk = (uint32_t)s.wave_chunk;
  return WT_OK;
}

wt_status wt_autotune_batch(wt_ctx* ctx, uint32_t N, uint32_t W, int with_hm, uint32_t* best_shape, float* best_ms) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (N < 2 || N > 8192 || W > WT_MAX_WAVES) return fail(ctx, WT_ERR_INVALID_ARG, "autotune size out of range");
  // enough patches of side N for about 4 M vertices: every shape then has thousands of
  // workgroup units, as in a real batch
  const uint64_t V1 = (uint64_t)N * N;
  const uint32_t P = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (4ull << 20) / V1));
  const uint64_t V = V1 * P;
  float *pos = nullptr, *nrm = nullptr, *tex = nullptr;
  wt_gerstner_wave* dw = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  wt_batch* b = nullptr;
  const uint32_t TW = 256;
  std::vector<wt_gerstner_wave> hw(W ? W : 1);
  for (uint32_t i = 0; i < W; ++i) {  // the forward autotuner's synthetic wave set
    const float L = 1.0f + 59.0f * (float)((i * 37u) % 101u) / 100.0f;
    const float ang = 0.7f * std::sin(1.7f * (float)i);
    hw[i] = wt_gerstner_wave{L / 90.0f, 6.2831853f / L, std::sqrt(9.81f * 6.2831853f / L), 0.5f / (float)W,
                             {std::cos(ang), std::sin(ang)}, 0.3f * (float)i, 0.0f};
  }
  std::vector<float> htex(TW * TW);
  for (uint32_t i = 0; i < TW * TW; ++i) htex[i] = 0.5f + 0.5f * std::sin(0.01f * (float)i);
  std::vector<wt_patch> patches(P);
  const double cc[4][4] = {{-60, 0, -10, 1}, {-900, 0, -1000, 1}, {60, 0, -10, 1}, {900, 0, -1000, 1}};
  for (uint32_t i = 0; i < P; ++i) {
    wt_patch& q = patches[i];
    q = wt_patch{};
    std::memcpy(q.corners, cc, sizeof(cc));
    q.cameraPosition[1] = 10.0f;
    q.gridLength = N;
    q.wave_offset = 0;
    q.wave_count = W;
    q.vertex_offset = (uint64_t)i * V1;
    q.waveAnimationTimeInSeconds = 10.0 + 0.37 * i;
    q.heightmapSampleScale = 0.02f;
    q.heightmapDisplacementScale = 0.05f;
  }
  wt_status rs = WT_OK;
  auto cleanup = [&]() {
    if (b) wt_batch_destroy(b);
    if (pos) (void)hipFree(pos);
    if (nrm) (void)hipFree(nrm);
    if (tex) (void)hipFree(tex);
    if (dw) (void)hipFree(dw);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
#define WT_TRY(x)                 \
  do {                            \
    hipError_t _e = (x);          \
    if (_e != hipSuccess) {       \
      rs = hip_fail(ctx, _e, #x); \
      cleanup();                  \
      return rs;                  \
    }                             \
  } while (0)
  WT_TRY(hipMalloc(&pos, V * 12));
  WT_TRY(hipMalloc(&nrm, V * 12));
  WT_TRY(hipMalloc(&dw, sizeof(wt_gerstner_wave) * hw.size()));
  WT_TRY(hipMalloc(&tex, sizeof(float) * htex.size()));
  WT_TRY(hipMemcpy(dw, hw.data(), sizeof(wt_gerstner_wave) * hw.size(), hipMemcpyHostToDevice));
  WT_TRY(hipMemcpy(tex, htex.data(), sizeof(float) * htex.size(), hipMemcpyHostToDevice));
  WT_TRY(hipStreamCreate(&st));
  WT_TRY(hipEventCreate(&e0));
  WT_TRY(hipEventCreate(&e1));
  const wt_heightmap hm{tex, TW, TW, WT_TEXEL_R32F, WT_WRAP_REPEAT};
  const wt_vertex_buffers vb{pos, nrm, nullptr, nullptr};
  float best_t = 1e30f;
  int best_s = wt::batch_shape_for((int)N);
  for (int sh = 0; sh < wt::kBatchShapes; ++sh) {
    rs = batch_create(ctx, patches.data(), P, sh, &b);
    if (rs != WT_OK) {
      cleanup();
      return rs;
    }
    rs = wt_water_displace_batched(ctx, b, dw, W, with_hm ? &hm : nullptr, &vb, 0.0, st);  // warm-up
    if (rs != WT_OK) {
      cleanup();
      return rs;
    }
    const int reps = 5;
    WT_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) (void)wt_water_displace_batched(ctx, b, dw, W, with_hm ? &hm : nullptr, &vb, r / 60.0, st);
    WT_TRY(hipEventRecord(e1, st));
    WT_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    WT_TRY(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    if (ms < best_t) {
      best_t = ms;
      best_s = sh;
    }
    wt_batch_destroy(b);
    b = nullptr;
  }
#undef WT_TRY
  ctx->batch_table[{bucket_log2(N), bucket_log2(W ? W : 1)}] = best_s;
  if (best_shape) *best_shape = (uint32_t)best_s;
  if (best_ms) *best_ms = best_t;
  cleanup();
  return WT_OK;
}

}  // extern "C"

namespace {

// Shared checks of the two device-projector entry points: 4-byte aligned device arrays
// (device memory is only checked outside stream capture, like the other entry points).
wt_status projector_args(wt_ctx* ctx, const wt_camera* cams, const float* amps, const void* corners,
                         const int32_t* status, hipStream_t st) {
  if (!cams || !amps) return fail(ctx, WT_ERR_INVALID_ARG, "cameras and amplitudes are required");
  const void* ptrs[4] = {cams, amps, corners, status};
  for (const void* p : ptrs)
    if (p && (reinterpret_cast<uintptr_t>(p) & 3)) return fail(ctx, WT_ERR_INVALID_ARG, "arrays must be 4-byte aligned");
  if (corners && (reinterpret_cast<uintptr_t>(corners) & 7))
    return fail(ctx, WT_ERR_INVALID_ARG, "corners must be 8-byte aligned");
  if (!stream_capturing(st))
    for (const void* p : ptrs)
      if (p && !is_device_pointer(p)) return fail(ctx, WT_ERR_INVALID_ARG, "arrays must be device memory");
  return WT_OK;
}

}  // namespace

extern "C" {

wt_status wt_projector_corners_batched(wt_ctx* ctx, const wt_camera* d_cameras, const float* d_amplitudes,
                                       uint32_t count, double* d_corners, int32_t* d_status, void* stream) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (count == 0) return WT_OK;
  if (count > WT_MAX_CAMERAS) return fail(ctx, WT_ERR_INVALID_ARG, "count exceeds WT_MAX_CAMERAS");
  if (!d_corners || !d_status) return fail(ctx, WT_ERR_INVALID_ARG, "corners and status are required");
  hipStream_t st = (hipStream_t)stream;
  wt_status s = projector_args(ctx, d_cameras, d_amplitudes, d_corners, d_status, st);
  if (s != WT_OK) return s;
  wt::ProjectorArgs a{d_cameras, d_amplitudes, (int)count, d_corners, d_status, nullptr};
  hipError_t e = wt::launch_projector(a, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "projector_kernel launch");
  return WT_OK;
}

wt_status wt_batch_set_cameras(wt_ctx* ctx, wt_batch* b, const wt_camera* d_cameras, const float* d_amplitudes,
                               int32_t* d_status, void* stream) {
  if (!ctx || !b) return WT_ERR_INVALID_ARG;
  hipStream_t st = (hipStream_t)stream;
  wt_status s = projector_args(ctx, d_cameras, d_amplitudes, nullptr, d_status, st);
  if (s != WT_OK) return s;
  wt::ProjectorArgs a{d_cameras, d_amplitudes, (int)b->patches.size(), nullptr, d_status, b->d_patches};
  hipError_t e = wt::launch_projector(a, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "projector_kernel launch");
  return WT_OK;
}

wt_status wt_grid_triangle_indices(wt_ctx* ctx, uint32_t grid_length, uint32_t base_vertex, uint32_t* d_indices,
                                   void* stream) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (grid_length < 2 || grid_length > 32768) return fail(ctx, WT_ERR_INVALID_ARG, "gridLength must be in [2, 32768]");
  if ((uint64_t)base_vertex + (uint64_t)grid_length * grid_length > (1ull << 32))
    return fail(ctx, WT_ERR_INVALID_ARG, "base_vertex + gridLength^2 exceeds the 32-bit index range");
  if (!d_indices || (reinterpret_cast<uintptr_t>(d_indices) & 3))
    return fail(ctx, WT_ERR_INVALID_ARG, "indices must be a 4-byte aligned device array");
  hipStream_t st = (hipStream_t)stream;
  if (!stream_capturing(st) && !is_device_pointer(d_indices))
    return fail(ctx, WT_ERR_INVALID_ARG, "indices must be device memory");
  hipError_t e = wt::launch_grid_indices(grid_length, base_vertex, d_indices, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "grid_indices_kernel launch");
  return WT_OK;
}

static bool side_ok(uint32_t v) { return v >= 1 && v <= 32768; }

wt_status wt_water_shade(wt_ctx* ctx, const wt_shade_params* p, const wt_shade_textures* tx,
                         const wt_water_fragment* fragments, const uint8_t* coverage, uint32_t width, uint32_t height,
                         float* colour, int blend, void* stream) {
  static_assert(sizeof(wt_water_fragment) == 48, "the kernel reads a fragment as three float4");
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (!p || !tx || !fragments || !colour) return fail(ctx, WT_ERR_INVALID_ARG, "params, textures, fragments and colour must be non-NULL");
  if (!side_ok(width) || !side_ok(height)) return fail(ctx, WT_ERR_INVALID_ARG, "width and height must be in [1, 32768]");
  if (blend != 0 && blend != 1) return fail(ctx, WT_ERR_INVALID_ARG, "blend must be 0 or 1");
  const float fs[20] = {p->fogColourFarAtCurrentTime[0], p->fogColourFarAtCurrentTime[1], p->fogColourFarAtCurrentTime[2],
                        p->fogColourFarAtCurrentTime[3], p->fogDepthRadiusFar, p->fogDepthRadiusNear,
                        p->softEdgesDeltaDepthThreshold, p->sunPosition[0], p->sunPosition[1], p->sunPosition[2],
                        p->sunShininess, p->sunStrength, p->tintDeltaDepthThreshold, p->viewportWidthHeight[0],
                        p->viewportWidthHeight[1], p->waterClarity, p->zFar, p->zNear, 0.0f, 0.0f};
  for (float v : fs)
    if (!std::isfinite(v)) return fail(ctx, WT_ERR_INVALID_ARG, "uniforms must be finite");
  if (!(p->zFar > p->zNear && p->zNear > 0.0f)) return fail(ctx, WT_ERR_INVALID_ARG, "need zFar > zNear > 0");
  if (!(p->viewportWidthHeight[0] > 0.0f && p->viewportWidthHeight[1] > 0.0f))
    return fail(ctx, WT_ERR_INVALID_ARG, "viewportWidthHeight must be positive");
  if (p->sunShininess < 0.0f) return fail(ctx, WT_ERR_INVALID_ARG, "sunShininess must be >= 0");
  if (!tx->depth || !tx->local_reflections.texels || !tx->local_refractions.texels || !tx->skybox)
    return fail(ctx, WT_ERR_INVALID_ARG, "every texture is required");
  if (!side_ok(tx->depth_width) || !side_ok(tx->depth_height) || !side_ok(tx->local_reflections.width) ||
      !side_ok(tx->local_reflections.height) || !side_ok(tx->local_refractions.width) ||
      !side_ok(tx->local_refractions.height) || !side_ok(tx->skybox_length))
    return fail(ctx, WT_ERR_INVALID_ARG, "texture sides must be in [1, 32768]");
  if ((reinterpret_cast<uintptr_t>(fragments) & 15) || (reinterpret_cast<uintptr_t>(colour) & 15))
    return fail(ctx, WT_ERR_INVALID_ARG, "fragments and colour must be 16-byte aligned");
  const void* words[4] = {tx->depth, tx->local_reflections.texels, tx->local_refractions.texels, tx->skybox};
  for (const void* b : words)
    if (reinterpret_cast<uintptr_t>(b) & 3) return fail(ctx, WT_ERR_INVALID_ARG, "textures must be 4-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  if (!stream_capturing(st)) {
    const void* bufs[7] = {fragments, colour, tx->depth, tx->local_reflections.texels, tx->local_refractions.texels,
                           tx->skybox, coverage};
    for (const void* b : bufs)
      if (b && !is_device_pointer(b)) return fail(ctx, WT_ERR_INVALID_ARG, "buffers must be device memory");
  }
  // FP32 uniforms and the reciprocals the kernel multiplies by in place of the shader's
  // divisions (a reciprocal-multiply is within GLSL's 2.5 ulp for division)
  wt::ShadeArgs a{};
  for (int k = 0; k < 4; ++k) a.fog[k] = p->fogColourFarAtCurrentTime[k];
  a.fog_far = p->fogDepthRadiusFar;
  a.fog_near = p->fogDepthRadiusNear;
  a.inv_fog_range = p->fogDepthRadiusFar > p->fogDepthRadiusNear
                        ? 1.0f / (p->fogDepthRadiusFar - p->fogDepthRadiusNear) : 0.0f;
  a.tint_off = p->tintDeltaDepthThreshold == 0.0f;
  a.inv_tint = a.tint_off ? 0.0f : 1.0f / p->tintDeltaDepthThreshold;
  a.soft_off = p->softEdgesDeltaDepthThreshold == 0.0f;
  a.inv_soft = a.soft_off ? 0.0f : 1.0f / p->softEdgesDeltaDepthThreshold;
  for (int k = 0; k < 3; ++k) a.sun[k] = p->sunPosition[k];
  a.shininess = p->sunShininess;
  a.strength = p->sunStrength;
  a.clarity = p->waterClarity;
  a.zfar = p->zFar;
  a.znear = p->zNear;
  a.zfar_m_znear = p->zFar - p->zNear;
  a.inv_zfar = 1.0f / p->zFar;
  a.inv_viewport_w = 1.0f / p->viewportWidthHeight[0];
  a.inv_viewport_h = 1.0f / p->viewportWidthHeight[1];
  a.depth = tx->depth;
  a.depth_w = (int)tx->depth_width;
  a.depth_h = (int)tx->depth_height;
  a.refl = reinterpret_cast<const unsigned*>(tx->local_reflections.texels);
  a.refl_w = (int)tx->local_reflections.width;
  a.refl_h = (int)tx->local_reflections.height;
  a.refr = reinterpret_cast<const unsigned*>(tx->local_refractions.texels);
  a.refr_w = (int)tx->local_refractions.width;
  a.refr_h = (int)tx->local_refractions.height;
  a.sky = reinterpret_cast<const unsigned*>(tx->skybox);
  a.sky_len = (int)tx->skybox_length;
  a.frags = reinterpret_cast<const float4*>(fragments);
  a.coverage = coverage;
  a.colour = reinterpret_cast<float4*>(colour);
  a.width = (int)width;
  a.count = (int)(width * height);
  a.blend = blend;
  hipError_t e = wt::launch_shade(a, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "water_shade_kernel launch");
  return WT_OK;
}

wt_status wt_water_displace_batched(wt_ctx* ctx, wt_batch* b, const wt_gerstner_wave* d_waves, uint32_t total_waves,
                                    const wt_heightmap* hm, const wt_vertex_buffers* out, double time_offset,
                                    void* stream) {
  if (!ctx || !b || !out) return WT_ERR_INVALID_ARG;
  if (total_waves < b->max_wave_end) return fail(ctx, WT_ERR_INVALID_ARG, "a patch indexes past total_waves");
  if (total_waves && !d_waves) return fail(ctx, WT_ERR_INVALID_ARG, "d_waves is NULL");
  if (!out->position || !out->normal) return fail(ctx, WT_ERR_INVALID_ARG, "position and normal buffers are required");
  if (!std::isfinite(time_offset)) return fail(ctx, WT_ERR_INVALID_ARG, "time_offset must be finite");
  hipStream_t st = (hipStream_t)stream;
  const bool capturing = stream_capturing(st);
  wt_status hs = wt::validate_heightmap(ctx, hm, !capturing);
  if (hs != WT_OK) return hs;
  const void* bufs[4] = {out->position, out->normal, out->tangent, out->bitangent};
  for (const void* p : bufs)
    if (p && (reinterpret_cast<uintptr_t>(p) & 3)) return fail(ctx, WT_ERR_INVALID_ARG, "output buffers must be 4-byte aligned");
  if (!capturing) {
    for (const void* p : bufs)
      if (p && !is_device_pointer(p)) return fail(ctx, WT_ERR_INVALID_ARG, "output buffers must be device memory");
    if (total_waves && !is_device_pointer(d_waves)) return fail(ctx, WT_ERR_INVALID_ARG, "d_waves must be device memory");
  }
  wt::BatchArgs a{};
  a.patches = b->d_patches;
  a.units = b->d_units;
  a.time_offset = time_offset;
  a.waves = d_waves;
  if (hm) {
    a.texels = hm->texels;
    a.tex_w = (int)hm->width;
    a.tex_h = (int)hm->height;
    a.tex_fmt = (int)hm->format;
    a.tex_wrap = (int)hm->wrap;
  }
  a.pos = out->position;
  a.nrm = out->normal;
  a.tan = out->tangent;
  a.bit = out->bitangent;
  for (int sh = 0; sh < wt::kBatchShapes; ++sh) {
    if (b->shape_units[sh] == 0) continue;
    a.unit_begin = b->shape_begin[sh];
    a.n_units = b->shape_units[sh];
    hipError_t e = wt::launch_forward_batched(a, sh, st);
    if (e != hipSuccess) return hip_fail(ctx, e, "water_forward_batched_kernel launch");
  }
  return WT_OK;
}

wt_status wt_graph_begin(wt_ctx* ctx, void* stream) {
  if (!ctx) return WT_ERR_INVALID_ARG;
  if (!stream) return fail(ctx, WT_ERR_INVALID_ARG, "graph capture needs a non-default stream");
  hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
  return WT_OK;
}

wt_status wt_graph_end(wt_ctx* ctx, void* stream, wt_graph** out) {
  if (!ctx || !out) return WT_ERR_INVALID_ARG;
  *out = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(ctx, e, "hipGraphInstantiate");
  }
  wt_graph* r = new (std::nothrow) wt_graph();
  if (!r) {
    (void)hipGraphExecDestroy(x);
    (void)hipGraphDestroy(g);
    return fail(ctx, WT_ERR_OOM, "graph handle");
  }
  r->graph = g;
  r->exec = x;
  *out = r;
  return WT_OK;
}

wt_status wt_graph_launch(wt_graph* g, void* stream) {
  if (!g) return WT_ERR_INVALID_ARG;
  return hipGraphLaunch(g->exec, (hipStream_t)stream) == hipSuccess ? WT_OK : WT_ERR_HIP;
}

void wt_graph_destroy(wt_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

}  // extern "C"
