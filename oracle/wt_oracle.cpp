// C entry points of the ORACLE (test infrastructure only; see wt_oracle.hpp).
// Loaded through ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg. The float instantiation run over std::threads is the CPU
// baseline; the double instantiation is the parity oracle.
#include "wt_oracle.hpp"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {

typedef struct {
  double corners[16];  // BL, TL, BR, TR as (x, y, z, w)
  double camera[3];
  int32_t N;
  int32_t W;
  double sample_scale, displacement_scale, bounce, scroll_x, scroll_z, time;
  const float* waves;  // W x 8 floats {A, w, phi, Q, Dx, Dz, phi0, pad}
  const void* texels;  // null: no heightmap
  int32_t tex_w, tex_h, tex_format, tex_wrap;
} wto_scene;

typedef struct {
  int32_t n;
  int32_t S;
  double spacing, x0, z0, t0, dt, g;
  const float* target;  // [S][n*n][3]
} wto_fit;

int wto_abi_version(void) { return 3; }

static wto::Camera<double> make_camera(const double c[9]) {
  return wto::Camera<double>(c[0], c[1], c[2], c[3], wto::V3<double>{c[4], c[5], c[6]}, c[7], c[8]);
}

// Camera basis and matrices (src/camera.cpp:130-146). cam = {fov_deg, aspect, near,
// far, px, py, pz, yaw_deg, pitch_deg}; the camera is built with the default yaw and
// pitch and then setRotation(yaw, pitch) is applied, as a caller would.
int wto_camera(const double cam[9], double view[16], double proj[16], double forward[3], double right[3],
               double up[3]) {
  wto::Camera<double> c = make_camera(cam);
  c.setRotation(cam[7], cam[8]);
  for (int i = 0; i < 16; ++i) {
    view[i] = c.view.m[i / 4][i % 4];
    proj[i] = c.proj.m[i / 4][i % 4];
  }
  forward[0] = c.forward.x; forward[1] = c.forward.y; forward[2] = c.forward.z;
  right[0] = c.right.x; right[1] = c.right.y; right[2] = c.right.z;
  up[0] = c.up.x; up[1] = c.up.y; up[2] = c.up.z;
  return 0;
}

// General 4x4 inverse (column-major in and out).
void wto_inverse(const double m[16], double out[16]) {
  wto::M4<double> a;
  for (int i = 0; i < 16; ++i) a.m[i / 4][i % 4] = m[i];
  wto::M4<double> r = wto::inverse(a);
  for (int i = 0; i < 16; ++i) out[i] = r.m[i / 4][i % 4];
}

// Projector corners (src/render-engine.cpp:797-1042). info (13 doubles):
// xmin, xmax, ymin, ymax, projector pos xyz, projector pitch, aim xyz, raised, npoints.
// slab_pts / ndc_pts (may be null) receive up to max_pts points as (x, y, z, w).
// pad / swap_aim_mix: the padding scalar and aim-mix order (1.2 and 0 restate the
// reference; other values plant a mistake for the tests that show a pin catches it).
int wto_projector_variant(const double cam[9], double amp, double pad, int swap_aim_mix, double corners[16],
                          double info[13], double* slab_pts, double* ndc_pts, int max_pts) {
  wto::Camera<double> c = make_camera(cam);
  c.setRotation(cam[7], cam[8]);
  wto::V4<double> co[4];
  wto::ProjectorDebug<double> dbg{};
  wto::ProjectorMutation<double> mut;
  mut.pad = pad;
  mut.swapAimMix = swap_aim_mix != 0;
  int st = wto::projectorCorners(c, amp, co, &dbg, &mut);
  if (info) {
    for (int i = 0; i < 13; ++i) info[i] = 0;
    info[12] = (double)dbg.slabPoints.size();
  }
  if (slab_pts)
    for (int i = 0; i < (int)dbg.slabPoints.size() && i < max_pts; ++i) {
      slab_pts[4 * i] = dbg.slabPoints[i].x; slab_pts[4 * i + 1] = dbg.slabPoints[i].y;
      slab_pts[4 * i + 2] = dbg.slabPoints[i].z; slab_pts[4 * i + 3] = dbg.slabPoints[i].w;
    }
  if (st != wto::WTO_OK) return st;
  for (int i = 0; i < 4; ++i) {
    corners[4 * i] = co[i].x; corners[4 * i + 1] = co[i].y; corners[4 * i + 2] = co[i].z; corners[4 * i + 3] = co[i].w;
  }
  if (info) {
    info[0] = dbg.xmin; info[1] = dbg.xmax; info[2] = dbg.ymin; info[3] = dbg.ymax;
    info[4] = dbg.projectorPosition.x; info[5] = dbg.projectorPosition.y; info[6] = dbg.projectorPosition.z;
    info[7] = dbg.projectorPitch;
    info[8] = dbg.aimpoint.x; info[9] = dbg.aimpoint.y; info[10] = dbg.aimpoint.z;
    info[11] = dbg.raised ? 1.0 : 0.0;
  }
  if (ndc_pts)
    for (int i = 0; i < (int)dbg.ndcPoints.size() && i < max_pts; ++i) {
      ndc_pts[4 * i] = dbg.ndcPoints[i].x; ndc_pts[4 * i + 1] = dbg.ndcPoints[i].y;
      ndc_pts[4 * i + 2] = dbg.ndcPoints[i].z; ndc_pts[4 * i + 3] = dbg.ndcPoints[i].w;
    }
  return 0;
}

int wto_projector(const double cam[9], double amp, double corners[16], double info[13], double* slab_pts,
                  double* ndc_pts, int max_pts) {
  return wto_projector_variant(cam, amp, 1.2, 0, corners, info, slab_pts, ndc_pts, max_pts);
}

long long wto_wrap_index(long long i, long long size, int wrap) { return wto::wrapIndex(i, size, wrap); }

double wto_sample_bilinear(const void* texels, int w, int h, int format, int wrap, double s, double t, double* dds,
                           double* ddt) {
  wto::Texture tx;
  tx.texels = texels; tx.width = w; tx.height = h; tx.format = format; tx.wrap = wrap;
  return wto::sampleBilinear<double>(tx, s, t, dds, ddt);
}

}  // extern "C"

template <class T> struct SceneHolder {
  std::vector<wto::Wave<T>> waves;
  wto::Texture tex;
  wto::Scene<T> sc;
  explicit SceneHolder(const wto_scene* s) {
    waves.resize(s->W > 0 ? s->W : 0);
    for (int i = 0; i < s->W; ++i) {
      const float* p = s->waves + 8 * i;
      waves[i] = {T(p[0]), T(p[1]), T(p[2]), T(p[3]), T(p[4]), T(p[5]), T(p[6])};
    }
    tex.texels = s->texels; tex.width = s->tex_w; tex.height = s->tex_h;
    tex.format = s->tex_format; tex.wrap = s->tex_wrap;
    for (int i = 0; i < 4; ++i)
      sc.corners[i] = {T(s->corners[4 * i]), T(s->corners[4 * i + 1]), T(s->corners[4 * i + 2]), T(s->corners[4 * i + 3])};
    sc.cameraPosition = {T(s->camera[0]), T(s->camera[1]), T(s->camera[2])};
    sc.N = s->N;
    sc.sampleScale = T(s->sample_scale);
    sc.displacementScale = T(s->displacement_scale);
    sc.bounce = T(s->bounce);
    sc.scrollX = T(s->scroll_x);
    sc.scrollZ = T(s->scroll_z);
    sc.time = T(s->time);
    sc.waves = waves.data();
    sc.W = s->W;
    sc.tex = s->texels ? &tex : nullptr;
  }
};

template <class T, class O>
static int run_vertices(const wto_scene* s, int mode, long long first, long long count, O* pos, O* nrm, O* tan, O* bit,
                        int nthreads) {
  if (!s || s->N < 2 || count < 0) return 1;
  SceneHolder<T> h(s);
  auto work = [&](long long lo, long long hi) {
    for (long long k = lo; k < hi; ++k) {
      long long id = first + k;
      wto::VertexOut<T> o = mode == 0 ? wto::vertexAnalytic<T>(h.sc, id) : wto::vertexReference<T>(h.sc, id);
      if (pos) { pos[3 * k] = O(o.pos.x); pos[3 * k + 1] = O(o.pos.y); pos[3 * k + 2] = O(o.pos.z); }
      if (nrm) { nrm[3 * k] = O(o.nrm.x); nrm[3 * k + 1] = O(o.nrm.y); nrm[3 * k + 2] = O(o.nrm.z); }
      if (tan) { tan[3 * k] = O(o.tan.x); tan[3 * k + 1] = O(o.tan.y); tan[3 * k + 2] = O(o.tan.z); }
      if (bit) { bit[3 * k] = O(o.bit.x); bit[3 * k + 1] = O(o.bit.y); bit[3 * k + 2] = O(o.bit.z); }
    }
  };
  if (nthreads <= 1 || count < 1024) {
    work(0, count);
    return 0;
  }
  std::vector<std::thread> th;
  long long chunk = (count + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    long long lo = t * chunk, hi = std::min(count, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
  return 0;
}

extern "C" {

// mode 0 = analytic normals (the build's contract), 1 = the reference's finite-
// difference normals. Outputs are count x 3 arrays (each may be null).
int wto_vertices_f64(const wto_scene* s, int mode, long long first, long long count, double* pos, double* nrm,
                     double* tan, double* bit, int nthreads) {
  return run_vertices<double, double>(s, mode, first, count, pos, nrm, tan, bit, nthreads);
}
int wto_vertices_f32(const wto_scene* s, int mode, long long first, long long count, float* pos, float* nrm,
                     float* tan, float* bit, int nthreads) {
  return run_vertices<float, float>(s, mode, first, count, pos, nrm, tan, bit, nthreads);
}
int wto_vertices_f80(const wto_scene* s, int mode, long long first, long long count, double* pos, double* nrm,
                     double* tan, double* bit, int nthreads) {
  return run_vertices<long double, double>(s, mode, first, count, pos, nrm, tan, bit, nthreads);
}

// Gerstner sum and analytic partials at one world point (x, z) (no heightmap).
void wto_gerstner_point(const wto_scene* s, double x, double z, double P[3], double dPdx[3], double dPdz[3]) {
  SceneHolder<double> h(s);
  wto::GerstnerOut<double> g = wto::gerstner(h.sc, x, z);
  P[0] = g.P.x; P[1] = g.P.y; P[2] = g.P.z;
  dPdx[0] = g.dPdx.x; dPdx[1] = g.dPdx.y; dPdx[2] = g.dPdx.z;
  dPdz[0] = g.dPdz.x; dPdz[1] = g.dPdz.y; dPdz[2] = g.dPdz.z;
}

// Displaced position (Gerstner + heightmap + bounce) of world point (x, z).
void wto_displaced_point(const wto_scene* s, double x, double z, double P[3]) {
  SceneHolder<double> h(s);
  wto::V3<double> p = wto::displacedPosition(h.sc, x, z);
  P[0] = p.x; P[1] = p.y; P[2] = p.z;
}

// FP64 heightfield-fit loss and gradient of the mean misfit (1/(2 S V)) sum |r|^2
// (W x 5 params {A, L, alpha, Q, phi0}). The vertex range is split over threads and
// the partial sums are added in thread order, so the result does not depend on timing.
double wto_fit_loss_grad_f64(const wto_fit* f, const double* params, int W, double* grad, int nthreads) {
  const long long V = (long long)f->n * f->n;
  if (nthreads < 1) nthreads = 1;
  wto::FitProblem<double> p;
  p.n = f->n; p.spacing = f->spacing; p.x0 = f->x0; p.z0 = f->z0; p.S = f->S;
  p.t0 = f->t0; p.dt = f->dt; p.g = f->g; p.target = f->target;
  std::vector<double> losses(nthreads, 0.0);
  std::vector<std::vector<double>> grads(nthreads, std::vector<double>(5 * W, 0.0));
  auto work = [&](int t, long long lo, long long hi) {
    losses[t] = wto::fitLossGradRange<double>(p, params, W, grad ? grads[t].data() : nullptr, lo, hi);
  };
  std::vector<std::thread> th;
  long long chunk = (V + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    long long lo = t * chunk, hi = std::min(V, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(work, t, lo, hi);
  }
  for (auto& x : th) x.join();
  const double inv = 1.0 / (double(f->S) * double(V));
  double loss = 0.0;
  for (int t = 0; t < nthreads; ++t) loss += losses[t];
  if (grad)
    for (int k = 0; k < 5 * W; ++k) {
      double s = 0.0;
      for (int t = 0; t < nthreads; ++t) s += grads[t][k];
      grad[k] = s * inv;
    }
  return loss * inv;
}

// FP32 instantiation of the same fit loss (used as the timed CPU adjoint baseline).
double wto_fit_loss_grad_f32(const wto_fit* f, const double* params, int W, double* grad, int nthreads) {
  const long long V = (long long)f->n * f->n;
  if (nthreads < 1) nthreads = 1;
  wto::FitProblem<float> p;
  p.n = f->n; p.spacing = (float)f->spacing; p.x0 = (float)f->x0; p.z0 = (float)f->z0; p.S = f->S;
  p.t0 = (float)f->t0; p.dt = (float)f->dt; p.g = (float)f->g; p.target = f->target;
  std::vector<float> pf(5 * W);
  for (int k = 0; k < 5 * W; ++k) pf[k] = (float)params[k];
  std::vector<double> losses(nthreads, 0.0);
  std::vector<std::vector<float>> grads(nthreads, std::vector<float>(5 * W, 0.0f));
  auto work = [&](int t, long long lo, long long hi) {
    losses[t] = wto::fitLossGradRange<float>(p, pf.data(), W, grad ? grads[t].data() : nullptr, lo, hi);
  };
  std::vector<std::thread> th;
  long long chunk = (V + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    long long lo = t * chunk, hi = std::min(V, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(work, t, lo, hi);
  }
  for (auto& x : th) x.join();
  const double inv = 1.0 / (double(f->S) * double(V));
  double loss = 0.0;
  for (int t = 0; t < nthreads; ++t) loss += losses[t];
  if (grad)
    for (int k = 0; k < 5 * W; ++k) {
      double s = 0.0;
      for (int t = 0; t < nthreads; ++t) s += grads[t][k];
      grad[k] = s * inv;
    }
  return loss * inv;
}

}  // extern "C"
