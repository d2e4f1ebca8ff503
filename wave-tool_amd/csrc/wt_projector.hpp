// Camera and projected-grid projector of the water stage, in double precision, compiled
// for both the host (wt_projector_corners, the host water stage) and the device (one
// thread per camera for batched patches).
//
// Camera restates src/camera.cpp:40-146; projector_corners restates the water block
// of RenderEngine::render, src/render-engine.cpp:797-1042, with the reference's
// Line / Plane / linePlaneIntersection (src/render-engine.h:53-124). The matrix
// helpers are written from their textbook definitions (gluPerspective, gluLookAt,
// Gauss-Jordan elimination).
//
// Floating-point contraction is off in this header so the host and device builds
// round every operation alike; they then differ only where the C library and the
// device math library round tan, sin, cos and acos differently.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/wt.h"

#define WT_HD __host__ __device__ inline

#pragma clang fp contract(off)

namespace wt {
namespace proj {

struct Vec3 {
  double x, y, z;
};
struct Vec4 {
  double x, y, z, w;
};
// Column-major 4x4: c[col][row], as GLM stores it.
struct Mat4 {
  double c[4][4];
};

WT_HD Vec3 add(const Vec3& a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
WT_HD Vec3 sub(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
WT_HD Vec3 scale(const Vec3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
WT_HD double dot3(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WT_HD Vec3 cross3(const Vec3& a, const Vec3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
WT_HD Vec3 unit(const Vec3& a) { return scale(a, 1.0 / sqrt(dot3(a, a))); }

WT_HD Mat4 mat_identity() {
  Mat4 m{};
  for (int i = 0; i < 4; ++i) m.c[i][i] = 1.0;
  return m;
}

WT_HD Mat4 mat_mul(const Mat4& a, const Mat4& b) {
  Mat4 r{};
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a.c[k][row] * b.c[col][k];
      r.c[col][row] = s;
    }
  return r;
}

WT_HD Vec4 mat_apply(const Mat4& m, const Vec4& v) {
  const double in[4] = {v.x, v.y, v.z, v.w};
  double o[4];
  for (int row = 0; row < 4; ++row) {
    o[row] = 0.0;
    for (int col = 0; col < 4; ++col) o[row] += m.c[col][row] * in[col];
  }
  return {o[0], o[1], o[2], o[3]};
}

// Gauss-Jordan inverse with partial pivoting; returns false for a singular matrix.
WT_HD bool mat_inverse(const Mat4& m, Mat4* out) {
  // augmented [A | I] in row-major working storage
  double a[4][8];
  for (int row = 0; row < 4; ++row)
    for (int col = 0; col < 4; ++col) {
      a[row][col] = m.c[col][row];
      a[row][4 + col] = row == col ? 1.0 : 0.0;
    }
  for (int p = 0; p < 4; ++p) {
    int best = p;
    double big = fabs(a[p][p]);
    for (int r = p + 1; r < 4; ++r)
      if (fabs(a[r][p]) > big) {
        best = r;
        big = fabs(a[r][p]);
      }
    if (big == 0.0) return false;
    // row swap written with constant row indices so the device keeps `a` in registers
    for (int r = p + 1; r < 4; ++r)
      if (r == best)
        for (int k = 0; k < 8; ++k) {
          const double t = a[p][k];
          a[p][k] = a[r][k];
          a[r][k] = t;
        }
    const double inv = 1.0 / a[p][p];
    for (int k = 0; k < 8; ++k) a[p][k] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == p) continue;
      const double f = a[r][p];
      if (f == 0.0) continue;
      for (int k = 0; k < 8; ++k) a[r][k] -= f * a[p][k];
    }
  }
  for (int row = 0; row < 4; ++row)
    for (int col = 0; col < 4; ++col) out->c[col][row] = a[row][4 + col];
  return true;
}

// gluPerspective (right-handed, clip z in [-1, 1]), as glm::perspective at
// src/camera.cpp:141.
WT_HD Mat4 perspective_rh(double fovy, double aspect, double zn, double zf) {
  Mat4 m{};
  const double f = 1.0 / tan(0.5 * fovy);
  m.c[0][0] = f / aspect;
  m.c[1][1] = f;
  m.c[2][2] = (zf + zn) / (zn - zf);
  m.c[2][3] = -1.0;
  m.c[3][2] = 2.0 * zf * zn / (zn - zf);
  return m;
}

// gluLookAt (right-handed), as glm::lookAt at src/camera.cpp:145.
WT_HD Mat4 look_at_rh(const Vec3& eye, const Vec3& center, const Vec3& up) {
  const Vec3 f = unit(sub(center, eye));
  const Vec3 s = unit(cross3(f, up));
  const Vec3 u = cross3(s, f);
  Mat4 m = mat_identity();
  const Vec3 rows[3] = {s, u, {-f.x, -f.y, -f.z}};
  for (int r = 0; r < 3; ++r) {
    m.c[0][r] = rows[r].x;
    m.c[1][r] = rows[r].y;
    m.c[2][r] = rows[r].z;
  }
  m.c[3][0] = -dot3(s, eye);
  m.c[3][1] = -dot3(u, eye);
  m.c[3][2] = dot3(f, eye);
  return m;
}

constexpr double kPi = 3.14159265358979323846;
WT_HD double radians(double d) { return d * kPi / 180.0; }
// glm::mod(x, y) = x - y * floor(x / y)
WT_HD double glsl_mod(double x, double y) { return x - y * floor(x / y); }

// The reference Camera (src/camera.h / src/camera.cpp) reduced to what the water
// stage reads: position, yaw/pitch basis, projection and view matrices.
struct Camera {
  double fov_deg, aspect, znear, zfar;
  Vec3 position, forward, up, up_init, right;
  double yaw_deg, pitch_deg;
  Mat4 proj, view;

  // Camera ctor (camera.cpp:47-54) with the default forward (0,0,-1), up (0,1,0),
  // yaw -90, pitch 0 (camera.cpp:40-45).
  WT_HD Camera(double fov, double aspect_, double n, double f, const Vec3& pos)
      : fov_deg(fov), aspect(aspect_), znear(n), zfar(f), position(pos), forward{0, 0, -1}, up{0, 1, 0},
        up_init{0, 1, 0}, right{1, 0, 0}, yaw_deg(-90.0), pitch_deg(0.0) {
    update_basis();
    proj = perspective_rh(radians(fov_deg), aspect, znear, zfar);
  }

  // camera.cpp:93-98
  WT_HD void set_rotation(double yaw, double pitch) {
    yaw_deg = glsl_mod(yaw, 360.0);
    pitch_deg = pitch < -89.0 ? -89.0 : (pitch > 89.0 ? 89.0 : pitch);
    update_basis();
  }

  // camera.cpp:100-104
  WT_HD void translate(const Vec3& d) {
    position = add(position, d);
    view = look_at_rh(position, add(position, forward), up);
  }

  // camera.cpp:130-138
  WT_HD void update_basis() {
    const double cy = cos(radians(yaw_deg)), sy = sin(radians(yaw_deg));
    const double cp = cos(radians(pitch_deg)), sp = sin(radians(pitch_deg));
    forward = unit(Vec3{cy * cp, sp, sy * cp});
    right = unit(cross3(forward, up_init));
    up = unit(cross3(right, forward));
    view = look_at_rh(position, add(position, forward), up);
  }
};

// The ABI camera is usable: finite fields, fov in (0, 180) degrees, positive aspect
// and near plane, far beyond near.
WT_HD bool camera_valid(const wt_camera& c) {
  const float v[9] = {c.fov_deg, c.aspect, c.near_clip, c.far_clip, c.position[0],
                      c.position[1], c.position[2], c.yaw_deg, c.pitch_deg};
  for (float x : v)
    if (!__builtin_isfinite(x)) return false;
  return c.fov_deg > 0.0f && c.fov_deg < 180.0f && c.aspect > 0.0f && c.near_clip > 0.0f && c.far_clip > c.near_clip;
}

// The camera the reference builds, then setRotation(yaw, pitch).
WT_HD Camera camera_from(const wt_camera& c) {
  Camera cam(c.fov_deg, c.aspect, c.near_clip, c.far_clip, Vec3{c.position[0], c.position[1], c.position[2]});
  cam.set_rotation(c.yaw_deg, c.pitch_deg);
  return cam;
}

// Line through p0, p1 intersected with the plane y = d (unit normal (0,1,0)):
// render-engine.h:111-124. The parametrisation along the unit direction is the
// reference's; a line parallel to the plane has no single intersection.
WT_HD bool hit_plane_y(const Vec3& p0, const Vec3& p1, double d, Vec3* out) {
  const Vec3 dir = unit(sub(p1, p0));
  const double denom = dir.y;
  if (denom == 0.0) return false;
  const double t = (d - p0.y) / denom;
  *out = add(p0, scale(dir, t));
  return true;
}

// Running x / y bounds of the slab points in projector NDC. The reference only
// raises the maximum when the point did not lower the minimum (render-engine.cpp:
// 980-990); the first point sets all four.
struct NdcBounds {
  double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
  int count = 0;

  WT_HD void add_point(const Mat4& pvp, const Vec3& p) {
    const Vec4 q = mat_apply(pvp, Vec4{p.x, 0.0, p.z, 1.0});
    const double nx = q.x / q.w, ny = q.y / q.w;
    if (count++ == 0) {
      xmin = xmax = nx;
      ymin = ymax = ny;
      return;
    }
    if (nx < xmin) xmin = nx;
    else if (nx > xmax) xmax = nx;
    if (ny < ymin) ymin = ny;
    else if (ny > ymax) ymax = ny;
  }
};

// Grid corners of the projected grid, render-engine.cpp:797-1042. corners[i] =
// (x, y, z, w) in the order BL, TL, BR, TR. Returns WT_OK, WT_ERR_NO_INTERSECTION
// (the frustum misses the slab |y| <= amp; the reference skips the water, :894) or
// WT_ERR_DEGENERATE (a singular matrix or a projector ray parallel to y = 0; the
// reference asserts, :1034).
//
// The projector camera depends on the camera alone, so it is built first and every
// frustum-slab point goes straight into the NDC bounds, in the reference's point
// order (edges, then in-slab frustum corners); no point list is kept.
WT_HD int projector_corners(const Camera& cam, double amp, double corners[4][4]) {
  // frustum corner points in NDC, padded by 1.2 in x and y (render-engine.cpp:815-832)
  const double pad = 1.2;
  Mat4 inv_vp;
  if (!mat_inverse(mat_mul(cam.proj, cam.view), &inv_vp)) return WT_ERR_DEGENERATE;
  Vec3 fc[8];
  for (int i = 0; i < 8; ++i) {
    // NDC corner i = (+-1, +-1, +-1) with bit 2 -> x, bit 1 -> y, bit 0 -> z
    const double nx = (i & 4) ? 1.0 : -1.0, ny = (i & 2) ? 1.0 : -1.0, nz = (i & 1) ? 1.0 : -1.0;
    const Vec4 q = mat_apply(inv_vp, Vec4{nx * pad, ny * pad, nz, 1.0});
    fc[i] = {q.x / q.w, q.y / q.w, q.z / q.w};
  }

  // projector = camera clone lifted above the slab and re-aimed (render-engine.cpp:905-962)
  Camera pj = cam;
  const double cam_y = cam.position.y;
  const bool underwater = cam_y < 0.0;
  const double min_height = amp + 1.0;
  if (cam_y < min_height) pj.translate(Vec3{0.0, underwater ? min_height - 2.0 * cam_y : min_height - cam_y, 0.0});
  Vec3 fwd = cam.forward;
  const double eps = 0.001;
  if (fabs(fwd.y) < eps) fwd.y = (fwd.y >= 0.0 ? 1.0 : -1.0) * eps;
  const bool looking_down = fwd.y < 0.0;
  // aim point 1: the forward ray (mirrored in y when it cannot reach y = 0) on y = 0
  const Vec3 ray = (looking_down != underwater) ? fwd : Vec3{fwd.x, -fwd.y, fwd.z};
  Vec3 aim1;
  const bool aim_ok = hit_plane_y(cam.position, add(cam.position, ray), 0.0, &aim1);
  if (aim_ok) {
    // aim point 2: one unit along the true forward, dropped onto y = 0
    Vec3 aim2 = add(cam.position, cam.forward);
    aim2.y = 0.0;
    const double a = fabs(fwd.y);
    const Vec3 aim = add(scale(aim2, 1.0 - a), scale(aim1, a));
    const Vec3 nf = unit(sub(aim, pj.position));
    const Vec3 nf_xz = unit(Vec3{nf.x, 0.0, nf.z});
    double c = dot3(nf, nf_xz);
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    pj.set_rotation(pj.yaw_deg, -acos(c) * 180.0 / kPi);
  }
  const Mat4 pvp = mat_mul(pj.proj, pj.view);

  // points of the frustum inside the slab |y| <= amp (render-engine.cpp:860-891), dropped
  // onto y = 0 and into projector NDC (:966-992)
  const int edges[12][2] = {{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3},
                            {2, 6}, {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7}};
  NdcBounds b;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const Vec3 &p = fc[edges[e][0]], &q = fc[edges[e][1]];
    const double lo = fmin(p.y, q.y), hi = fmax(p.y, q.y);
    Vec3 ip;
    if (lo <= amp && amp <= hi && hit_plane_y(p, q, amp, &ip)) b.add_point(pvp, ip);
    if (lo <= -amp && -amp <= hi && hit_plane_y(p, q, -amp, &ip)) b.add_point(pvp, ip);
  }
  for (int i = 0; i < 8; ++i)
    if (-amp <= fc[i].y && fc[i].y <= amp) b.add_point(pvp, fc[i]);
  if (b.count == 0) return WT_ERR_NO_INTERSECTION;  // the reference skips the water (:894)
  if (!aim_ok) return WT_ERR_DEGENERATE;

  Mat4 range = mat_identity();  // render-engine.cpp:1002-1005
  range.c[0][0] = b.xmax - b.xmin;
  range.c[1][1] = b.ymax - b.ymin;
  range.c[3][0] = b.xmin;
  range.c[3][1] = b.ymin;
  Mat4 inv_pv;
  if (!mat_inverse(mat_mul(cam.proj, pj.view), &inv_pv)) return WT_ERR_DEGENERATE;
  const Mat4 projector_mat = mat_mul(inv_pv, range);  // :1008
  for (int i = 0; i < 4; ++i) {
    // range-space corners BL (0,0), TL (0,1), BR (1,0), TR (1,1) (:1012-1015)
    const double u = (i & 2) ? 1.0 : 0.0, v = (i & 1) ? 1.0 : 0.0;
    const Vec4 n = mat_apply(projector_mat, Vec4{u, v, -1.0, 1.0});
    const Vec4 f = mat_apply(projector_mat, Vec4{u, v, 1.0, 1.0});
    const Vec3 p0{n.x / n.w, n.y / n.w, n.z / n.w}, p1{f.x / f.w, f.y / f.w, f.z / f.w};
    Vec3 ip;
    if (!hit_plane_y(p0, p1, 0.0, &ip)) return WT_ERR_DEGENERATE;  // the reference asserts (:1034)
    corners[i][0] = ip.x;
    corners[i][1] = ip.y;
    corners[i][2] = ip.z;
    corners[i][3] = 1.0;
  }
  return WT_OK;
}

}  // namespace proj
}  // namespace wt

// back to the HIP default (fused multiply-add allowed) for code after this header
#pragma clang fp contract(fast)
