// Host-side geometry of the water stage: camera, projector, steepness.
// Double precision throughout; no device code.
#pragma once

#include <cstdint>

namespace wt {
namespace host {

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

Mat4 mat_identity();
Mat4 mat_mul(const Mat4& a, const Mat4& b);
Vec4 mat_apply(const Mat4& m, const Vec4& v);
// Gauss-Jordan inverse with partial pivoting; returns false for a singular matrix.
bool mat_inverse(const Mat4& m, Mat4* out);
Mat4 perspective_rh(double fovy_rad, double aspect, double znear, double zfar);
Mat4 look_at_rh(const Vec3& eye, const Vec3& center, const Vec3& up);

// The reference Camera (src/camera.h / src/camera.cpp) reduced to what the water
// stage reads: position, yaw/pitch basis, projection and view matrices.
struct Camera {
  double fov_deg, aspect, znear, zfar;
  Vec3 position, forward, up, up_init, right;
  double yaw_deg, pitch_deg;
  Mat4 proj, view;

  Camera(double fov, double aspect_, double n, double f, const Vec3& pos);
  void set_rotation(double yaw, double pitch);  // camera.cpp:93-98
  void translate(const Vec3& d);                // camera.cpp:100-104
  void update_basis();                          // camera.cpp:130-138
};

enum ProjectorStatus { PROJ_OK = 0, PROJ_NO_INTERSECTION = 2, PROJ_DEGENERATE = 6 };

// Grid corners of the projected grid, render-engine.cpp:797-1042. corners[i] =
// (x, y, z, w) in the order BL, TL, BR, TR.
int projector_corners(const Camera& cam, double displaceable_amplitude, double corners[4][4]);

}  // namespace host
}  // namespace wt
