// wave-tool water-displacement ORACLE — test infrastructure only.
//
// A scalar CPU restatement of the reference's water-displacement path, templated on
// the scalar type: `double` is the parity oracle, `float` is the multithreaded CPU
// baseline that is timed beside the GPU, `long double` is the brute-force check of the
// double instantiation. Nothing in the shipped library links, includes or calls this
// file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker or the timed CPU baseline.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference repository root). The reference has no tests or golden vectors; the
// parity pins for each function are the closed forms and invariants exercised in
// tests/test_oracle_*.py (listed next to each function below).
//
// The GLM arithmetic the reference calls (perspective, lookAt, inverse, mix,
// normalize, cross, reflect, mod, clamp) is written here from the textbook
// definitions (gluPerspective / gluLookAt matrices, cofactor inverse, GLSL spec).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace wto {

// ---------------------------------------------------------------- vector math
template <class T> struct V2 { T x, y; };
template <class T> struct V3 {
  T x, y, z;
  V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  V3 operator*(T s) const { return {x * s, y * s, z * s}; }
  V3 operator-() const { return {-x, -y, -z}; }
};
template <class T> struct V4 {
  T x, y, z, w;
  V4 operator+(const V4& o) const { return {x + o.x, y + o.y, z + o.z, w + o.w}; }
  V4 operator-(const V4& o) const { return {x - o.x, y - o.y, z - o.z, w - o.w}; }
  V4 operator*(T s) const { return {x * s, y * s, z * s, w * s}; }
  V4 operator/(T s) const { return {x / s, y / s, z / s, w / s}; }
  V3<T> xyz() const { return {x, y, z}; }
};
// Column-major 4x4 (m[col][row]) — the GLM storage the reference's matrices use.
template <class T> struct M4 {
  T m[4][4];
  static M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = T(1);
    return r;
  }
  V4<T> operator*(const V4<T>& v) const {
    T in[4] = {v.x, v.y, v.z, v.w}, out[4];
    for (int r = 0; r < 4; ++r) {
      out[r] = T(0);
      for (int c = 0; c < 4; ++c) out[r] += m[c][r] * in[c];
    }
    return {out[0], out[1], out[2], out[3]};
  }
  M4 operator*(const M4& b) const {
    M4 r{};
    for (int c = 0; c < 4; ++c)
      for (int rr = 0; rr < 4; ++rr) {
        T s = T(0);
        for (int k = 0; k < 4; ++k) s += m[k][rr] * b.m[c][k];
        r.m[c][rr] = s;
      }
    return r;
  }
};

template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline V3<T> normalize(const V3<T>& a) {
  T inv = T(1) / std::sqrt(dot(a, a));
  return a * inv;
}
template <class T> inline T length(const V3<T>& a) { return std::sqrt(dot(a, a)); }
// GLSL mix(x, y, a) = x*(1-a) + y*a.
template <class T> inline V4<T> mix(const V4<T>& a, const V4<T>& b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline V3<T> mix(const V3<T>& a, const V3<T>& b, T t) { return a * (T(1) - t) + b * t; }
// GLSL/GLM mod(x, y) = x - y*floor(x/y).
template <class T> inline T mod(T x, T y) { return x - y * std::floor(x / y); }
template <class T> inline T clamp(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }
template <class T> inline T radians(T d) { return d * T(3.14159265358979323846264338327950288L) / T(180); }
template <class T> inline T degrees(T r) { return r * T(180) / T(3.14159265358979323846264338327950288L); }
// reflect(I, N) = I - 2 dot(N, I) N.
template <class T> inline V3<T> reflect(const V3<T>& i, const V3<T>& n) { return i - n * (T(2) * dot(n, i)); }

// gluPerspective, right-handed, clip z in [-1, 1] (the reference never defines
// GLM_FORCE_DEPTH_ZERO_TO_ONE). Used at src/camera.cpp:141.
template <class T> M4<T> perspective(T fovy, T aspect, T znear, T zfar) {
  M4<T> r{};
  T f = T(1) / std::tan(fovy / T(2));
  r.m[0][0] = f / aspect;
  r.m[1][1] = f;
  r.m[2][2] = (zfar + znear) / (znear - zfar);
  r.m[2][3] = T(-1);
  r.m[3][2] = (T(2) * zfar * znear) / (znear - zfar);
  return r;
}
// gluLookAt, right-handed. Used at src/camera.cpp:145.
template <class T> M4<T> lookAt(const V3<T>& eye, const V3<T>& center, const V3<T>& up) {
  V3<T> f = normalize(center - eye);
  V3<T> s = normalize(cross(f, up));
  V3<T> u = cross(s, f);
  M4<T> r = M4<T>::identity();
  r.m[0][0] = s.x; r.m[1][0] = s.y; r.m[2][0] = s.z;
  r.m[0][1] = u.x; r.m[1][1] = u.y; r.m[2][1] = u.z;
  r.m[0][2] = -f.x; r.m[1][2] = -f.y; r.m[2][2] = -f.z;
  r.m[3][0] = -dot(s, eye);
  r.m[3][1] = -dot(u, eye);
  r.m[3][2] = dot(f, eye);
  return r;
}
// General 4x4 inverse by cofactor expansion (adjugate / determinant). Used at
// src/render-engine.cpp:286 and :1008.
template <class T> M4<T> inverse(const M4<T>& A) {
  // flatten to row-major a[r][c]
  T a[4][4];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) a[r][c] = A.m[c][r];
  auto minor3 = [&](int skipr, int skipc) {
    T b[3][3];
    for (int r = 0, rr = 0; r < 4; ++r) {
      if (r == skipr) continue;
      for (int c = 0, cc = 0; c < 4; ++c) {
        if (c == skipc) continue;
        b[rr][cc++] = a[r][c];
      }
      ++rr;
    }
    return b[0][0] * (b[1][1] * b[2][2] - b[1][2] * b[2][1]) - b[0][1] * (b[1][0] * b[2][2] - b[1][2] * b[2][0]) +
           b[0][2] * (b[1][0] * b[2][1] - b[1][1] * b[2][0]);
  };
  T cof[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) cof[r][c] = (((r + c) & 1) ? T(-1) : T(1)) * minor3(r, c);
  T det = T(0);
  for (int c = 0; c < 4; ++c) det += a[0][c] * cof[0][c];
  M4<T> inv{};
  // inverse[r][c] = cof[c][r] / det  (adjugate = transpose of cofactor matrix)
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv.m[c][r] = cof[c][r] / det;
  return inv;
}

// ---------------------------------------------------------------- camera
// Restates src/camera.cpp:40-146 (defaults :40-45, ctor :47-54, setRotation :93-98,
// translate :100-104, updateBasisVectors :130-138, projection :140-142, view :144-146).
template <class T> struct Camera {
  T fov, aspect, znear, zfar;  // fov in degrees (src/camera.cpp:141 converts)
  V3<T> position, forward, up, upInit, right;
  T yaw, pitch;                // degrees
  M4<T> proj, view;

  Camera(T fov_, T aspect_, T n, T f, V3<T> pos, T yaw_ = T(-90), T pitch_ = T(0))
      : fov(fov_), aspect(aspect_), znear(n), zfar(f), position(pos), forward{T(0), T(0), T(-1)}, up{T(0), T(1), T(0)},
        yaw(yaw_), pitch(pitch_) {
    upInit = up;
    right = cross(forward, up);
    updateBasisVectors();
    updateProjectionMat();
  }
  void setRotation(T yawDeg, T pitchDeg) {
    yaw = mod(yawDeg, T(360));
    pitch = clamp(pitchDeg, T(-89), T(89));
    updateBasisVectors();
  }
  void translate(const V3<T>& d) {
    position = position + d;
    updateViewMat();
  }
  void updateBasisVectors() {
    T cy = std::cos(radians(yaw)), sy = std::sin(radians(yaw));
    T cp = std::cos(radians(pitch)), sp = std::sin(radians(pitch));
    forward = normalize(V3<T>{cy * cp, sp, sy * cp});
    right = normalize(cross(forward, upInit));
    up = normalize(cross(right, forward));
    updateViewMat();
  }
  void updateProjectionMat() { proj = perspective(radians(fov), aspect, znear, zfar); }
  void updateViewMat() { view = lookAt(position, position + forward, up); }
};

// ---------------------------------------------------------------- projector
// geometry::Line / Plane / utils::linePlaneIntersection, src/render-engine.h:55-124.
// Plane: a x + b y + c z = d with unit normal; the segment is an infinite line.
template <class T> bool linePlaneIntersection(V3<T>& out, const V3<T>& p0, const V3<T>& p1, const V3<T>& n, T d) {
  V3<T> dir = normalize(p1 - p0);
  T denom = dot(n, dir);
  if (denom == T(0)) return false;  // parallel: 0 or infinitely many points (:118)
  T t = (d - dot(n, p0)) / denom;
  out = p0 + dir * t;
  return true;
}

enum { WTO_OK = 0, WTO_NO_INTERSECTION = 2, WTO_DEGENERATE = 6 };

template <class T> struct ProjectorDebug {
  std::vector<V4<T>> slabPoints;   // frustum ∩ slab points, world (before y=0 projection)
  std::vector<V4<T>> ndcPoints;    // the same points on y=0 in projector NDC
  T xmin, xmax, ymin, ymax;        // range bounds (render-engine.cpp:966-992)
  V3<T> projectorPosition;
  T projectorPitch;
  V3<T> aimpoint;
  bool raised;
};

// Deliberate departures from the reference's projector, used only by tests that show
// a pin fails when the restatement is wrong: the frustum padding scalar (1.2 at
// src/render-engine.cpp:828) and the operand order of the aim-point mix (:953).
template <class T> struct ProjectorMutation {
  T pad = T(1.2);
  bool swapAimMix = false;
};

// Restates the water block of RenderEngine::render, src/render-engine.cpp:797-1042:
// displaceable slab, padded frustum corners, edge/slab intersections, projector
// placement and aim, range matrix, four corner rays intersected with y = 0.
// Returns the corners in the order of waterGridCornerPoints (:1012-1015):
// [0] bottom-left, [1] top-left, [2] bottom-right, [3] top-right.
template <class T>
int projectorCorners(const Camera<T>& cam, T displaceableAmplitude, V4<T> corners[4], ProjectorDebug<T>* dbg = nullptr,
                     const ProjectorMutation<T>* mutation = nullptr) {
  const T amp = displaceableAmplitude;  // :803
  const V3<T> up{T(0), T(1), T(0)};
  const M4<T> viewProjection = cam.proj * cam.view;  // :284
  const M4<T> invVP = inverse(viewProjection);       // :286
  V4<T> fc[8] = {{-1, -1, -1, 1}, {-1, -1, 1, 1}, {-1, 1, -1, 1}, {-1, 1, 1, 1},
                 {1, -1, -1, 1},  {1, -1, 1, 1},  {1, 1, -1, 1},  {1, 1, 1, 1}};  // :815-822
  const T PAD = mutation ? mutation->pad : T(1.2);                                // :828
  for (auto& p : fc) { p.x *= PAD; p.y *= PAD; }                                   // :829-832
  for (auto& p : fc) { V4<T> q = invVP * p; p = q / q.w; }                         // :835-838
  static const int E[24] = {0, 1, 0, 2, 0, 4, 1, 3, 1, 5, 2, 3, 2, 6, 3, 7, 4, 5, 4, 6, 5, 7, 6, 7};  // :842-853
  std::vector<V4<T>> pts;
  for (int i = 0; i < 12; ++i) {  // :860-884
    V3<T> p0 = fc[E[2 * i]].xyz(), p1 = fc[E[2 * i + 1]].xyz();
    T lo = std::fmin(p0.y, p1.y), hi = std::fmax(p0.y, p1.y);
    V3<T> ip;
    if (lo <= amp && amp <= hi && linePlaneIntersection(ip, p0, p1, up, amp)) pts.push_back({ip.x, ip.y, ip.z, T(1)});
    if (lo <= -amp && -amp <= hi && linePlaneIntersection(ip, p0, p1, up, -amp)) pts.push_back({ip.x, ip.y, ip.z, T(1)});
  }
  for (auto& p : fc)  // :887-891
    if (-amp <= p.y && p.y <= amp) pts.push_back(p);
  if (dbg) dbg->slabPoints = pts;
  if (pts.empty()) return WTO_NO_INTERSECTION;  // :894 silently skips the water

  Camera<T> projector = cam;  // :905
  const T camY = cam.position.y;
  const bool underwater = camY < T(0);        // :908
  const T ELEV = T(1);                        // :910
  const T MINDIST = amp + ELEV;               // :911
  bool raised = false;
  if (camY < MINDIST) {                       // :914-917
    raised = true;
    if (underwater) projector.translate(V3<T>{T(0), MINDIST - T(2) * camY, T(0)});
    else projector.translate(V3<T>{T(0), MINDIST - camY, T(0)});
  }
  V3<T> fwdSafe = cam.forward;                // :920-926
  const T EPS = T(0.001);
  if (std::fabs(fwdSafe.y) < EPS) fwdSafe.y = (fwdSafe.y >= T(0) ? T(1) : T(-1)) * EPS;
  V3<T> aim1;                                 // :929-939
  const bool lookingDown = fwdSafe.y < T(0);
  if (lookingDown != underwater) {
    if (!linePlaneIntersection(aim1, cam.position, cam.position + fwdSafe, up, T(0))) return WTO_DEGENERATE;
  } else {
    V3<T> mirrored = reflect(fwdSafe, up);
    if (!linePlaneIntersection(aim1, cam.position, cam.position + mirrored, up, T(0))) return WTO_DEGENERATE;
  }
  const T FWDLEN = T(1);                      // :943
  V3<T> aim2 = cam.position + cam.forward * FWDLEN;  // :944
  aim2.y = T(0);                              // :946
  const T a = std::fabs(fwdSafe.y);           // :950
  const V3<T> aim = (mutation && mutation->swapAimMix) ? mix(aim1, aim2, a) : mix(aim2, aim1, a);  // :953
  const V3<T> nf = normalize(aim - projector.position);              // :956
  const V3<T> nfxz = normalize(V3<T>{nf.x, T(0), nf.z});             // :957
  T pitchDeg = -degrees(std::acos(clamp(dot(nf, nfxz), T(-1), T(1))));  // :959
  projector.setRotation(projector.yaw, pitchDeg);                    // :962
  for (auto& p : pts) p.y = T(0);                                    // :966-969
  const M4<T> pvp = projector.proj * projector.view;                 // :974
  std::vector<V4<T>> ndc;
  for (auto& p : pts) { V4<T> q = pvp * p; ndc.push_back(q / q.w); }  // :975-978
  T xmin = ndc[0].x, xmax = ndc[0].x, ymin = ndc[0].y, ymax = ndc[0].y;  // :981-992
  for (size_t i = 1; i < ndc.size(); ++i) {
    // the reference's else-if bound update (:986, :989) is kept: a point can only
    // update the max when it did not update the min
    if (ndc[i].x < xmin) xmin = ndc[i].x;
    else if (ndc[i].x > xmax) xmax = ndc[i].x;
    if (ndc[i].y < ymin) ymin = ndc[i].y;
    else if (ndc[i].y > ymax) ymax = ndc[i].y;
  }
  M4<T> range = M4<T>::identity();            // :1002-1005 (column-major)
  range.m[0][0] = xmax - xmin;
  range.m[1][1] = ymax - ymin;
  range.m[3][0] = xmin;
  range.m[3][1] = ymin;
  const M4<T> projectorMat = inverse(cam.proj * projector.view) * range;  // :1008
  V4<T> gc[4] = {{0, 0, -1, 1}, {0, 1, -1, 1}, {1, 0, -1, 1}, {1, 1, -1, 1}};  // :1012-1015
  for (int i = 0; i < 4; ++i) {               // :1019-1036
    V4<T> p0 = gc[i], p1 = gc[i];
    p1.z = T(1);
    p0 = projectorMat * p0; p0 = p0 / p0.w;
    p1 = projectorMat * p1; p1 = p1 / p1.w;
    V3<T> ip;
    if (!linePlaneIntersection(ip, p0.xyz(), p1.xyz(), up, T(0))) return WTO_DEGENERATE;  // :1034 asserts
    corners[i] = {ip.x, ip.y, ip.z, T(1)};
  }
  if (dbg) {
    dbg->ndcPoints = ndc;
    dbg->xmin = xmin; dbg->xmax = xmax; dbg->ymin = ymin; dbg->ymax = ymax;
    dbg->projectorPosition = projector.position;
    dbg->projectorPitch = projector.pitch;
    dbg->aimpoint = aim;
    dbg->raised = raised;
  }
  return WTO_OK;
}

// ---------------------------------------------------------------- heightmap
enum { WTO_TEX_R8_UNORM = 0, WTO_TEX_R32F = 1 };
enum { WTO_WRAP_REPEAT = 0, WTO_WRAP_MIRRORED_REPEAT = 1 };

struct Texture {
  const void* texels = nullptr;  // row-major, row 0 = texture t in [0, 1/H) (GL memory order)
  int width = 0, height = 0, format = WTO_TEX_R32F, wrap = WTO_WRAP_REPEAT;
};

// GL texel-index wrap (OpenGL 4.x spec, "Coordinate Wrapping and Texel Selection"):
// REPEAT: i mod size; MIRRORED_REPEAT: (size-1) - mirror((i mod 2size) - size) with
// mirror(a) = a >= 0 ? a : -(1+a). The reference sets MIRRORED_REPEAT at
// src/texture.cpp:63-64 (GL_REPEAT is the commented alternative at :61-62).
inline long long wrapIndex(long long i, long long size, int wrap) {
  if (wrap == WTO_WRAP_REPEAT) {
    long long r = i % size;
    return r < 0 ? r + size : r;
  }
  long long two = 2 * size;
  long long m = i % two;
  if (m < 0) m += two;
  long long a = m - size;
  long long mir = a >= 0 ? a : -(1 + a);
  return (size - 1) - mir;
}

template <class T> inline T texel(const Texture& tx, long long i, long long j) {
  long long ii = wrapIndex(i, tx.width, tx.wrap), jj = wrapIndex(j, tx.height, tx.wrap);
  size_t idx = (size_t)jj * (size_t)tx.width + (size_t)ii;
  if (tx.format == WTO_TEX_R8_UNORM) return T(((const uint8_t*)tx.texels)[idx]) / T(255);  // unorm8 -> c/255
  return T(((const float*)tx.texels)[idx]);
}

// Bilinear GL_LINEAR filtering at texture coordinates (s, t) with texel centres at
// (i + 1/2)/size (GL spec; src/texture.cpp:65-66 sets GL_LINEAR min/mag, mipmaps are
// generated but never selected because the shader samples the base level through
// texture() in a vertex shader, where the implicit LOD is 0). Returns the value and
// its derivatives with respect to s and t (piecewise, from the same texel cell).
template <class T> inline T sampleBilinear(const Texture& tx, T s, T t, T* dds, T* ddt) {
  T u = s * T(tx.width) - T(0.5), v = t * T(tx.height) - T(0.5);
  T fu = std::floor(u), fv = std::floor(v);
  long long i0 = (long long)fu, j0 = (long long)fv;
  T a = u - fu, b = v - fv;
  T t00 = texel<T>(tx, i0, j0), t10 = texel<T>(tx, i0 + 1, j0);
  T t01 = texel<T>(tx, i0, j0 + 1), t11 = texel<T>(tx, i0 + 1, j0 + 1);
  T top = t00 + (t10 - t00) * a, bot = t01 + (t11 - t01) * a;
  if (dds) *dds = T(tx.width) * ((T(1) - b) * (t10 - t00) + b * (t11 - t01));
  if (ddt) *ddt = T(tx.height) * ((T(1) - a) * (t01 - t00) + a * (t11 - t10));
  return top + (bot - top) * b;
}

// ---------------------------------------------------------------- vertex program
// GLSL struct GerstnerWave (assets/shaders/water-grid.vert:45-51), member order kept,
// with the appended phase offset phi0 (zero in the reference).
template <class T> struct Wave { T A, w, phi, Q, Dx, Dz, phi0; };

template <class T> struct Scene {
  V4<T> corners[4];   // [0] BL, [1] TL, [2] BR, [3] TR (render-engine.cpp:1039-1042)
  V3<T> cameraPosition;
  int N;              // gridLength (render-engine.cpp:1055)
  T sampleScale;      // heightmapSampleScale (render-engine.h:225)
  T displacementScale;// heightmapDisplacementScale (render-engine.h:224)
  T bounce;           // verticalBounceWaveDisplacement (render-engine.cpp:299-300)
  T scrollX, scrollZ; // heightmap scroll velocity in world m/s (zero in the reference)
  T time;             // waveAnimationTimeInSeconds
  const Wave<T>* waves;
  int W;
  const Texture* tex; // nullptr: no heightmap term
};

// computeInterpolatedGridPosition, water-grid.vert:87-91.
template <class T> inline V4<T> gridPosition(const Scene<T>& sc, T u, T v) {
  V4<T> a = mix(sc.corners[0], sc.corners[2], u);  // BL -> BR
  V4<T> b = mix(sc.corners[1], sc.corners[3], u);  // TL -> TR
  return mix(a, b, v);
}

// computeGerstnerSurfacePosition, water-grid.vert:74-85, plus the analytic partial
// derivatives of the displaced position with respect to the undisplaced grid point
// (x, z) (derived from the same sum; no reference counterpart).
template <class T> struct GerstnerOut {
  V3<T> P;
  V3<T> dPdx, dPdz;
};
template <class T> inline GerstnerOut<T> gerstner(const Scene<T>& sc, T x, T z) {
  GerstnerOut<T> o;
  o.P = {x, T(0), z};
  o.dPdx = {T(1), T(0), T(0)};
  o.dPdz = {T(0), T(0), T(1)};
  for (int i = 0; i < sc.W; ++i) {
    const Wave<T>& wv = sc.waves[i];
    // :78  theta = w * dot(D, xz) + phi * t   (+ phi0, the appended phase offset)
    T th = wv.w * (wv.Dx * x + wv.Dz * z) + wv.phi * sc.time + wv.phi0;
    T c = std::cos(th), s = std::sin(th);
    T qc = wv.Q * c;                                                  // :79
    o.P = o.P + V3<T>{wv.Dx * qc, s, wv.Dz * qc} * wv.A;              // :80-81
    // d/dx: d(theta)/dx = w Dx, d(theta)/dz = w Dz
    T qaws = wv.Q * wv.A * wv.w * s, awc = wv.A * wv.w * c;
    o.dPdx = o.dPdx + V3<T>{-qaws * wv.Dx * wv.Dx, awc * wv.Dx, -qaws * wv.Dz * wv.Dx};
    o.dPdz = o.dPdz + V3<T>{-qaws * wv.Dx * wv.Dz, awc * wv.Dz, -qaws * wv.Dz * wv.Dz};
  }
  return o;
}

// sampleHeightmap + computeHeightmapDisplacement, water-grid.vert:93-111:
// r = texture(hm, s_h * (x, -z)).r ; dy = s_d * 2 * (r - 0.5). The scroll term
// subtracts v*t from the world sample point (zero in the reference).
template <class T> inline T heightmapDisplacement(const Scene<T>& sc, T Px, T Pz, T* ddx, T* ddz) {
  if (!sc.tex || !sc.tex->texels) {
    if (ddx) *ddx = T(0);
    if (ddz) *ddz = T(0);
    return T(0);
  }
  T s = sc.sampleScale * (Px - sc.scrollX * sc.time);
  T t = sc.sampleScale * -(Pz - sc.scrollZ * sc.time);
  T dHds, dHdt;
  T r = sampleBilinear<T>(*sc.tex, s, t, &dHds, &dHdt);
  T k = sc.displacementScale * T(2);
  if (ddx) *ddx = k * dHds * sc.sampleScale;     // d/dPx
  if (ddz) *ddz = -k * dHdt * sc.sampleScale;    // d/dPz
  return sc.displacementScale * (T(2) * (r - T(0.5)));
}

// Full displaced position of one grid point (x, z): Gerstner, heightmap bump, bounce
// (water-grid.vert:138, 141-147).
template <class T> inline V3<T> displacedPosition(const Scene<T>& sc, T x, T z) {
  GerstnerOut<T> g = gerstner(sc, x, z);
  V3<T> P = g.P;
  P.y += heightmapDisplacement(sc, P.x, P.z, (T*)nullptr, (T*)nullptr);
  P.y += sc.bounce;
  return P;
}

template <class T> struct VertexOut { V3<T> pos, nrm, tan, bit; };

// Vertex id -> (u, v), water-grid.vert:123-125: du = dv = 1/(N-1),
// u = (id mod N) * du, v = (id / N) * dv (integer division).
template <class T> inline void vertexUV(int N, long long id, T* u, T* v) {
  T d = T(1) / T(N - 1);
  *u = T(id % N) * d;
  *v = T(id / N) * d;
}

// Analytic mode (the build's output contract): position as in water-grid.vert:135-147,
// normal = normalize(dP/dz x dP/dx) including the heightmap term through the chain
// rule, flipped toward the camera as in :190; tangent = normalize(dP/dx),
// bitangent = normalize(dP/dz) (not flipped).
template <class T> inline VertexOut<T> vertexAnalytic(const Scene<T>& sc, long long id) {
  T u, v;
  vertexUV<T>(sc.N, id, &u, &v);
  V4<T> gp = gridPosition(sc, u, v);
  GerstnerOut<T> g = gerstner(sc, gp.x, gp.z);
  V3<T> P = g.P;
  T hx, hz;
  T h = heightmapDisplacement(sc, P.x, P.z, &hx, &hz);
  P.y += h;
  P.y += sc.bounce;
  // chain rule: dh/dx = dh/dPx * dPx/dx + dh/dPz * dPz/dx
  V3<T> Tx = g.dPdx, Tz = g.dPdz;
  Tx.y += hx * g.dPdx.x + hz * g.dPdx.z;
  Tz.y += hx * g.dPdz.x + hz * g.dPdz.z;
  VertexOut<T> o;
  o.pos = P;
  V3<T> n = normalize(cross(Tz, Tx));
  V3<T> view = sc.cameraPosition - P;                         // :158
  if (dot(n, normalize(view)) < T(0)) n = -n;                 // :190
  o.nrm = n;
  o.tan = normalize(Tx);
  o.bit = normalize(Tz);
  return o;
}

// Reference mode: the four-neighbour finite-difference normal of water-grid.vert:164-190
// (neighbour uvs extrapolate past the grid edge, :164-167).
template <class T> inline VertexOut<T> vertexReference(const Scene<T>& sc, long long id) {
  T u, v;
  vertexUV<T>(sc.N, id, &u, &v);
  T d = T(1) / T(sc.N - 1);
  auto at = [&](T uu, T vv) {
    V4<T> gp = gridPosition(sc, uu, vv);
    return displacedPosition(sc, gp.x, gp.z);
  };
  VertexOut<T> o;
  o.pos = at(u, v);
  V3<T> mu = at(u - d, v), pu = at(u + d, v), mv = at(u, v - d), pv = at(u, v + d);
  V3<T> n = normalize(cross(pu - mu, pv - mv));              // :187
  V3<T> view = sc.cameraPosition - o.pos;
  if (dot(n, normalize(view)) < T(0)) n = -n;                 // :190
  o.nrm = n;
  o.tan = normalize(pu - mu);
  o.bit = normalize(pv - mv);
  return o;
}

// ---------------------------------------------------------------- adjoint (extension)
// Heightfield fit: world grid (x_j, z_j), snapshots t_s = t0 + s dt. Model per wave
// (A, L, alpha, Q, phi0): w = 2 pi / L, omega = sqrt(g w), D = (cos alpha, sin alpha),
// theta = w D.x + omega t + phi0 (the reference's sign convention, water-grid.vert:78),
// d = (Q A Dx cos, A sin, Q A Dz cos). Loss = (1 / (2 S V)) sum |d - target|^2.
// No reference counterpart; pinned by central finite differences of this loss.
template <class T> struct FitProblem {
  int n;            // grid side
  T spacing;        // metres between grid points
  T x0, z0;         // world position of grid point (0, 0)
  int S;            // snapshots
  T t0, dt;
  T g;              // gravity
  const float* target;  // [S][n*n][3] displacement (dx, h, dz)
};
template <class T> inline void gridPoint(const FitProblem<T>& p, long long v, T* x, T* z) {
  *x = p.x0 + T(v % p.n) * p.spacing;
  *z = p.z0 + T(v / p.n) * p.spacing;
}
// Unnormalised sums over the vertex range [lo, hi): returns 1/2 sum |r|^2 and adds
// sum r . dd/dp into grad (W x 5, may be null). params: W x 5 {A, L, alpha, Q, phi0}.
// The caller divides by S*V to get the loss and gradient of the mean misfit.
template <class T>
T fitLossGradRange(const FitProblem<T>& p, const T* params, int W, T* grad, long long lo, long long hi) {
  const long long V = (long long)p.n * p.n;
  const T PI = T(3.14159265358979323846264338327950288L);
  std::vector<T> w(W), om(W), dx(W), dz(W), th(W);
  for (int i = 0; i < W; ++i) {
    T L = params[5 * i + 1];
    w[i] = T(2) * PI / L;
    om[i] = std::sqrt(p.g * w[i]);
    dx[i] = std::cos(params[5 * i + 2]);
    dz[i] = std::sin(params[5 * i + 2]);
  }
  T loss = T(0);
  for (long long v = lo; v < hi; ++v) {
    T x, z;
    gridPoint(p, v, &x, &z);
    for (int s = 0; s < p.S; ++s) {
      T t = p.t0 + T(s) * p.dt;
      T d[3] = {T(0), T(0), T(0)};
      for (int i = 0; i < W; ++i) {
        const T A = params[5 * i], Q = params[5 * i + 3], ph0 = params[5 * i + 4];
        th[i] = w[i] * (dx[i] * x + dz[i] * z) + om[i] * t + ph0;
        T c = std::cos(th[i]), sn = std::sin(th[i]);
        d[0] += Q * A * dx[i] * c;
        d[1] += A * sn;
        d[2] += Q * A * dz[i] * c;
      }
      const float* tg = p.target + ((size_t)s * V + v) * 3;
      T r[3] = {d[0] - T(tg[0]), d[1] - T(tg[1]), d[2] - T(tg[2])};
      loss += T(0.5) * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      if (!grad) continue;
      for (int i = 0; i < W; ++i) {
        const T A = params[5 * i], L = params[5 * i + 1], Q = params[5 * i + 3];
        T c = std::cos(th[i]), sn = std::sin(th[i]);
        T e = r[0] * dx[i] + r[2] * dz[i];        // residual along D
        T ep = -r[0] * dz[i] + r[2] * dx[i];      // residual along D-perp
        T gth = A * (r[1] * c - Q * e * sn);      // r . dd/dtheta
        T dthdL = -(w[i] / L) * (dx[i] * x + dz[i] * z) - om[i] / (T(2) * L) * t;
        T dthda = w[i] * (-dz[i] * x + dx[i] * z);
        grad[5 * i + 0] += Q * e * c + r[1] * sn;
        grad[5 * i + 1] += gth * dthdL;
        grad[5 * i + 2] += gth * dthda + Q * A * ep * c;
        grad[5 * i + 3] += A * e * c;
        grad[5 * i + 4] += gth;
      }
    }
  }
  return loss;
}

}  // namespace wto
