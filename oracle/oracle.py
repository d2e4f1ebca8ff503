"""ctypes loader for the ORACLE (test infrastructure only).

The oracle is a scalar CPU restatement of the reference's water-displacement path
(wt_oracle.hpp). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import this module, and only as the checker or the timed CPU baseline; the
shipped library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WT_ORACLE_LIB names another build of the oracle (the sanitizer build)
LIB_PATH = os.environ.get("WT_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

MODE_ANALYTIC = 0
MODE_REFERENCE = 1
TEX_R8_UNORM = 0
TEX_R32F = 1
WRAP_REPEAT = 0
WRAP_MIRRORED_REPEAT = 1
NO_INTERSECTION = 2


class Scene(ctypes.Structure):
    _fields_ = [
        ("corners", ctypes.c_double * 16),
        ("camera", ctypes.c_double * 3),
        ("N", ctypes.c_int32),
        ("W", ctypes.c_int32),
        ("sample_scale", ctypes.c_double),
        ("displacement_scale", ctypes.c_double),
        ("bounce", ctypes.c_double),
        ("scroll_x", ctypes.c_double),
        ("scroll_z", ctypes.c_double),
        ("time", ctypes.c_double),
        ("waves", ctypes.c_void_p),
        ("texels", ctypes.c_void_p),
        ("tex_w", ctypes.c_int32),
        ("tex_h", ctypes.c_int32),
        ("tex_format", ctypes.c_int32),
        ("tex_wrap", ctypes.c_int32),
    ]


class Fit(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("S", ctypes.c_int32),
        ("spacing", ctypes.c_double),
        ("x0", ctypes.c_double),
        ("z0", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("g", ctypes.c_double),
        ("target", ctypes.c_void_p),
    ]


class Shade(ctypes.Structure):
    _fields_ = [
        ("fog", ctypes.c_double * 4),
        ("fog_far", ctypes.c_double),
        ("fog_near", ctypes.c_double),
        ("soft_edges", ctypes.c_double),
        ("sun", ctypes.c_double * 3),
        ("shininess", ctypes.c_double),
        ("strength", ctypes.c_double),
        ("tint", ctypes.c_double),
        ("viewport", ctypes.c_double * 2),
        ("clarity", ctypes.c_double),
        ("zfar", ctypes.c_double),
        ("znear", ctypes.c_double),
        ("depth", ctypes.c_void_p),
        ("depth_w", ctypes.c_int32),
        ("depth_h", ctypes.c_int32),
        ("refl", ctypes.c_void_p),
        ("refl_w", ctypes.c_int32),
        ("refl_h", ctypes.c_int32),
        ("refr", ctypes.c_void_p),
        ("refr_w", ctypes.c_int32),
        ("refr_h", ctypes.c_int32),
        ("sky", ctypes.c_void_p),
        ("sky_len", ctypes.c_int32),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        L.wto_abi_version.restype = ctypes.c_int
        L.wto_camera.argtypes = [dp] * 6
        L.wto_inverse.argtypes = [dp, dp]
        L.wto_projector.argtypes = [dp, ctypes.c_double, dp, dp, dp, dp, ctypes.c_int]
        L.wto_projector.restype = ctypes.c_int
        L.wto_projector_variant.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp,
                                            ctypes.c_int]
        L.wto_projector_variant.restype = ctypes.c_int
        L.wto_wrap_index.argtypes = [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int]
        L.wto_wrap_index.restype = ctypes.c_longlong
        L.wto_sample_bilinear.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, dp, dp]
        L.wto_sample_bilinear.restype = ctypes.c_double
        for name, ptr in (("wto_vertices_f64", dp), ("wto_vertices_f32", fp), ("wto_vertices_f80", dp)):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(Scene), ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong,
                          ptr, ptr, ptr, ptr, ctypes.c_int]
            f.restype = ctypes.c_int
        L.wto_gerstner_point.argtypes = [ctypes.POINTER(Scene), ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.wto_displaced_point.argtypes = [ctypes.POINTER(Scene), ctypes.c_double, ctypes.c_double, dp]
        for name in ("wto_fit_loss_grad_f64", "wto_fit_loss_grad_f32"):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(Fit), dp, ctypes.c_int, dp, ctypes.c_int]
            f.restype = ctypes.c_double
        for name, ptr in (("wto_water_shade_f64", dp), ("wto_water_shade_f32", fp)):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(Shade), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                          ctypes.c_longlong, ctypes.c_longlong, ptr, ptr, ctypes.c_int]
            f.restype = ctypes.c_int
        L.wto_cube_face.argtypes = [dp, dp, dp]
        L.wto_cube_face.restype = ctypes.c_int
        L.wto_sample_rgba8.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       dp]
        L.wto_sample_cube.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.wto_sample_depth.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double]
        L.wto_sample_depth.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a, ct=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


def camera(cam):
    """cam = (fov_deg, aspect, near, far, px, py, pz, yaw_deg, pitch_deg)."""
    c = np.ascontiguousarray(cam, dtype=np.float64)
    view, proj = np.zeros(16), np.zeros(16)
    fwd, right, up = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().wto_camera(_p(c), _p(view), _p(proj), _p(fwd), _p(right), _p(up))
    # column-major storage -> numpy matrix M[row, col]
    return dict(view=view.reshape(4, 4).T.copy(), proj=proj.reshape(4, 4).T.copy(), forward=fwd, right=right, up=up)


def inverse(m):
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float64).T.reshape(16))
    out = np.zeros(16)
    lib().wto_inverse(_p(a), _p(out))
    return out.reshape(4, 4).T.copy()


def projector(cam, amp, max_pts=64, pad=None, swap_aim_mix=False):
    """Returns (status, corners[4,4] (BL, TL, BR, TR), info dict). pad / swap_aim_mix
    plant a mistake (another padding scalar, the aim mix's operands swapped) for tests
    that show a pin catches it; the defaults restate the reference."""
    c = np.ascontiguousarray(cam, dtype=np.float64)
    corners = np.zeros(16)
    info = np.zeros(13)
    slab = np.zeros(4 * max_pts)
    ndc = np.zeros(4 * max_pts)
    if pad is None and not swap_aim_mix:
        st = lib().wto_projector(_p(c), float(amp), _p(corners), _p(info), _p(slab), _p(ndc), max_pts)
    else:
        st = lib().wto_projector_variant(_p(c), float(amp), 1.2 if pad is None else float(pad), int(bool(swap_aim_mix)),
                                         _p(corners), _p(info), _p(slab), _p(ndc), max_pts)
    n = int(info[12])
    d = dict(xmin=info[0], xmax=info[1], ymin=info[2], ymax=info[3], projector_pos=info[4:7].copy(),
             projector_pitch=info[7], aim=info[8:11].copy(), raised=bool(info[11]), npoints=n,
             slab_points=slab.reshape(-1, 4)[:n].copy(), ndc_points=ndc.reshape(-1, 4)[:n].copy())
    return st, corners.reshape(4, 4).copy(), d


def wrap_index(i, size, wrap):
    return lib().wto_wrap_index(int(i), int(size), int(wrap))


def sample_bilinear(texels, fmt, wrap, s, t):
    tex = np.ascontiguousarray(texels)
    h, w = tex.shape
    dds, ddt = ctypes.c_double(), ctypes.c_double()
    v = lib().wto_sample_bilinear(tex.ctypes.data, w, h, fmt, wrap, float(s), float(t), ctypes.byref(dds),
                                  ctypes.byref(ddt))
    return v, dds.value, ddt.value


class SceneData:
    """Keeps the numpy arrays a Scene points to alive."""

    def __init__(self, corners, camera_pos, N, waves, time, sample_scale=0.0, displacement_scale=0.0, bounce=0.0,
                 scroll=(0.0, 0.0), texels=None, tex_format=TEX_R32F, tex_wrap=WRAP_REPEAT):
        self.waves = np.ascontiguousarray(np.asarray(waves, dtype=np.float32).reshape(-1, 8))
        self.texels = None if texels is None else np.ascontiguousarray(texels)
        s = Scene()
        s.corners[:] = [float(v) for v in np.asarray(corners, dtype=np.float64).reshape(16)]
        s.camera[:] = [float(v) for v in camera_pos]
        s.N = int(N)
        s.W = int(self.waves.shape[0])
        s.sample_scale = float(sample_scale)
        s.displacement_scale = float(displacement_scale)
        s.bounce = float(bounce)
        s.scroll_x, s.scroll_z = float(scroll[0]), float(scroll[1])
        s.time = float(time)
        s.waves = self.waves.ctypes.data if s.W else None
        if self.texels is not None:
            s.texels = self.texels.ctypes.data
            s.tex_h, s.tex_w = self.texels.shape
            s.tex_format = int(tex_format)
            s.tex_wrap = int(tex_wrap)
        else:
            s.texels = None
        self.scene = s


def vertices(sd, mode=MODE_ANALYTIC, first=0, count=None, precision="f64", nthreads=1, want=("pos", "nrm", "tan", "bit")):
    N = sd.scene.N
    if count is None:
        count = N * N - first
    dt = np.float32 if precision == "f32" else np.float64
    ct = ctypes.c_float if precision == "f32" else ctypes.c_double
    out = {k: (np.zeros((count, 3), dtype=dt) if k in want else None) for k in ("pos", "nrm", "tan", "bit")}
    fn = {"f64": lib().wto_vertices_f64, "f32": lib().wto_vertices_f32, "f80": lib().wto_vertices_f80}[precision]
    rc = fn(ctypes.byref(sd.scene), mode, first, count, _p(out["pos"], ct), _p(out["nrm"], ct), _p(out["tan"], ct),
            _p(out["bit"], ct), nthreads)
    if rc != 0:
        raise RuntimeError("oracle vertices failed")
    return out


def gerstner_point(sd, x, z):
    P, dx, dz = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().wto_gerstner_point(ctypes.byref(sd.scene), float(x), float(z), _p(P), _p(dx), _p(dz))
    return P, dx, dz


def displaced_point(sd, x, z):
    P = np.zeros(3)
    lib().wto_displaced_point(ctypes.byref(sd.scene), float(x), float(z), _p(P))
    return P


class FitData:
    def __init__(self, n, S, spacing, x0, z0, t0, dt, target, g=9.81):
        self.target = np.ascontiguousarray(np.asarray(target, dtype=np.float32).reshape(S, n * n, 3))
        f = Fit()
        f.n, f.S = int(n), int(S)
        f.spacing, f.x0, f.z0, f.t0, f.dt, f.g = float(spacing), float(x0), float(z0), float(t0), float(dt), float(g)
        f.target = self.target.ctypes.data
        self.fit = f


def fit_loss_grad(fd, params, nthreads=1, precision="f64", grad=True):
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1, 5))
    W = p.shape[0]
    g = np.zeros((W, 5)) if grad else None
    fn = lib().wto_fit_loss_grad_f64 if precision == "f64" else lib().wto_fit_loss_grad_f32
    loss = fn(ctypes.byref(fd.fit), _p(p), W, _p(g), nthreads)
    return loss, g


# ---------------------------------------------------------------- fragment stage
SHADE_UNIFORMS = ("fogColourFarAtCurrentTime", "fogDepthRadiusFar", "fogDepthRadiusNear",
                  "softEdgesDeltaDepthThreshold", "sunPosition", "sunShininess", "sunStrength",
                  "tintDeltaDepthThreshold", "viewportWidthHeight", "waterClarity", "zFar", "zNear")


class ShadeData:
    """The fragment shader's uniforms (dict keyed by the GLSL names) and textures:
    depth (h, w) float32, reflections / refractions (h, w, 4) uint8, skybox
    (6, L, L, 4) uint8. Keeps the arrays the Shade struct points to alive."""

    def __init__(self, uniforms, depth, reflections, refractions, skybox):
        self.depth = np.ascontiguousarray(depth, dtype=np.float32)
        self.refl = np.ascontiguousarray(reflections, dtype=np.uint8)
        self.refr = np.ascontiguousarray(refractions, dtype=np.uint8)
        self.sky = np.ascontiguousarray(skybox, dtype=np.uint8)
        u = uniforms
        s = Shade()
        s.fog[:] = [float(v) for v in u["fogColourFarAtCurrentTime"]]
        s.fog_far, s.fog_near = float(u["fogDepthRadiusFar"]), float(u["fogDepthRadiusNear"])
        s.soft_edges = float(u["softEdgesDeltaDepthThreshold"])
        s.sun[:] = [float(v) for v in u["sunPosition"]]
        s.shininess, s.strength = float(u["sunShininess"]), float(u["sunStrength"])
        s.tint = float(u["tintDeltaDepthThreshold"])
        s.viewport[:] = [float(v) for v in u["viewportWidthHeight"]]
        s.clarity, s.zfar, s.znear = float(u["waterClarity"]), float(u["zFar"]), float(u["zNear"])
        s.depth = self.depth.ctypes.data
        s.depth_h, s.depth_w = self.depth.shape
        s.refl = self.refl.ctypes.data
        s.refl_h, s.refl_w = self.refl.shape[:2]
        s.refr = self.refr.ctypes.data
        s.refr_h, s.refr_w = self.refr.shape[:2]
        s.sky = self.sky.ctypes.data
        s.sky_len = self.sky.shape[1]
        self.shade = s


def water_shade(sd, fragments, coverage=None, dst=None, precision="f64", nthreads=1, first=0, count=None, out=None):
    """Fragment colours of a (h, w, 12) float32 fragment buffer; returns (h, w, 4).
    Pixels outside [first, first + count) and uncovered pixels keep out's values (dst's
    with blending, NaN otherwise)."""
    frags = np.ascontiguousarray(fragments, dtype=np.float32)
    h, w = frags.shape[:2]
    cov = None if coverage is None else np.ascontiguousarray(coverage, dtype=np.uint8)
    dt = np.float32 if precision == "f32" else np.float64
    ct = ctypes.c_float if precision == "f32" else ctypes.c_double
    d = None if dst is None else np.ascontiguousarray(dst, dtype=dt)
    if out is None:
        out = d.copy() if d is not None else np.full((h, w, 4), np.nan, dtype=dt)
    if count is None:
        count = h * w - first
    fn = lib().wto_water_shade_f64 if precision == "f64" else lib().wto_water_shade_f32
    rc = fn(ctypes.byref(sd.shade), frags.ctypes.data, None if cov is None else cov.ctypes.data, w, first, count,
            _p(out, ct), _p(d, ct), nthreads)
    if rc != 0:
        raise RuntimeError("oracle water_shade failed")
    return out


def cube_face(r):
    s, t = ctypes.c_double(), ctypes.c_double()
    f = lib().wto_cube_face(_p(np.ascontiguousarray(r, dtype=np.float64)), ctypes.byref(s), ctypes.byref(t))
    return f, s.value, t.value


def sample_rgba8(texels, s, t):
    tex = np.ascontiguousarray(texels, dtype=np.uint8)
    out = np.zeros(4)
    lib().wto_sample_rgba8(tex.ctypes.data, tex.shape[1], tex.shape[0], float(s), float(t), _p(out))
    return out


def sample_cube(faces, r):
    f = np.ascontiguousarray(faces, dtype=np.uint8)
    out = np.zeros(4)
    lib().wto_sample_cube(f.ctypes.data, f.shape[1], _p(np.ascontiguousarray(r, dtype=np.float64)), _p(out))
    return out


def sample_depth(texels, s, t):
    tex = np.ascontiguousarray(texels, dtype=np.float32)
    return lib().wto_sample_depth(tex.ctypes.data, tex.shape[1], tex.shape[0], float(s), float(t))
