"""ctypes binding of include/wt.h (libwavetool.so, built in-tree for gfx950).

The shared library links the HIP runtime by soname. torch is imported first so
that the library binds to the same runtime instance torch already loaded, and
device pointers from torch tensors are valid in both.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# WT_LIBWAVETOOL names another build of the library (the host-sanitizer build of the
# CPU test run); the default is the in-tree build
LIB_PATH = os.environ.get("WT_LIBWAVETOOL") or os.path.join(HERE, "libwavetool.so")

WT_OK = 0
WT_ERR_INVALID_ARG = 1
WT_ERR_NO_INTERSECTION = 2
WT_ERR_HIP = 3
WT_ERR_OOM = 4
WT_ERR_UNSUPPORTED = 5
WT_ERR_DEGENERATE = 6

WT_WRAP_REPEAT = 0
WT_WRAP_MIRRORED_REPEAT = 1
WT_TEXEL_R8_UNORM = 0
WT_TEXEL_R32F = 1

WT_FLOP_PER_VERTEX_WAVE = 26
WT_BYTES_PER_VERTEX_PN = 24
WT_BYTES_PER_VERTEX_PNTB = 48
WT_ADJOINT_FLOP_PER_VERTEX_WAVE_SNAPSHOT = 50
WT_MAX_WAVES = 1024
WT_MAX_VARYING_VERTICES = 2 * (2 ** 32 - 256)
WT_FIT_STATIC_SCRATCH_BYTES = 16 << 20


class GerstnerWave(ctypes.Structure):
    _fields_ = [
        ("amplitude_A", ctypes.c_float),
        ("frequency_w", ctypes.c_float),
        ("phaseConstant_phi", ctypes.c_float),
        ("steepness_Q_i", ctypes.c_float),
        ("xzDirection_D", ctypes.c_float * 2),
        ("phaseOffset_phi0", ctypes.c_float),
        ("reserved", ctypes.c_float),
    ]


class Camera(ctypes.Structure):
    _fields_ = [
        ("fov_deg", ctypes.c_float),
        ("aspect", ctypes.c_float),
        ("near_clip", ctypes.c_float),
        ("far_clip", ctypes.c_float),
        ("position", ctypes.c_float * 3),
        ("yaw_deg", ctypes.c_float),
        ("pitch_deg", ctypes.c_float),
    ]


class Heightmap(ctypes.Structure):
    _fields_ = [
        ("texels", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("format", ctypes.c_uint32),
        ("wrap", ctypes.c_uint32),
    ]


class WaterParams(ctypes.Structure):
    _fields_ = [
        ("corners", (ctypes.c_double * 4) * 4),
        ("cameraPosition", ctypes.c_float * 3),
        ("gridLength", ctypes.c_uint32),
        ("heightmapSampleScale", ctypes.c_float),
        ("heightmapDisplacementScale", ctypes.c_float),
        ("verticalBounceWaveDisplacement", ctypes.c_float),
        ("heightmapScrollVelocity", ctypes.c_float * 2),
        ("reserved", ctypes.c_float),
        ("waveAnimationTimeInSeconds", ctypes.c_double),
    ]


class VertexBuffers(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("tangent", ctypes.c_void_p),
        ("bitangent", ctypes.c_void_p),
    ]


class VaryingBuffers(ctypes.Structure):
    _fields_ = [
        ("clip_position", ctypes.c_void_p),
        ("ndc_xy_height0", ctypes.c_void_p),
        ("view_vector", ctypes.c_void_p),
        ("normal_view_only_yaw", ctypes.c_void_p),
        ("heightmap_sample", ctypes.c_void_p),
    ]


class LaunchConfig(ctypes.Structure):
    _fields_ = [
        ("wave_split", ctypes.c_uint32),
        ("steps_per_lane", ctypes.c_uint32),
        ("lanes_per_row", ctypes.c_uint32),
        ("wave_chunk", ctypes.c_uint32),
    ]


class Patch(ctypes.Structure):
    _fields_ = [
        ("corners", (ctypes.c_double * 4) * 4),
        ("cameraPosition", ctypes.c_float * 3),
        ("gridLength", ctypes.c_uint32),
        ("wave_offset", ctypes.c_uint32),
        ("wave_count", ctypes.c_uint32),
        ("vertex_offset", ctypes.c_uint64),
        ("waveAnimationTimeInSeconds", ctypes.c_double),
        ("heightmapSampleScale", ctypes.c_float),
        ("heightmapDisplacementScale", ctypes.c_float),
        ("verticalBounceWaveDisplacement", ctypes.c_float),
        ("heightmapScrollVelocity", ctypes.c_float * 2),
        ("reserved", ctypes.c_float),
    ]


class FitProblem(ctypes.Structure):
    _fields_ = [
        ("grid_side", ctypes.c_uint32),
        ("snapshots", ctypes.c_uint32),
        ("wave_count", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("spacing", ctypes.c_double),
        ("x0", ctypes.c_double),
        ("z0", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("target", ctypes.c_void_p),
        ("params", ctypes.c_void_p),
    ]


class AdamConfig(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float)]


class ShadeParams(ctypes.Structure):
    _fields_ = [("fogColourFarAtCurrentTime", ctypes.c_float * 4), ("fogDepthRadiusFar", ctypes.c_float),
                ("fogDepthRadiusNear", ctypes.c_float), ("softEdgesDeltaDepthThreshold", ctypes.c_float),
                ("sunPosition", ctypes.c_float * 3), ("sunShininess", ctypes.c_float), ("sunStrength", ctypes.c_float),
                ("tintDeltaDepthThreshold", ctypes.c_float), ("viewportWidthHeight", ctypes.c_float * 2),
                ("waterClarity", ctypes.c_float), ("zFar", ctypes.c_float), ("zNear", ctypes.c_float)]


class Rgba8Image(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class ShadeTextures(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_void_p), ("depth_width", ctypes.c_uint32), ("depth_height", ctypes.c_uint32),
                ("local_reflections", Rgba8Image), ("local_refractions", Rgba8Image), ("skybox", ctypes.c_void_p),
                ("skybox_length", ctypes.c_uint32)]


WT_BYTES_PER_FRAGMENT = 48  # sizeof(wt_water_fragment)


# (name, restype, argtypes) for every entry point declared in include/wt.h
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
PROTOTYPES = [
    ("wt_abi_version", ctypes.c_int, []),
    ("wt_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("wt_create", ctypes.c_int, [ctypes.POINTER(_vp)]),
    ("wt_destroy", None, [_vp]),
    ("wt_last_error", ctypes.c_char_p, [_vp]),
    ("wt_camera_matrices", ctypes.c_int, [ctypes.POINTER(Camera), _dp, _dp, _dp]),
    ("wt_projector_corners", ctypes.c_int, [ctypes.POINTER(Camera), ctypes.c_float, _dp]),
    ("wt_projector_corners_batched", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp, _vp]),
    ("wt_normalize_steepness", ctypes.c_int,
     [ctypes.POINTER(GerstnerWave), ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, ctypes.c_uint32]),
    ("wt_water_displace", ctypes.c_int,
     [_vp, ctypes.POINTER(WaterParams), _vp, ctypes.c_uint32, ctypes.POINTER(Heightmap),
      ctypes.POINTER(VertexBuffers), _vp]),
    ("wt_vertex_varyings", ctypes.c_int,
     [_vp, ctypes.POINTER(WaterParams), ctypes.POINTER(Heightmap), ctypes.POINTER(ctypes.c_float),
      ctypes.POINTER(ctypes.c_float), _vp, _vp, ctypes.c_uint64, ctypes.POINTER(VaryingBuffers), _vp]),
    ("wt_set_launch_config", ctypes.c_int, [_vp, ctypes.POINTER(LaunchConfig)]),
    ("wt_get_launch_config", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(LaunchConfig)]),
    ("wt_autotune_forward", ctypes.c_int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(LaunchConfig),
      ctypes.POINTER(ctypes.c_float)]),
    ("wt_autotune_forward_overlapped", ctypes.c_int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(LaunchConfig),
      ctypes.POINTER(ctypes.c_float)]),
    ("wt_autotune_batch", ctypes.c_int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
      ctypes.POINTER(ctypes.c_float)]),
    ("wt_autotune_save", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("wt_autotune_load", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("wt_graph_begin", ctypes.c_int, [_vp, _vp]),
    ("wt_graph_end", ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    ("wt_graph_launch", ctypes.c_int, [_vp, _vp]),
    ("wt_graph_destroy", None, [_vp]),
    ("wt_batch_create", ctypes.c_int, [_vp, ctypes.POINTER(Patch), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("wt_batch_destroy", None, [_vp]),
    ("wt_batch_vertex_count", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    ("wt_batch_patch_launch_config", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(LaunchConfig)]),
    ("wt_batch_set_cameras", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("wt_grid_triangle_indices", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp]),
    ("wt_water_shade", ctypes.c_int,
     [_vp, ctypes.POINTER(ShadeParams), ctypes.POINTER(ShadeTextures), _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp,
      ctypes.c_int, _vp]),
    ("wt_water_displace_batched", ctypes.c_int,
     [_vp, _vp, _vp, ctypes.c_uint32, ctypes.POINTER(Heightmap), ctypes.POINTER(VertexBuffers), ctypes.c_double, _vp]),
    ("wt_fit_create", ctypes.c_int, [_vp, ctypes.POINTER(FitProblem), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("wt_fit_destroy", None, [_vp]),
    ("wt_fit_reset", ctypes.c_int, [_vp, _vp]),
    ("wt_fit_iterate", ctypes.c_int, [_vp, ctypes.POINTER(AdamConfig), ctypes.c_uint32, _vp, _vp]),
    ("wt_fit_loss_grad", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("wt_fit_loss_grad_static", ctypes.c_int, [_vp, ctypes.POINTER(FitProblem), _vp, _vp, _vp]),
    ("wt_fit_synthesize", ctypes.c_int,
     [_vp, ctypes.POINTER(FitProblem), ctypes.c_float, ctypes.c_uint64, _vp, _vp]),
    ("wt_fit_phase_plan", ctypes.c_int,
     [ctypes.POINTER(FitProblem), ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int)]),
]

_lib = None


class WtError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__("wt status %d (%s)%s" % (status, STATUS_NAMES.get(status, "?"), (": " + msg) if msg else ""))


STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "no intersection", 3: "HIP error", 4: "out of memory",
                5: "unsupported", 6: "degenerate"}


def lib(load_torch=True):
    """Load libwavetool.so. Raises if the library has not been built: there is no
    fallback path."""
    global _lib
    if _lib is None:
        if load_torch:
            import torch  # noqa: F401  (binds the HIP runtime torch uses first)
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libwavetool.so is missing (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in PROTOTYPES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != WT_OK:
        msg = ""
        if ctx is not None:
            m = lib().wt_last_error(ctx)
            msg = m.decode() if m else ""
        raise WtError(status, msg)
