"""Python side of the MI355X water-displacement stage (ctypes over libwavetool.so).

The product path is the C ABI in include/wt.h; this module only marshals
arguments (device memory and streams come from torch tensors) and raises on
non-OK status. There is no CPU fallback: every compute call goes to the HIP
kernels in libwavetool.so.
"""
import ctypes

from . import scenes
from ._abi import (LaunchConfig, Camera, GerstnerWave, Heightmap, VertexBuffers, VaryingBuffers, WaterParams,
                   FitProblem, AdamConfig, Patch, ShadeParams, ShadeTextures, Rgba8Image, WtError, check, lib,
                   WT_OK, WT_ERR_INVALID_ARG, WT_ERR_NO_INTERSECTION, WT_ERR_DEGENERATE, WT_WRAP_REPEAT,
                   WT_WRAP_MIRRORED_REPEAT, WT_TEXEL_R8_UNORM, WT_TEXEL_R32F, WT_FLOP_PER_VERTEX_WAVE,
                   WT_BYTES_PER_VERTEX_PN, WT_BYTES_PER_VERTEX_PNTB, WT_BYTES_PER_FRAGMENT)

__all__ = ["Context", "camera", "projector_corners", "projector_corners_batched", "camera_rows", "water_params", "normalize_steepness",
           "shade_params", "scenes", "lib", "check", "WtError"]


def camera(cam_tuple):
    """Camera from (fov_deg, aspect, near, far, px, py, pz, yaw_deg, pitch_deg)."""
    c = Camera()
    c.fov_deg, c.aspect, c.near_clip, c.far_clip = cam_tuple[0:4]
    c.position[:] = list(cam_tuple[4:7])
    c.yaw_deg, c.pitch_deg = cam_tuple[7], cam_tuple[8]
    return c


def projector_corners(cam_tuple, amp):
    """World grid corners [4][4] (BL, TL, BR, TR) from the host projector (double)."""
    import numpy as np
    out = (ctypes.c_double * 16)()
    st = lib().wt_projector_corners(ctypes.byref(camera(cam_tuple)), float(amp), out)
    check(st)
    return np.array(out[:], dtype=np.float64).reshape(4, 4)


def camera_rows(cam_tuples):
    """(count, 9) float32 array laid out as wt_camera structs, one row per camera tuple
    (fov_deg, aspect, near, far, px, py, pz, yaw_deg, pitch_deg)."""
    import numpy as np
    return np.asarray([list(c) for c in cam_tuples], dtype=np.float32).reshape(-1, 9)


def projector_corners_batched(ctx, cameras, amplitudes, corners, status, stream=None):
    """Device projector over cameras (device float32 (count, 9), wt_camera rows) with
    slab half-heights amplitudes (device float32 (count,)); fills corners (device float64
    (count, 4, 4)) and status (device int32 (count,))."""
    import torch
    count = int(cameras.shape[0])
    _check_array(cameras, torch.float32, count * 9, "cameras")
    _check_array(amplitudes, torch.float32, count, "amplitudes")
    _check_array(corners, torch.float64, count * 16, "corners")
    _check_array(status, torch.int32, count, "status")
    ctx._check(lib().wt_projector_corners_batched(ctx.handle, _ptr(cameras), _ptr(amplitudes), count, _ptr(corners),
                                                  _ptr(status), _stream(stream)))


def camera_matrices(cam_tuple):
    import numpy as np
    view, proj, fwd = (ctypes.c_double * 16)(), (ctypes.c_double * 16)(), (ctypes.c_double * 3)()
    check(lib().wt_camera_matrices(ctypes.byref(camera(cam_tuple)), view, proj, fwd))
    return (np.array(view[:]).reshape(4, 4).T.copy(), np.array(proj[:]).reshape(4, 4).T.copy(),
            np.array(fwd[:]))


def normalize_steepness(waves, q, normaliser):
    """In-place Q_i = q / (w A normaliser) on a (W, 8) float32 wave array."""
    import numpy as np
    W = waves.shape[0]
    arr = (GerstnerWave * W)()
    ctypes.memmove(arr, waves.ctypes.data, W * 32)
    qa = np.ascontiguousarray(q, dtype=np.float32)
    check(lib().wt_normalize_steepness(arr, qa.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), W, normaliser))
    ctypes.memmove(waves.ctypes.data, arr, W * 32)
    return waves


def water_params(corners, cam_pos, N, time, sample_scale=0.0, displacement_scale=0.0, bounce=0.0,
                 scroll=(0.0, 0.0)):
    p = WaterParams()
    for i in range(4):
        for k in range(4):
            p.corners[i][k] = float(corners[i][k])
    p.cameraPosition[:] = [float(v) for v in cam_pos]
    p.gridLength = int(N)
    p.heightmapSampleScale = float(sample_scale)
    p.heightmapDisplacementScale = float(displacement_scale)
    p.verticalBounceWaveDisplacement = float(bounce)
    p.heightmapScrollVelocity[:] = [float(scroll[0]), float(scroll[1])]
    p.waveAnimationTimeInSeconds = float(time)
    return p


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def shade_params(uniforms):
    """wt_shade_params from a dict keyed by the fragment shader's uniform names."""
    p = ShadeParams()
    for name, _ in ShadeParams._fields_:
        v = uniforms[name]
        if isinstance(v, (list, tuple)) or hasattr(v, "__len__"):
            getattr(p, name)[:] = [float(x) for x in v]
        else:
            setattr(p, name, float(v))
    return p


def _check_array(t, dtype, numel, name):
    """The kernels index count entries of each array: refuse a tensor that holds fewer."""
    if t.dtype != dtype or not t.is_contiguous() or t.numel() != numel:
        raise ValueError("%s: need a contiguous %s tensor of %d elements, got %s %s" %
                         (name, dtype, numel, t.dtype, tuple(t.shape)))


def _check_min_array(t, dtype, numel, name):
    """A tensor the kernels index up to numel entries of (a view into a larger buffer
    is fine): refuse a wrong dtype, a non-contiguous view or one that holds fewer."""
    if t.dtype != dtype or not t.is_contiguous() or t.numel() < numel:
        raise ValueError("%s: need a contiguous %s tensor of at least %d elements, got %s %s" %
                         (name, dtype, numel, t.dtype, tuple(t.shape)))


def _check_waves(waves, count):
    import torch
    if count and (waves is None or waves.dim() != 2 or waves.shape[1] != 8 or waves.shape[0] < count):
        raise ValueError("waves: need a (>= %d, 8) tensor, got %s" % (count, None if waves is None
                                                                         else tuple(waves.shape)))
    if count:
        _check_min_array(waves, torch.float32, count * 8, "waves")


def _heightmap(heightmap, hm_format, hm_wrap):
    """wt_heightmap over a 2-D device tensor whose dtype matches the texel format
    (uint8 for WT_TEXEL_R8_UNORM, float32 for WT_TEXEL_R32F)."""
    import torch
    if heightmap is None:
        return None
    if heightmap.dim() != 2:
        raise ValueError("heightmap: need a 2-D (height, width) tensor, got %s" % (tuple(heightmap.shape),))
    want = {WT_TEXEL_R8_UNORM: torch.uint8, WT_TEXEL_R32F: torch.float32}.get(hm_format)
    if want is None:
        raise ValueError("heightmap: unknown texel format %r" % (hm_format,))
    h, w = heightmap.shape
    _check_array(heightmap, want, h * w, "heightmap")
    return Heightmap(heightmap.data_ptr(), w, h, hm_format, hm_wrap)


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Context:
    """Owns a wt_ctx bound to the current device."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().wt_create(ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().wt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        check(st, self.handle)

    def grid_indices(self, n, out, base=0, stream=None):
        """Write the triangle index list of an n x n grid (ids + base) into the device int32
        tensor `out` of 6 (n - 1)^2 elements (read it as uint32)."""
        import torch
        _check_array(out, torch.int32, 6 * (int(n) - 1) ** 2, "indices")
        self._check(lib().wt_grid_triangle_indices(self.handle, int(n), int(base), _ptr(out), _stream(stream)))

    def shade(self, uniforms, depth, reflections, refractions, skybox, fragments, colour, coverage=None, blend=False,
              stream=None):
        """The water fragment stage (wt_water_shade) on device tensors: depth (h, w)
        float32, reflections / refractions (h, w, 4) uint8, skybox (6, L, L, 4) uint8,
        fragments (H, W, 12) float32 (wt_water_fragment records), colour (H, W, 4)
        float32, coverage (H, W) uint8 or None. uniforms: dict by GLSL name or a
        ShadeParams."""
        import torch
        H, W = int(fragments.shape[0]), int(fragments.shape[1])
        _check_array(fragments, torch.float32, H * W * 12, "fragments")
        _check_array(colour, torch.float32, H * W * 4, "colour")
        if coverage is not None:
            _check_array(coverage, torch.uint8, H * W, "coverage")
        dh, dw = int(depth.shape[0]), int(depth.shape[1])
        _check_array(depth, torch.float32, dh * dw, "depth")
        imgs = []
        for name, t in (("reflections", reflections), ("refractions", refractions)):
            h, w = int(t.shape[0]), int(t.shape[1])
            _check_array(t, torch.uint8, h * w * 4, name)
            imgs.append(Rgba8Image(t.data_ptr(), w, h))
        L = int(skybox.shape[1])
        _check_array(skybox, torch.uint8, 6 * L * L * 4, "skybox")
        tx = ShadeTextures(depth.data_ptr(), dw, dh, imgs[0], imgs[1], skybox.data_ptr(), L)
        p = uniforms if isinstance(uniforms, ShadeParams) else shade_params(uniforms)
        self._check(lib().wt_water_shade(self.handle, ctypes.byref(p), ctypes.byref(tx), _ptr(fragments),
                                         _ptr(coverage) if coverage is not None else None, W, H, _ptr(colour),
                                         1 if blend else 0, _stream(stream)))

    def displace(self, params, waves, pos, nrm, tan=None, bit=None, heightmap=None, hm_format=WT_TEXEL_R32F,
                 hm_wrap=WT_WRAP_REPEAT, stream=None, wave_count=None):
        """waves: device tensor (W, 8) float32; pos/nrm/tan/bit: device float tensors
        with gridLength^2 * 3 elements; heightmap: device tensor (H, W) or None."""
        import torch
        W = int(waves.shape[0]) if wave_count is None else int(wave_count)
        _check_waves(waves, W)
        n3 = int(params.gridLength) ** 2 * 3
        for name, t in (("pos", pos), ("nrm", nrm), ("tan", tan), ("bit", bit)):
            if t is not None or name in ("pos", "nrm"):
                if t is None:
                    raise ValueError("%s is required" % name)
                _check_min_array(t, torch.float32, n3, name)
        vb = VertexBuffers(_ptr(pos), _ptr(nrm), _ptr(tan), _ptr(bit))
        hm = _heightmap(heightmap, hm_format, hm_wrap)
        st = lib().wt_water_displace(self.handle, ctypes.byref(params), _ptr(waves) if W else None, W,
                                     ctypes.byref(hm) if hm is not None else None, ctypes.byref(vb), _stream(stream))
        self._check(st)

    def varyings(self, params, view_projection, view_only_yaw, pos, nrm, vertex_count, clip=None, ndc0=None,
                 view=None, nview=None, sample=None, heightmap=None, hm_format=WT_TEXEL_R32F, hm_wrap=WT_WRAP_REPEAT,
                 stream=None):
        """The water program's varyings from positions / normals (wt_vertex_varyings).
        Matrices: 4x4 arrays in the usual row / column indexing (m[r][c]); they are
        passed column-major as the reference uploads them."""
        import numpy as np
        import torch
        n = int(vertex_count)
        for name, t, k in (("pos", pos, 3), ("nrm", nrm, 3), ("clip", clip, 4), ("ndc0", ndc0, 2), ("view", view, 3),
                           ("nview", nview, 3), ("sample", sample, 1)):
            if t is not None:
                _check_min_array(t, torch.float32, n * k, name)
        f16 = ctypes.c_float * 16
        vp = f16(*[float(v) for v in np.asarray(view_projection, dtype=np.float64).T.reshape(-1)])
        vy = f16(*[float(v) for v in np.asarray(view_only_yaw, dtype=np.float64).T.reshape(-1)])
        out = VaryingBuffers(_ptr(clip), _ptr(ndc0), _ptr(view), _ptr(nview), _ptr(sample))
        hm = _heightmap(heightmap, hm_format, hm_wrap)
        st = lib().wt_vertex_varyings(self.handle, ctypes.byref(params), ctypes.byref(hm) if hm is not None else None,
                                      vp, vy, _ptr(pos), _ptr(nrm), int(vertex_count), ctypes.byref(out),
                                      _stream(stream))
        self._check(st)

    def set_launch_config(self, wave_split=0, steps=0, lanes_per_row=0, wave_chunk=0):
        cfg = LaunchConfig(wave_split, steps, lanes_per_row, wave_chunk)
        self._check(lib().wt_set_launch_config(self.handle, ctypes.byref(cfg)))

    def get_launch_config(self, N, W):
        cfg = LaunchConfig()
        self._check(lib().wt_get_launch_config(self.handle, N, W, ctypes.byref(cfg)))
        return dict(wave_split=cfg.wave_split, steps=cfg.steps_per_lane, lanes_per_row=cfg.lanes_per_row,
                    wave_chunk=cfg.wave_chunk)

    def autotune(self, N, W, with_heightmap=True, frames_in_flight=1):
        """Tune the plain kernel's launch shape for (N, W); with frames_in_flight > 1 each
        shape is timed with that many frames' launches overlapping."""
        cfg = LaunchConfig()
        ms = ctypes.c_float()
        if frames_in_flight == 1:
            st = lib().wt_autotune_forward(self.handle, N, W, int(bool(with_heightmap)), ctypes.byref(cfg),
                                           ctypes.byref(ms))
        else:
            st = lib().wt_autotune_forward_overlapped(self.handle, N, W, int(bool(with_heightmap)),
                                                      int(frames_in_flight), ctypes.byref(cfg), ctypes.byref(ms))
        self._check(st)
        return dict(wave_split=cfg.wave_split, steps=cfg.steps_per_lane, lanes_per_row=cfg.lanes_per_row,
                    wave_chunk=cfg.wave_chunk), ms.value

    def autotune_batch(self, N, W, with_heightmap=True):
        """Batched segment shape (index) for the (N, W) size bucket, and its time (ms)."""
        shape = ctypes.c_uint32()
        ms = ctypes.c_float()
        self._check(lib().wt_autotune_batch(self.handle, int(N), int(W), int(bool(with_heightmap)),
                                            ctypes.byref(shape), ctypes.byref(ms)))
        return shape.value, ms.value

    def autotune_save(self, path):
        self._check(lib().wt_autotune_save(self.handle, str(path).encode()))

    def autotune_load(self, path):
        self._check(lib().wt_autotune_load(self.handle, str(path).encode()))

    def graph_begin(self, stream):
        self._check(lib().wt_graph_begin(self.handle, ctypes.c_void_p(stream.cuda_stream)))

    def graph_end(self, stream):
        g = ctypes.c_void_p()
        self._check(lib().wt_graph_end(self.handle, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(g)))
        return Graph(g)


def patch(corners, cam_pos, N, wave_offset, wave_count, vertex_offset, time, sample_scale=0.0,
          displacement_scale=0.0, bounce=0.0, scroll=(0.0, 0.0)):
    p = Patch()
    for i in range(4):
        for k in range(4):
            p.corners[i][k] = float(corners[i][k])
    p.cameraPosition[:] = [float(v) for v in cam_pos]
    p.gridLength, p.wave_offset, p.wave_count = int(N), int(wave_offset), int(wave_count)
    p.vertex_offset = int(vertex_offset)
    p.waveAnimationTimeInSeconds = float(time)
    p.heightmapSampleScale = float(sample_scale)
    p.heightmapDisplacementScale = float(displacement_scale)
    p.verticalBounceWaveDisplacement = float(bounce)
    p.heightmapScrollVelocity[:] = [float(scroll[0]), float(scroll[1])]
    return p


def np_corners(p):
    """Corners [4][4] of a Patch as a numpy array."""
    import numpy as np
    return np.ctypeslib.as_array(p.corners).copy()


def batch_inputs(specs, amp=2.0, sample_scale=0.02, displacement_scale=0.05, scroll=(0.3, 0.0)):
    """Patch structs and the concatenated wave array for specs from
    scenes.batched_patches (each patch's corners from its own camera)."""
    import numpy as np
    patches, waves = [], []
    woff = voff = 0
    for s in specs:
        corners = projector_corners(s["camera"], amp)
        w = scenes.gen_waves(s["wave_seed"], s["W"])
        patches.append(patch(corners, s["camera"][4:7], s["N"], woff, s["W"], voff, s["time"], sample_scale,
                             displacement_scale, 0.0, scroll))
        waves.append(w)
        woff += s["W"]
        voff += s["N"] * s["N"]
    return patches, np.concatenate(waves, axis=0), voff


class Batch:
    """A batch of patches for the batched forward kernels (wt_batch)."""

    def __init__(self, ctx, patches):
        self.ctx = ctx
        self.count = len(patches)
        arr = (Patch * len(patches))(*patches)
        h = ctypes.c_void_p()
        ctx._check(lib().wt_batch_create(ctx.handle, arr, len(patches), ctypes.byref(h)))
        self.handle = h
        v = ctypes.c_uint64()
        check(lib().wt_batch_vertex_count(h, ctypes.byref(v)))
        self.vertices = v.value

    def displace(self, waves, pos, nrm, tan=None, bit=None, heightmap=None, hm_format=WT_TEXEL_R32F,
                 hm_wrap=WT_WRAP_REPEAT, time_offset=0.0, stream=None):
        import torch
        _check_waves(waves, int(waves.shape[0]))
        n3 = int(self.vertices) * 3
        for name, t in (("pos", pos), ("nrm", nrm), ("tan", tan), ("bit", bit)):
            if t is not None or name in ("pos", "nrm"):
                if t is None:
                    raise ValueError("%s is required" % name)
                _check_min_array(t, torch.float32, n3, name)
        vb = VertexBuffers(_ptr(pos), _ptr(nrm), _ptr(tan), _ptr(bit))
        hm = _heightmap(heightmap, hm_format, hm_wrap)
        st = lib().wt_water_displace_batched(self.ctx.handle, self.handle, _ptr(waves), int(waves.shape[0]),
                                             ctypes.byref(hm) if hm is not None else None, ctypes.byref(vb),
                                             float(time_offset), _stream(stream))
        self.ctx._check(st)

    def set_cameras(self, cameras, amplitudes, status=None, stream=None):
        """Re-aim every patch from its own camera on the device (wt_batch_set_cameras):
        cameras device float32 (patches, 9), amplitudes device float32 (patches,),
        status optional device int32 (patches,)."""
        import torch
        _check_array(cameras, torch.float32, self.count * 9, "cameras")
        _check_array(amplitudes, torch.float32, self.count, "amplitudes")
        if status is not None:
            _check_array(status, torch.int32, self.count, "status")
        self.ctx._check(lib().wt_batch_set_cameras(self.ctx.handle, self.handle, _ptr(cameras), _ptr(amplitudes),
                                                   _ptr(status), _stream(stream)))

    def patch_launch_config(self, i):
        """(wave_split, steps, lanes_per_row, wave_chunk) the batch runs patch i with."""
        cfg = LaunchConfig()
        check(lib().wt_batch_patch_launch_config(self.handle, int(i), ctypes.byref(cfg)))
        return (cfg.wave_split, cfg.steps_per_lane, cfg.lanes_per_row, cfg.wave_chunk)

    def close(self):
        if self.handle:
            lib().wt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_problem(n, S, W, spacing, x0, z0, t0, dt, target, params, gravity=9.81):
    """FitProblem over device tensors target (S, n*n, 3) and params (W, 5), float32."""
    p = FitProblem()
    p.grid_side, p.snapshots, p.wave_count = int(n), int(S), int(W)
    p.spacing, p.x0, p.z0, p.t0, p.dt, p.gravity = (float(v) for v in (spacing, x0, z0, t0, dt, gravity))
    p.target = None if target is None else target.data_ptr()
    p.params = params.data_ptr()
    return p


def synthesize(ctx, problem, sigma, seed, target, stream=None):
    """Fill the device tensor `target` with the displacement of problem.params plus noise."""
    ctx._check(lib().wt_fit_synthesize(ctx.handle, ctypes.byref(problem), float(sigma), int(seed),
                                       ctypes.c_void_p(target.data_ptr()), _stream(stream)))


def fit_loss_grad_static(ctx, problem, grad, loss=None, stream=None):
    """Loss and gradient of one problem through wt_fit_loss_grad_static (no wt_fit, no
    allocation): grad a device float32 tensor of W*5 elements, loss None or one element."""
    import torch
    _check_array(grad, torch.float32, int(problem.wave_count) * 5, "grad")
    if loss is not None:
        _check_array(loss, torch.float32, 1, "loss")
    ctx._check(lib().wt_fit_loss_grad_static(ctx.handle, ctypes.byref(problem), _ptr(grad), _ptr(loss),
                                              _stream(stream)))


def phase_plan(problem, budget_bytes=0):
    b = ctypes.c_uint64()
    r = ctypes.c_int()
    check(lib().wt_fit_phase_plan(ctypes.byref(problem), int(budget_bytes), ctypes.byref(b), ctypes.byref(r)))
    return b.value, bool(r.value)


class Fit:
    """A batch of heightfield-fitting problems (wt_fit)."""

    def __init__(self, ctx, problems):
        self.ctx = ctx
        self.count = len(problems)
        arr = (FitProblem * self.count)(*problems)
        h = ctypes.c_void_p()
        ctx._check(lib().wt_fit_create(ctx.handle, arr, self.count, ctypes.byref(h)))
        self.handle = h

    def reset(self, stream=None):
        self.ctx._check(lib().wt_fit_reset(self.handle, _stream(stream)))

    def iterate(self, iterations, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8, losses=None, stream=None):
        cfg = AdamConfig(lr, beta1, beta2, eps)
        self.ctx._check(lib().wt_fit_iterate(self.handle, ctypes.byref(cfg), int(iterations), _ptr(losses),
                                             _stream(stream)))

    def loss_grad(self, grad, losses=None, stream=None):
        self.ctx._check(lib().wt_fit_loss_grad(self.handle, _ptr(grad), _ptr(losses), _stream(stream)))

    def close(self):
        if self.handle:
            lib().wt_fit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        check(lib().wt_graph_launch(self.handle, ctypes.c_void_p(stream.cuda_stream)))

    def close(self):
        if self.handle:
            lib().wt_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
