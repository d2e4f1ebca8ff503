"""Seeded random scenes through wt_water_displace against the FP64 oracle.

Each seed draws a camera (height, pitch, yaw), a wave count and set, a time up to one
hour, a grid side (including the smallest grids), a heightmap in one of the five
sampling modes with its scales, scroll and bounce, and asks for tangents and
bitangents. Tolerance as everywhere: positions within 1e-4 m, normals, tangents and
bitangents within 1e-4 rad, within 1 km of the camera.
"""
import numpy as np
import pytest

import wavetool_amd as wt
from wavetool_amd import scenes
from helpers import NRM_TOL, POS_TOL
from test_gpu_forward import check_parity, say

pytestmark = pytest.mark.gpu

MODES = [None, (wt.WT_TEXEL_R32F, wt.WT_WRAP_REPEAT), (wt.WT_TEXEL_R32F, wt.WT_WRAP_MIRRORED_REPEAT),
         (wt.WT_TEXEL_R8_UNORM, wt.WT_WRAP_REPEAT), (wt.WT_TEXEL_R8_UNORM, wt.WT_WRAP_MIRRORED_REPEAT)]
SIDES = [2, 3, 33, 128, 200]


def draw_scene(seed):
    rng = np.random.default_rng(1000 + seed)
    cam = scenes.camera_tuple(height=rng.uniform(2.0, 50.0), pitch_deg=rng.uniform(-35.0, -5.0),
                              yaw_deg=rng.uniform(0.0, 360.0))
    W = int(rng.integers(1, 97))
    waves = scenes.gen_waves(5000 + seed, W)
    t = float(rng.uniform(0.0, 3600.0))
    N = SIDES[(seed // len(MODES) + seed) % len(SIDES)]
    mode = MODES[seed % len(MODES)]
    texels, fmt, wrap = None, wt.WT_TEXEL_R32F, wt.WT_WRAP_REPEAT
    s_h = s_d = 0.0
    scroll = (0.0, 0.0)
    if mode is not None:
        fmt, wrap = mode
        h, w = int(rng.integers(8, 65)), int(rng.integers(8, 65))
        if fmt == wt.WT_TEXEL_R8_UNORM:
            texels = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
        else:
            texels = scenes.value_noise(seed, 64)[:h, :w].copy()
        s_h = float(np.float32(rng.uniform(0.01, 0.05)))
        s_d = float(np.float32(rng.uniform(0.02, 0.5)))
        scroll = (float(np.float32(rng.uniform(-0.5, 0.5))), float(np.float32(rng.uniform(-0.5, 0.5))))
    bounce = float(np.float32(rng.uniform(-0.1, 0.1)))
    cfg = dict(N=N, sample_scale=s_h, displacement_scale=s_d, scroll=scroll)
    return cam, waves, t, cfg, texels, fmt, wrap, bounce


@pytest.mark.parametrize("seed", range(15))
def test_random_scene(torch_cuda, wtctx, seed):
    cam, waves, t, cfg, texels, fmt, wrap, bounce = draw_scene(seed)
    say("random scene seed %d: N=%d W=%d t=%.2f texels=%s fmt=%d wrap=%d" % (
        seed, cfg["N"], waves.shape[0], t, None if texels is None else texels.shape, fmt, wrap))
    corners = wt.projector_corners(cam, scenes.DISPLACEMENT_BOUND)
    res = check_parity(torch_cuda, wtctx, cfg, corners, waves, cam[4:7], t, texels, True, fmt=fmt, wrap=wrap,
                       bounce=bounce)
    print(res, flush=True)
    assert res["pos_max"] <= POS_TOL and res["nrm_max"] <= NRM_TOL
    assert res["tan_max"] <= NRM_TOL and res["bit_max"] <= NRM_TOL
    assert res["n_excluded"] <= max(4, cfg["N"] ** 2 // 50)
