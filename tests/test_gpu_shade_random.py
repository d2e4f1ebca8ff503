"""Seeded random frames through wt_water_shade against the FP64 oracle.

Each seed draws a camera (height, pitch, yaw, field of view), the skybox side, the
resolution of the local reflection / refraction targets and every free uniform of the
fragment shader (sun direction and highlight, fog colour and radii, soft-edge and tint
thresholds, clarity), and checks the frame with the tolerance model of
test_gpu_shade.py.
"""
import math

import numpy as np
import pytest

from wavetool_amd import scenes
from test_gpu_shade import check_frame, comparison_mask, oracle_frame, run, say

pytestmark = pytest.mark.gpu


def f32(x):
    return float(np.float32(x))


def draw_frame(seed, width=128, height=72):
    rng = np.random.default_rng(2000 + seed)
    sc = scenes.shading_scene(width, height, seed=100 + seed, sky_length=int(rng.choice([16, 32, 64])),
                              local_scale=float(rng.choice([0.5, 1.0])), eye_height=rng.uniform(2.0, 30.0),
                              pitch_deg=rng.uniform(-40.0, -5.0), yaw_deg=rng.uniform(0.0, 360.0),
                              fov_deg=rng.uniform(40.0, 90.0))
    el = rng.uniform(-0.3, 1.3)
    az = rng.uniform(0.0, 2.0 * math.pi)
    sun = [f32(math.cos(el) * math.cos(az)), f32(math.sin(el)), f32(math.cos(el) * math.sin(az))]
    near = rng.uniform(0.0, 0.5)
    u = dict(sc["uniforms"])
    u.update(sunPosition=sun, sunShininess=f32(rng.uniform(1.0, 100.0)), sunStrength=f32(rng.uniform(0.0, 2.0)),
             fogColourFarAtCurrentTime=[f32(v) for v in rng.uniform(0.0, 1.0, size=4)],
             fogDepthRadiusNear=f32(near), fogDepthRadiusFar=f32(near + rng.uniform(0.05, 0.5)),
             softEdgesDeltaDepthThreshold=f32(rng.uniform(0.01, 0.2)),
             tintDeltaDepthThreshold=f32(rng.uniform(0.01, 0.2)), waterClarity=f32(rng.uniform(0.0, 1.0)))
    sc["uniforms"] = u
    return sc


@pytest.mark.parametrize("seed", range(8))
def test_random_frame(torch_cuda, wtctx, seed):
    torch = torch_cuda
    sc = draw_frame(seed)
    H, W = sc["coverage"].shape
    say("random shade frame seed %d: %dx%d, sky %d, local %s" % (seed, W, H, sc["skybox"].shape[1],
                                                                 sc["reflections"].shape[:2]))
    colour = torch.zeros((H, W, 4), device="cuda")
    got = run(torch, wtctx, sc, sc["coverage"], colour)
    ref = oracle_frame(sc, sc["coverage"])
    stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], sc["coverage"])
    # the face-edge, sky-threshold and fog-step exclusions leave out under 0.1% of the
    # covered pixels, as on the fixed frames
    assert stable.sum() > 0.999 * sc["coverage"].astype(bool).sum()
    check_frame(got, ref, stable, tol)
