"""The reference's default scene at its own size on its own heightmap asset, against the
FP64 oracle.

Scene: the four default waves (src/render-engine.cpp:46-49) with the upload
normalisation (:1073), the default camera (:52; fov 72 deg, aspect 1 for the 1024 x 1024
window, near 0.1, far 100, at (0, 4, 70) looking down -z), the 513 x 513 grid
(src/program.cpp:688), s_h = 0.02 and s_d = 1.0 (src/render-engine.h:224-225), the
asset assets/textures/noise/waves/waves3/00.png as 8-bit texels under
GL_MIRRORED_REPEAT (src/texture.cpp:63-64; texels from the committed fixture), and the
vertical bounce 0.1 sin(2 pi phase) with the default 3 s period (render-engine.h:219,
237; program.cpp:452-464, phase = t / 3 mod 1).
"""
import importlib.util
import math
import os

import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from helpers import NRM_TOL, POS_TOL, compare, near_texel_edge

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
NTHREADS = min(16, os.cpu_count() or 1)
N_REF = 513


def say(msg):
    print("[case] " + msg, flush=True)


def maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def bounce_at(t, period=3.0, amplitude=0.1):
    f32 = np.float32
    phase = f32(math.fmod(t / period, 1.0))
    return float(f32(amplitude) * f32(math.sin(float(phase) * 2.0 * math.pi)))


@pytest.mark.parametrize("t", [0.0, 10.0, 3600.0])
def test_reference_default_scene_on_the_asset(torch_cuda, wtctx, t):
    torch = torch_cuda
    m = maker()
    texels = np.load(os.path.join(GOLDEN, "waves3_00_r8.npz"))["texels"]
    assert texels.shape == (512, 512) and texels.dtype == np.uint8
    waves = m.default_waves()
    cam = m.DEFAULT_CAMERA
    corners = wt.projector_corners(cam, m.DEFAULT_AMP)
    b = bounce_at(t)
    say("reference default scene N=%d t=%g bounce=%.4f, waves3/00.png R8 mirrored" % (N_REF, t, b))
    params = wt.water_params(corners, cam[4:7], N_REF, t, 0.02, 1.0, b)
    dev = torch.device("cuda")
    V = N_REF * N_REF
    bufs = [torch.full((V * 3,), float("nan"), device=dev) for _ in range(4)]
    wtctx.displace(params, torch.from_numpy(waves).to(dev), *bufs, heightmap=torch.from_numpy(texels).to(dev),
                   hm_format=wt.WT_TEXEL_R8_UNORM, hm_wrap=wt.WT_WRAP_MIRRORED_REPEAT)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy().reshape(V, 3) for k, v in zip(("pos", "nrm", "tan", "bit"), bufs)}
    sd = orc.SceneData(corners, cam[4:7], N_REF, waves, t, 0.02, 1.0, b, (0.0, 0.0), texels, orc.TEX_R8_UNORM,
                       orc.WRAP_MIRRORED_REPEAT)
    ref = orc.vertices(sd, nthreads=NTHREADS)
    edge = near_texel_edge(ref["pos"], 0.02, (0.0, 0.0), t, 512, 512)
    res = compare(out, ref, cam[4:7], edge)
    print(res, flush=True)
    # every vertex of this scene is within 1 km (far plane 100 m)
    assert res["n_near"] == V
    assert res["pos_max"] <= POS_TOL and res["nrm_max"] <= NRM_TOL
    assert res["tan_max"] <= NRM_TOL and res["bit_max"] <= NRM_TOL
    assert res["n_excluded"] <= V // 50
    # the bump is live: the asset moves the surface by up to s_d * 2 * (207/255 - 0.5)
    flat = orc.vertices(orc.SceneData(corners, cam[4:7], N_REF, waves, t, 0.0, 0.0, b), nthreads=NTHREADS,
                        want=("pos",))
    bump = ref["pos"][:, 1] - flat["pos"][:, 1]
    assert np.abs(bump).max() > 0.5
