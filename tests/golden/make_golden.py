"""Writes the golden fixtures of the water-displacement path from the FP64 oracle.

The reference ships no tests or recorded outputs, so these fixtures are regression
guards produced by the oracle (whose correctness is pinned by tests/test_oracle_*.py),
not independent truth. Run from the repository root:  python tests/golden/make_golden.py
Fixtures (little-endian, numpy .npz without pickles, plus a JSON header):
  default_scene.npz   the reference's default scene (src/render-engine.cpp:46-52,
                      src/camera.cpp:40-45, src/render-engine.h:214-237) at N = 17,
                      t in {0, 10, 3600}, analytic and finite-difference normals, with a
                      16x16 8-bit MIRRORED_REPEAT heightmap (s_h = 0.02, s_d = 1.0)
  w4_t10_n17.npz      the 256x256 scene's waves (4 seeded waves, t = 10 s) on a 17x17 grid, N/T/B
  bilinear.npz        64 samples of a 16x16 R8 texture under REPEAT and MIRRORED_REPEAT
  adjoint_small.npz   loss and gradient of a 16x16 x 2 snapshots x 3 waves fit
  waves3_00_r8.npz    the reference's own heightmap asset, assets/textures/noise/waves/
                      waves3/00.png (512 x 512 grey), decoded as RenderEngine::load2DTexture
                      does it (src/render-engine.cpp:1308-1322: stb_image forcing RGBA,
                      rows flipped so row 0 is the image's bottom), red channel kept as the
                      R8 texels the water program samples; decoded with PIL (PNG is
                      lossless, so any decoder gives these bytes)
Run with fixture names as arguments to rewrite only those.
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "wave-tool_amd")]

import oracle as orc  # noqa: E402
from wavetool_amd import scenes  # noqa: E402


def default_waves():
    """render-engine.cpp:46-49 with the upload normalisation of :1073 (Count() = 4)."""
    d1 = 1.0 / math.sqrt(2.0)
    raw = [(0.06, 1.0, 2.0, 1.0, (1.0, 0.0)), (0.1, 1.0, 0.2, 0.0, (d1, d1)),
           (0.0, 0.0, 0.0, 0.0, (0.0, 1.0)), (0.0, 0.0, 0.0, 0.0, (0.0, 1.0))]
    w = np.zeros((4, 8), np.float32)
    for i, (A, fw, phi, q, D) in enumerate(raw):
        wa = np.float32(np.float32(fw) * np.float32(A))
        w[i] = [A, fw, phi, np.float32(q) / (wa * np.float32(4)) if wa != 0 else 0.0, D[0], D[1], 0.0, 0.0]
    return w


DEFAULT_CAMERA = (72.0, 1.0, 0.1, 100.0, 0.0, 4.0, 70.0, -90.0, 0.0)
DEFAULT_AMP = 0.16 + 1.0 + 0.1  # TotalAmplitude + heightmapDisplacementScale + verticalBounceWaveAmplitude (:803)


def texture16():
    return ((np.arange(256).reshape(16, 16) * 37 + 11) % 256).astype(np.uint8)


def make_default_scene():
    st, corners, _ = orc.projector(DEFAULT_CAMERA, DEFAULT_AMP)
    assert st == 0
    waves = default_waves()
    tex = texture16()
    out = dict(corners=corners, waves=waves, texels=tex)
    for t in (0.0, 10.0, 3600.0):
        sd = orc.SceneData(corners, DEFAULT_CAMERA[4:7], 17, waves, t, 0.02, 1.0, 0.0, (0.0, 0.0), tex,
                           orc.TEX_R8_UNORM, orc.WRAP_MIRRORED_REPEAT)
        a = orc.vertices(sd, mode=orc.MODE_ANALYTIC)
        f = orc.vertices(sd, mode=orc.MODE_REFERENCE, want=("pos", "nrm"))
        for k in ("pos", "nrm", "tan", "bit"):
            out["t%g_%s" % (t, k)] = a[k]
        out["t%g_fd_nrm" % t] = f["nrm"]
    return out, dict(camera=DEFAULT_CAMERA, amp=DEFAULT_AMP, N=17, times=[0.0, 10.0, 3600.0], sample_scale=0.02,
                     displacement_scale=1.0, wrap="MIRRORED_REPEAT", format="R8_UNORM")


def make_w4_t10():
    cfg = scenes.scene_256_w4()
    st, corners, _ = orc.projector(cfg["camera"], cfg["amp"])
    waves = scenes.gen_waves(cfg["seed"], cfg["W"])
    sd = orc.SceneData(corners, cfg["camera"][4:7], 17, waves, cfg["t0"])
    a = orc.vertices(sd)
    return dict(corners=corners, waves=waves, **a), dict(camera=cfg["camera"], amp=cfg["amp"], N=17, t=cfg["t0"],
                                                         seed=cfg["seed"])


def make_bilinear():
    tex = texture16()
    pts = np.random.default_rng(2024).uniform(-2.5, 2.5, size=(64, 2))
    out = dict(texels=tex, points=pts)
    for name, wrap in (("repeat", orc.WRAP_REPEAT), ("mirrored", orc.WRAP_MIRRORED_REPEAT)):
        vals = np.array([orc.sample_bilinear(tex, orc.TEX_R8_UNORM, wrap, s, t) for s, t in pts])
        out[name] = vals
    return out, dict(columns=["value", "d/ds", "d/dt"])


def make_adjoint():
    prob = scenes.fit_problem(77, n=16, S=2, W_hidden=3, W_fit=3)
    tgt = scenes.fit_target(prob)
    fd = orc.FitData(prob["n"], prob["S"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], tgt)
    L, g = orc.fit_loss_grad(fd, prob["init"])
    return dict(target=tgt, params=prob["init"], loss=np.array([L]), grad=g), dict(
        n=16, S=2, W=3, spacing=prob["spacing"], x0=prob["x0"], z0=prob["z0"], t0=prob["t0"], dt=prob["dt"])


REFERENCE_ASSET = "/root/reference/assets/textures/noise/waves/waves3/00.png"


def stb_like_red(img):
    """What stb_image with STBI_rgb_alpha and vertical flip hands the reference's texture
    upload, red channel only: grey g -> (g, g, g, 255), palette -> its RGB, 16-bit -> the
    high byte; rows flipped so row 0 is the image's bottom."""
    if img.mode in ("I;16", "I;16B"):
        a = (np.asarray(img, dtype=np.uint16) >> 8).astype(np.uint8)
    else:
        a = np.asarray(img.convert("RGBA"))[:, :, 0]
    return a[::-1].copy()


def make_waves3_asset():
    import hashlib
    from PIL import Image
    raw = open(REFERENCE_ASSET, "rb").read()
    img = Image.open(REFERENCE_ASSET)
    tex = stb_like_red(img)
    return dict(texels=tex), dict(source="assets/textures/noise/waves/waves3/00.png", mode=img.mode,
                                  width=int(img.size[0]), height=int(img.size[1]),
                                  png_sha256=hashlib.sha256(raw).hexdigest(),
                                  texels_sha256=hashlib.sha256(tex.tobytes()).hexdigest())


MAKERS = {"default_scene": make_default_scene, "w4_t10_n17": make_w4_t10, "bilinear": make_bilinear,
          "adjoint_small": make_adjoint, "waves3_00_r8": make_waves3_asset}


def main(names):
    path = os.path.join(HERE, "golden.json")
    header = json.load(open(path)) if names and os.path.exists(path) else {}
    for name, fn in MAKERS.items():
        if names and name not in names:
            continue
        arrays, meta = fn()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays) if name == "waves3_00_r8" else \
            np.savez(os.path.join(HERE, name + ".npz"), **arrays)
        header[name] = meta
    with open(path, "w") as f:
        json.dump(header, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
