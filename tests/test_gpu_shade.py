"""GPU parity of the fragment stage (wt_water_shade, assets/shaders/water-grid.frag:70-162)
against the FP64 oracle (oracle/wt_oracle_shade.hpp) on synthetic water frames.

Tolerance. The kernel computes in FP32, as the shader does. Three places amplify FP32
rounding, all in the shader's own formulas:
- linearizeDepth's denominator zFar - d (zFar - zNear) cancels as d -> 1, so the
  linear depth of a far fragment carries an error of about eps32 * zFar * lin / den;
  the soft-edge and tint factors scale that by 1 / threshold (20 and 10 at the
  defaults);
- the sun term pow(x, 50) scales the relative error of x by 50;
- FP32 texture coordinates (about 1e-7 relative) move a bilinear lookup by up to
  4e-4 texel on a 3840-texel target; across the local reflection's 0/255 alpha edges
  that is up to 4e-4 in alpha, weighted by the Fresnel factor in the colour (about
  2e-5 at the benchmark's viewing angles).
The per-pixel bound is 5e-5 + 2 s (e_back + e_frag), with e the linear-depth error
bound above and s = 1 + 1 / soft-edge threshold + 0.6 / tint threshold while the depth
difference is under those thresholds (0.6 is the largest deep-to-shallow tint step).
On the benchmark frame the kernel's largest error is 0.65 of the bound (0.09 at the
99.99th percentile) before the texture-coordinate term was added. Pixels
whose reflected direction sits within 1e-4 of a cube-face edge (the face choice flips
between FP32 and FP64 there, and the texels on either side differ), whose back depth
sits at the sky threshold, or, with equal fog radii, whose view distance sits at the
fog step, are left out; they are under 0.1% of a frame. The median error must also
stay below 1e-5 (a few FP32 ulps of the colour), so a systematic slip cannot hide under
the far-field bound."""
import ctypes

import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from wavetool_amd import scenes

pytestmark = pytest.mark.gpu
EPS32 = float(np.finfo(np.float32).eps)


def say(msg):
    print("[case] " + msg, flush=True)


def upload(torch, sc, depth=None):
    dev = torch.device("cuda")
    d = sc["depth"] if depth is None else depth
    return dict(depth=torch.from_numpy(np.ascontiguousarray(d)).to(dev),
                reflections=torch.from_numpy(sc["reflections"]).to(dev),
                refractions=torch.from_numpy(sc["refractions"]).to(dev),
                skybox=torch.from_numpy(sc["skybox"]).to(dev),
                fragments=torch.from_numpy(sc["fragments"]).to(dev))


def lin_error(u, d):
    """Bound on the FP32 error of linearizeDepth(d): two roundings at the scale of zFar
    in the denominator (half an ulp each), relative to what the cancellation leaves,
    plus a few relative roundings."""
    n, f = u["zNear"], u["zFar"]
    den = f - d * (f - n)
    lin = n * d / den
    ulp_f = EPS32 * 2.0 ** np.floor(np.log2(f))
    return lin * (ulp_f / den + 4 * EPS32)


def comparison_mask(sc, u, depth, coverage):
    """(stable pixels, per-pixel tolerance) over the frame."""
    fr = sc["fragments"].astype(np.float64)
    H, W = fr.shape[:2]
    n, v = fr[..., 0:3], fr[..., 3:6]
    ln = np.linalg.norm(v, axis=-1, keepdims=True)
    ln[ln == 0] = 1.0
    vv = v / ln
    R = -vv + 2.0 * np.sum(n * vv, axis=-1, keepdims=True) * n
    a = np.sort(np.abs(R), axis=-1)
    face_ok = (a[..., 2] - a[..., 1]) > 1e-4 * a[..., 2]
    rows, cols = np.mgrid[0:H, 0:W]
    dh, dw = depth.shape
    di = np.clip(np.floor((cols + 0.5) / u["viewportWidthHeight"][0] * dw).astype(int), 0, dw - 1)
    dj = np.clip(np.floor((rows + 0.5) / u["viewportWidthHeight"][1] * dh).astype(int), 0, dh - 1)
    back = depth.astype(np.float64)[dj, di]
    nz, fz = u["zNear"], u["zFar"]
    lin_back = nz * back / (fz - back * (fz - nz))
    eb, ef = lin_error(u, back), lin_error(u, fr[..., 11])
    sky_ok = np.abs(lin_back - 0.999) > 2 * eb + 1e-7
    cov = np.ones((H, W), dtype=bool) if coverage is None else coverage.astype(bool)
    # how strongly the colour depends on the depth difference: the soft-edge and tint
    # ramps while the difference is under their thresholds, the refraction weight always
    lin_frag = nz * fr[..., 11] / (fz - fr[..., 11] * (fz - nz))
    delta = np.clip(lin_back - lin_frag, 0.0, 1.0)
    sens = np.full((H, W), 1.0)
    for key, weight in (("softEdgesDeltaDepthThreshold", 1.0), ("tintDeltaDepthThreshold", 0.6)):
        thr = u[key]
        if thr > 0:
            sens += np.where(delta < thr + 2 * (eb + ef), weight / thr, 0.0)
    tol = 5e-5 + 2.0 * sens * (eb + ef)
    fog_ok = np.ones((H, W), dtype=bool)
    if u["fogDepthRadiusNear"] == u["fogDepthRadiusFar"]:
        fog_ok = np.abs(ln[..., 0] / fz - u["fogDepthRadiusFar"]) > 1e-5
    return cov & face_ok & sky_ok & fog_ok, tol


def check_frame(got, ref, stable, tol):
    err = np.abs(got - ref).max(axis=-1)
    assert stable.sum() > 0
    bad = stable & ~(err <= tol)
    assert not bad.any(), "%d pixels over tolerance, worst %.3g at %s" % (
        bad.sum(), err[bad].max(), np.argwhere(bad)[0])
    assert np.median(err[stable]) < 1e-5


def run(torch, ctx, sc, coverage, colour, blend=False, depth=None):
    t = upload(torch, sc, depth)
    cov = None if coverage is None else torch.from_numpy(np.ascontiguousarray(coverage)).cuda()
    ctx.shade(sc["uniforms"], t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"], colour,
              coverage=cov, blend=blend)
    torch.cuda.synchronize()
    return colour.cpu().numpy().astype(np.float64)


def oracle_frame(sc, coverage, dst=None, depth=None, uniforms=None):
    sd = orc.ShadeData(uniforms or sc["uniforms"], sc["depth"] if depth is None else depth, sc["reflections"],
                       sc["refractions"], sc["skybox"])
    return orc.water_shade(sd, sc["fragments"], coverage, dst=dst, nthreads=8)


def test_shade_matches_oracle(torch_cuda, wtctx):
    torch = torch_cuda
    say("shade 192x108 sky 64, coverage mask")
    sc = scenes.shading_scene(192, 108, seed=1, sky_length=64)
    H, W = 108, 192
    colour = torch.full((H, W, 4), -7.0, device="cuda")
    got = run(torch, wtctx, sc, sc["coverage"], colour)
    ref = oracle_frame(sc, sc["coverage"])
    cov = sc["coverage"].astype(bool)
    assert (got[~cov] == -7.0).all()  # uncovered pixels untouched
    stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], sc["coverage"])
    assert stable.sum() > 0.999 * cov.sum()
    check_frame(got, ref, stable, tol)


def test_shade_blend_half_resolution_textures(torch_cuda, wtctx):
    """Blending over a destination, local textures at half the viewport and a
    half-resolution depth texture (GL_NEAREST at the pixel centre)."""
    torch = torch_cuda
    say("shade 160x96 blend, half-res local + depth textures")
    sc = scenes.shading_scene(160, 96, seed=2, sky_length=48, local_scale=0.5)
    depth = np.ascontiguousarray(sc["depth"][::2, ::2])
    dst = np.random.default_rng(3).uniform(0.0, 1.0, size=(96, 160, 4))
    colour = torch.from_numpy(dst.astype(np.float32)).cuda()
    got = run(torch, wtctx, sc, sc["coverage"], colour, blend=True, depth=depth)
    ref = oracle_frame(sc, sc["coverage"], dst=dst.astype(np.float32).astype(np.float64), depth=depth)
    cov = sc["coverage"].astype(bool)
    assert (got[~cov] == dst.astype(np.float32)[~cov]).all()
    stable, tol = comparison_mask(sc, sc["uniforms"], depth, sc["coverage"])
    check_frame(got, ref, stable, tol)


def test_shade_every_pixel_without_mask(torch_cuda, wtctx):
    """A camera looking steeply down sees water in every pixel: coverage NULL."""
    torch = torch_cuda
    say("shade 128x128 no coverage mask")
    sc = scenes.shading_scene(128, 128, seed=4, sky_length=32, pitch_deg=-70.0, eye_height=5.0)
    assert sc["coverage"].all()
    colour = torch.empty((128, 128, 4), device="cuda")
    got = run(torch, wtctx, sc, None, colour)
    ref = oracle_frame(sc, None)
    stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], None)
    check_frame(got, ref, stable, tol)


@pytest.mark.parametrize("override", [
    dict(tintDeltaDepthThreshold=0.0, softEdgesDeltaDepthThreshold=0.0),
    dict(sunShininess=0.0, sunStrength=0.7),
    dict(fogDepthRadiusNear=0.3, fogDepthRadiusFar=0.3, fogColourFarAtCurrentTime=[0.2, 0.4, 0.6, 0.9]),
    dict(waterClarity=1.0, sunPosition=[0.0, -1.0, 0.0]),
])
def test_shade_uniform_edge_cases(torch_cuda, wtctx, override):
    torch = torch_cuda
    say("shade uniform case %s" % sorted(override))
    sc = scenes.shading_scene(96, 64, seed=5, sky_length=32)
    sc["uniforms"] = dict(sc["uniforms"], **override)
    colour = torch.zeros((64, 96, 4), device="cuda")
    got = run(torch, wtctx, sc, sc["coverage"], colour)
    ref = oracle_frame(sc, sc["coverage"])
    stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], sc["coverage"])
    check_frame(got, ref, stable, tol)


def test_shade_single_pixel_and_edges(torch_cuda, wtctx):
    """1 x 1 and 1 x N frames: clamp-to-edge lookups at the image borders."""
    torch = torch_cuda
    say("shade 1x1 and 7x1 frames")
    for W, H in ((1, 1), (7, 1), (1, 5)):
        sc = scenes.shading_scene(W, H, seed=6, sky_length=4, pitch_deg=-80.0)
        colour = torch.zeros((H, W, 4), device="cuda")
        got = run(torch, wtctx, sc, sc["coverage"], colour)
        ref = oracle_frame(sc, sc["coverage"])
        stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], sc["coverage"])
        if stable.any():
            check_frame(got, ref, stable, tol)


def test_shade_graph_replay(torch_cuda, wtctx):
    """Captured in a HIP graph, replays match the eager call bit for bit."""
    torch = torch_cuda
    say("shade graph replay 128x72")
    sc = scenes.shading_scene(128, 72, seed=8, sky_length=32)
    t = upload(torch, sc)
    cov = torch.from_numpy(sc["coverage"]).cuda()
    eager = torch.zeros((72, 128, 4), device="cuda")
    wtctx.shade(sc["uniforms"], t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"], eager,
                coverage=cov)
    out = torch.zeros_like(eager)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    wtctx.graph_begin(stream)
    wtctx.shade(sc["uniforms"], t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"], out,
                coverage=cov, stream=stream)
    graph = wtctx.graph_end(stream)
    for k in range(3):
        out.zero_()
        torch.cuda.synchronize()
        graph.launch(stream)
        torch.cuda.synchronize()
        assert torch.equal(out, eager), k
    graph.close()


def test_shade_rejects_bad_arguments(torch_cuda, wtctx):
    torch = torch_cuda
    sc = scenes.shading_scene(16, 16, seed=9, sky_length=4)
    t = upload(torch, sc)
    colour = torch.zeros((16 * 16 * 4 + 4,), device="cuda")
    bad = dict(sc["uniforms"], zFar=0.05)
    with pytest.raises(wt.WtError):
        wtctx.shade(bad, t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"],
                    colour[:16 * 16 * 4].view(16, 16, 4))
    bad = dict(sc["uniforms"], sunShininess=-1.0)
    with pytest.raises(wt.WtError):
        wtctx.shade(bad, t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"],
                    colour[:16 * 16 * 4].view(16, 16, 4))
    # a colour buffer 4 bytes off 16-byte alignment
    p = wt.shade_params(sc["uniforms"])
    img = wt.Rgba8Image(t["reflections"].data_ptr(), 16, 16)
    tx = wt.ShadeTextures(t["depth"].data_ptr(), 16, 16, img, img, t["skybox"].data_ptr(), 4)
    st = wt.lib().wt_water_shade(wtctx.handle, ctypes.byref(p), ctypes.byref(tx), t["fragments"].data_ptr(), None,
                                 16, 16, colour.data_ptr() + 4, 0, None)
    assert st == wt.WT_ERR_INVALID_ARG
    st = wt.lib().wt_water_shade(wtctx.handle, ctypes.byref(p), ctypes.byref(tx), t["fragments"].data_ptr(), None,
                                 0, 16, colour.data_ptr(), 0, None)
    assert st == wt.WT_ERR_INVALID_ARG
    tx.skybox = None
    st = wt.lib().wt_water_shade(wtctx.handle, ctypes.byref(p), ctypes.byref(tx), t["fragments"].data_ptr(), None,
                                 16, 16, colour.data_ptr(), 0, None)
    assert st == wt.WT_ERR_INVALID_ARG


def test_shade_4k_frame(torch_cuda, wtctx):
    """The benchmark frame (3840 x 2160, 2048^2 skybox faces), every pixel against the
    oracle."""
    torch = torch_cuda
    say("shade 3840x2160 sky 2048 row bands")
    W, H = 3840, 2160
    sc = scenes.shading_scene(W, H, seed=11, sky_length=2048)
    colour = torch.full((H, W, 4), -1.0, device="cuda")
    got = run(torch, wtctx, sc, sc["coverage"], colour)
    sd = orc.ShadeData(sc["uniforms"], sc["depth"], sc["reflections"], sc["refractions"], sc["skybox"])
    ref = np.full((H, W, 4), -1.0)
    orc.water_shade(sd, sc["fragments"], sc["coverage"], nthreads=16, out=ref)
    stable, tol = comparison_mask(sc, sc["uniforms"], sc["depth"], sc["coverage"])
    assert stable.sum() > 0.999 * sc["coverage"].sum()
    check_frame(got, ref, stable, tol)
    cov = sc["coverage"].astype(bool)
    assert (got[~cov] == -1.0).all()
