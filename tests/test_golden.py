"""The oracle reproduces the committed golden fixtures (regression guard; the fixtures
were written by tests/golden/make_golden.py), and the GPU kernels match them at the
parity tolerance."""
import importlib.util
import json
import os

import numpy as np
import pytest

import wavetool_amd as wt
from helpers import NRM_TOL, POS_TOL, compare, near_texel_edge

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("name", ["default_scene", "w4_t10_n17", "bilinear", "adjoint_small"])
def test_oracle_reproduces_fixture(name):
    arrays, _ = maker().MAKERS[name]()
    ref = load(name)
    assert sorted(arrays) == sorted(ref)
    for k, v in arrays.items():
        np.testing.assert_allclose(v, ref[k], rtol=0, atol=1e-12, err_msg=k)


def test_fixture_header():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        h = json.load(f)
    assert h["default_scene"]["N"] == 17 and h["adjoint_small"]["W"] == 3


@pytest.mark.gpu
def test_gpu_matches_default_scene_fixture(torch_cuda, wtctx):
    """The reference's default scene (8-bit mirrored heightmap, reference scales) on the
    GPU against the fixture: positions within 1e-4 m, normals within 1e-4 rad."""
    torch = torch_cuda
    g = load("default_scene")
    N = 17
    dev = torch.device("cuda")
    waves = torch.from_numpy(g["waves"]).to(dev)
    tex = torch.from_numpy(g["texels"]).to(dev)
    cam_pos = (0.0, 4.0, 70.0)
    for t in (0.0, 10.0, 3600.0):
        p = wt.water_params(g["corners"], cam_pos, N, t, 0.02, 1.0, 0.0)
        pos = torch.empty(N * N * 3, device=dev)
        nrm = torch.empty_like(pos)
        tan = torch.empty_like(pos)
        bit = torch.empty_like(pos)
        wtctx.displace(p, waves, pos, nrm, tan, bit, heightmap=tex, hm_format=wt.WT_TEXEL_R8_UNORM,
                       hm_wrap=wt.WT_WRAP_MIRRORED_REPEAT)
        torch.cuda.synchronize()
        out = {k: v.cpu().numpy().reshape(-1, 3) for k, v in (("pos", pos), ("nrm", nrm), ("tan", tan), ("bit", bit))}
        ref = {k: g["t%g_%s" % (t, k)] for k in ("pos", "nrm", "tan", "bit")}
        edge = near_texel_edge(ref["pos"], 0.02, (0.0, 0.0), t, 16, 16)
        res = compare(out, ref, cam_pos, edge_mask=edge)
        print(t, res, flush=True)
        assert res["pos_max"] <= POS_TOL and res["nrm_max"] <= NRM_TOL
        assert res["tan_max"] <= NRM_TOL and res["bit_max"] <= NRM_TOL
        assert res["n_excluded"] + res["n_flip"] <= N * N // 50


@pytest.mark.gpu
def test_gpu_matches_adjoint_fixture(torch_cuda, wtctx):
    torch = torch_cuda
    g = load("adjoint_small")
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        m = json.load(f)["adjoint_small"]
    dev = torch.device("cuda")
    tgt = torch.from_numpy(g["target"]).to(dev)
    prm = torch.from_numpy(g["params"].astype(np.float32)).to(dev)
    fp = wt.fit_problem(m["n"], m["S"], m["W"], m["spacing"], m["x0"], m["z0"], m["t0"], m["dt"], tgt, prm)
    fit = wt.Fit(wtctx, [fp])
    grad = torch.zeros(m["W"] * 5, device=dev)
    loss = torch.zeros(1, device=dev)
    fit.loss_grad(grad, loss)
    torch.cuda.synchronize()
    ref = g["grad"]
    # the fixture's params are FP64; the device holds them in FP32
    assert abs(loss.item() - g["loss"][0]) <= 1e-5 * g["loss"][0]
    assert np.abs(grad.cpu().numpy().reshape(-1, 5) - ref).max() <= 2e-3 * np.abs(ref).max()
    fit.close()
