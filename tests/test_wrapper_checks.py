"""The Python wrappers refuse arrays the kernels would index out of bounds, before any
call reaches the C ABI (CPU tensors stand in for device ones: the checks run first)."""
import os

import numpy as np
import pytest
import torch

import wavetool_amd as wt
from wavetool_amd import scenes


def host_context():
    """A Context object with no device handle: enough to reach the argument checks."""
    ctx = wt.Context.__new__(wt.Context)
    ctx.handle = None
    return ctx


def forward_args(N=8, W=4):
    cam = scenes.camera_tuple()
    params = wt.water_params(wt.projector_corners(cam, 2.0), cam[4:7], N, 0.0, 0.02, 0.05)
    waves = torch.from_numpy(scenes.gen_waves(1, W))
    pos = torch.zeros(N * N * 3)
    return params, waves, pos, pos.clone()


def test_displace_rejects_undersized_outputs():
    ctx = host_context()
    params, waves, pos, nrm = forward_args()
    with pytest.raises(ValueError, match="pos"):
        ctx.displace(params, waves, pos[:-1], nrm)
    with pytest.raises(ValueError, match="nrm"):
        ctx.displace(params, waves, pos, nrm[:10])
    with pytest.raises(ValueError, match="tan"):
        ctx.displace(params, waves, pos, nrm, tan=torch.zeros(5))
    with pytest.raises(ValueError, match="bit"):
        ctx.displace(params, waves, pos, nrm, tan=pos.clone(), bit=torch.zeros(8 * 8 * 3, dtype=torch.float64))


def test_displace_rejects_wrong_dtype_and_layout():
    ctx = host_context()
    params, waves, pos, nrm = forward_args()
    with pytest.raises(ValueError, match="pos"):
        ctx.displace(params, waves, pos.double(), nrm)
    with pytest.raises(ValueError, match="pos"):  # a strided view
        ctx.displace(params, waves, torch.zeros(8 * 8 * 3 * 2)[::2], nrm)
    with pytest.raises(ValueError, match="waves"):
        ctx.displace(params, waves.double(), pos, nrm)
    with pytest.raises(ValueError, match="waves"):
        ctx.displace(params, waves[:, :6].contiguous(), pos, nrm)


def test_displace_rejects_more_waves_than_the_array_holds():
    ctx = host_context()
    params, waves, pos, nrm = forward_args(W=4)
    with pytest.raises(ValueError, match="waves"):
        ctx.displace(params, waves, pos, nrm, wave_count=5)


def test_heightmap_dtype_must_match_the_texel_format():
    ctx = host_context()
    params, waves, pos, nrm = forward_args()
    r8 = torch.zeros((16, 16), dtype=torch.uint8)
    r32 = torch.zeros((16, 16), dtype=torch.float32)
    with pytest.raises(ValueError, match="heightmap"):  # bytes read as R32F: 4x past the allocation
        ctx.displace(params, waves, pos, nrm, heightmap=r8, hm_format=wt.WT_TEXEL_R32F)
    with pytest.raises(ValueError, match="heightmap"):
        ctx.displace(params, waves, pos, nrm, heightmap=r32, hm_format=wt.WT_TEXEL_R8_UNORM)
    with pytest.raises(ValueError, match="heightmap"):
        ctx.displace(params, waves, pos, nrm, heightmap=r32.t(), hm_format=wt.WT_TEXEL_R32F)
    with pytest.raises(ValueError, match="heightmap"):
        ctx.displace(params, waves, pos, nrm, heightmap=r32.reshape(-1))


def test_varyings_check_sizes_against_the_vertex_count():
    ctx = host_context()
    params, waves, pos, nrm = forward_args()
    eye = np.eye(4)
    V = 64
    with pytest.raises(ValueError, match="pos"):
        ctx.varyings(params, eye, eye, pos, nrm, V + 1, view=torch.zeros((V + 1) * 3))
    with pytest.raises(ValueError, match="clip"):
        ctx.varyings(params, eye, eye, pos, nrm, V, clip=torch.zeros(V * 4 - 1))
    with pytest.raises(ValueError, match="ndc0"):
        ctx.varyings(params, eye, eye, pos, nrm, V, ndc0=torch.zeros(V))
    with pytest.raises(ValueError, match="sample"):
        ctx.varyings(params, eye, eye, pos, nrm, V, sample=torch.zeros(V, dtype=torch.float16),
                     heightmap=torch.zeros((4, 4)))
    with pytest.raises(ValueError, match="heightmap"):
        ctx.varyings(params, eye, eye, pos, nrm, V, sample=torch.zeros(V),
                     heightmap=torch.zeros((4, 4), dtype=torch.uint8))


def test_batch_displace_checks_against_the_batch_size():
    b = wt.Batch.__new__(wt.Batch)
    b.ctx, b.handle, b.count, b.vertices = host_context(), None, 2, 100
    waves = torch.from_numpy(scenes.gen_waves(1, 4))
    ok = torch.zeros(300)
    with pytest.raises(ValueError, match="pos"):
        b.displace(waves, torch.zeros(299), ok)
    with pytest.raises(ValueError, match="nrm"):
        b.displace(waves, ok, torch.zeros(300, dtype=torch.float64))
    with pytest.raises(ValueError, match="waves"):
        b.displace(waves.reshape(-1), ok, ok.clone())
    with pytest.raises(ValueError, match="heightmap"):
        b.displace(waves, ok, ok.clone(), heightmap=torch.zeros((4, 4), dtype=torch.uint8))


def test_abi_limits_in_the_python_mirror():
    assert wt._abi.WT_MAX_WAVES == 1024
    assert wt._abi.WT_MAX_VARYING_VERTICES == 2 * (2 ** 32 - 256)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "wt.h")).read()
    assert "#define WT_MAX_WAVES 1024" in text
    assert "#define WT_MAX_VARYING_VERTICES 8589934080ull" in text
