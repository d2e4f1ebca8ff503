"""GPU tests of the batched forward kernels (wt_water_displace_batched)."""
import os

import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from wavetool_amd import scenes
from helpers import POS_TOL, NRM_TOL, compare, near_texel_edge

pytestmark = pytest.mark.gpu
NTHREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print("[case] " + msg, flush=True)


def run_batch(torch, ctx, specs, texels, time_offset=0.0):
    dev = torch.device("cuda")
    patches, waves, V = wt.batch_inputs(specs)
    batch = wt.Batch(ctx, patches)
    assert batch.vertices == V
    w = torch.from_numpy(waves).to(dev)
    hm = torch.from_numpy(texels).to(dev)
    pos = torch.full((V * 3,), float("nan"), device=dev)
    nrm = torch.full((V * 3,), float("nan"), device=dev)
    batch.displace(w, pos, nrm, heightmap=hm, time_offset=time_offset)
    torch.cuda.synchronize()
    return batch, patches, waves, w, hm, pos, nrm


@pytest.mark.parametrize("count,seed", [(24, 5), (7, 6)])
def test_batched_equals_plain_per_patch(torch_cuda, wtctx, count, seed):
    """Every patch of a batch equals the plain kernel run on that patch alone with the
    launch shape the batch chose for it, bit for bit; so the batch inherits the plain
    kernel's parity."""
    torch = torch_cuda
    specs = scenes.batched_patches(seed=seed, count=count)
    if seed == 6:  # a batch of small grids only (sides 9-182: the two narrow shapes)
        for s in specs:
            s["N"] = max(2, s["N"] // 4)
    say("batched vs plain: %d patches, sides %s" % (count, [s["N"] for s in specs]))
    texels = scenes.value_noise(3, 128)
    batch, patches, waves, w, hm, pos, nrm = run_batch(torch, wtctx, specs, texels, time_offset=0.5)
    shapes = set()
    try:
        for i, (s, p) in enumerate(zip(specs, patches)):
            shape = batch.patch_launch_config(i)
            shapes.add(shape)
            wtctx.set_launch_config(*shape)
            N = s["N"]
            params = wt.water_params(np.ctypeslib.as_array(p.corners).copy(), s["camera"][4:7], N, s["time"] + 0.5, 0.02, 0.05, 0.0,
                                     (0.3, 0.0))
            P = torch.empty(N * N * 3, device="cuda")
            Nn = torch.empty_like(P)
            wtctx.displace(params, w[p.wave_offset:p.wave_offset + p.wave_count], P, Nn, heightmap=hm,
                           wave_count=p.wave_count)
            torch.cuda.synchronize()
            o = p.vertex_offset * 3
            assert torch.equal(P, pos[o:o + N * N * 3]), i
            assert torch.equal(Nn, nrm[o:o + N * N * 3]), i
    finally:
        wtctx.set_launch_config()
    print("shapes used:", sorted(shapes), flush=True)
    assert len(shapes) >= 2  # both batches mix segment shapes
    batch.close()


def test_batched_oracle_parity_long_times(torch_cuda, wtctx):
    """Patches at t up to 3600 s against the FP64 oracle (FP64 phase reduction)."""
    torch = torch_cuda
    specs = scenes.batched_patches(seed=9, count=6)
    for s in specs:
        s["N"] = min(s["N"], 200)
    say("batched oracle parity, times %s" % [round(s["time"], 1) for s in specs])
    texels = scenes.value_noise(4, 64)
    batch, patches, waves, w, hm, pos, nrm = run_batch(torch, wtctx, specs, texels)
    P = pos.cpu().numpy().reshape(-1, 3)
    Nn = nrm.cpu().numpy().reshape(-1, 3)
    for s, p in zip(specs, patches):
        N = s["N"]
        sd = orc.SceneData(np.ctypeslib.as_array(p.corners).copy(), s["camera"][4:7], N, waves[p.wave_offset:p.wave_offset + p.wave_count],
                           s["time"], 0.02, 0.05, 0.0, (0.3, 0.0), texels)
        ref = orc.vertices(sd, nthreads=NTHREADS, want=("pos", "nrm"))
        o = p.vertex_offset
        out = dict(pos=P[o:o + N * N], nrm=Nn[o:o + N * N])
        edge = near_texel_edge(ref["pos"], 0.02, (0.3, 0.0), s["time"], 64, 64)
        res = compare(out, ref, s["camera"][4:7], edge)
        print(N, s["W"], res, flush=True)
        assert res["pos_max"] <= POS_TOL and res["nrm_max"] <= NRM_TOL
    batch.close()


def test_batched_graph_replay(torch_cuda, wtctx):
    """Eight frames of the batch captured in one graph; replays equal eager frames."""
    torch = torch_cuda
    say("batched graph replay")
    specs = scenes.batched_patches(seed=11, count=12)
    texels = scenes.value_noise(3, 128)
    batch, patches, waves, w, hm, pos, nrm = run_batch(torch, wtctx, specs, texels)
    F = 8
    V = batch.vertices
    outp = torch.zeros((F, V * 3), device="cuda")
    outn = torch.zeros_like(outp)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    for f in range(F):
        batch.displace(w, outp[f], outn[f], heightmap=hm, time_offset=f / 60.0, stream=s)
    s.synchronize()
    ep, en = outp.clone(), outn.clone()
    wtctx.graph_begin(s)
    for f in range(F):
        batch.displace(w, outp[f], outn[f], heightmap=hm, time_offset=f / 60.0, stream=s)
    g = wtctx.graph_end(s)
    for rep in range(3):
        outp.zero_()
        outn.zero_()
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        assert torch.equal(outp, ep) and torch.equal(outn, en), rep
    g.close()
    batch.close()


def test_batch_rejects_bad_input(torch_cuda, wtctx):
    say("batch invalid input")
    specs = scenes.batched_patches(seed=2, count=3)
    patches, waves, V = wt.batch_inputs(specs)
    bad = list(patches)
    bad[1].gridLength = 1
    with pytest.raises(wt.WtError) as e:
        wt.Batch(wtctx, bad)
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    # a wave range that wraps 32 bits, and a vertex offset past the 2^48 limit
    patches, waves, V = wt.batch_inputs(specs)
    patches[2].wave_offset = 2 ** 32 - 2
    with pytest.raises(wt.WtError) as e:
        wt.Batch(wtctx, patches)
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    patches, waves, V = wt.batch_inputs(specs)
    patches[0].wave_offset = 2 ** 31 - 2
    with pytest.raises(wt.WtError) as e:
        wt.Batch(wtctx, patches)
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    patches, waves, V = wt.batch_inputs(specs)
    patches[1].vertex_offset = 2 ** 48 + 1
    with pytest.raises(wt.WtError) as e:
        wt.Batch(wtctx, patches)
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    patches, waves, V = wt.batch_inputs(specs)
    b = wt.Batch(wtctx, patches)
    short = torch_cuda.from_numpy(waves[:2]).cuda()
    pos = torch_cuda.zeros(V * 3, device="cuda")
    with pytest.raises(wt.WtError) as e:
        b.displace(short, pos, pos.clone())
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    b.close()


def moved_camera(c, f):
    """Camera c after f frames of a slow fly-by: yaw turns 0.05 deg, the eye moves 0.1 m
    along x and -0.05 m along z and rises 0.2% per frame (float32-representable)."""
    return scenes.camera_tuple(height=c[5] * (1.0 + 0.002 * f), pitch_deg=c[8], yaw_deg=c[7] + 0.05 * f,
                               fov_deg=c[0], aspect=c[1], near=c[2], far=c[3], x=c[4] + 0.1 * f, z=c[6] - 0.05 * f)


def test_config2_full_batch_reaimed_on_device(torch_cuda, wtctx):
    """The 512-patch batch of the benchmark (seed 2: sides 32-1024, 4-128 waves, t in
    [0, 3600) s) re-aimed by wt_batch_set_cameras and displaced at frames 0 and 119 of
    its 120-frame loop, the cameras moving between the two. 18 patches per frame -
    the largest grid, the most waves and 16 drawn at random - against the FP64 oracle,
    whose corners come from the oracle's own projector for the same cameras."""
    torch = torch_cuda
    dev = torch.device("cuda")
    specs = scenes.batched_patches(seed=2, count=512)
    patches, waves_np, V = wt.batch_inputs(specs)
    batch = wt.Batch(wtctx, patches)
    waves = torch.from_numpy(waves_np).to(dev)
    texels = scenes.value_noise(1, 512)
    hm = torch.from_numpy(texels).to(dev)
    pos = torch.full((V * 3,), float("nan"), device=dev)
    nrm = torch.full((V * 3,), float("nan"), device=dev)
    amps = torch.full((len(specs),), 2.0, dtype=torch.float32, device=dev)
    status = torch.full((len(specs),), -1, dtype=torch.int32, device=dev)
    rng = np.random.default_rng(22)
    big = max(range(len(specs)), key=lambda i: specs[i]["N"])
    many = max(range(len(specs)), key=lambda i: specs[i]["W"])
    for f in (0, 119):
        cams = [moved_camera(s["camera"], f) for s in specs]
        batch.set_cameras(torch.from_numpy(wt.camera_rows(cams)).to(dev), amps, status)
        batch.displace(waves, pos, nrm, heightmap=hm, time_offset=f / 60.0)
        torch.cuda.synchronize()
        st = status.cpu().numpy()
        assert (st == wt.WT_OK).all(), np.unique(st)
        pick = [big, many] + [int(i) for i in rng.choice(len(specs), size=16, replace=False)]
        say("config 2 frame %d: %d patches, checking %s" % (f, len(specs), pick))
        for i in pick:
            s, p = specs[i], patches[i]
            N = s["N"]
            ost, corners, _ = orc.projector(cams[i], 2.0)
            assert ost == 0
            t = s["time"] + f / 60.0
            sd = orc.SceneData(corners, cams[i][4:7], N, waves_np[p.wave_offset:p.wave_offset + p.wave_count], t,
                               0.02, 0.05, 0.0, (0.3, 0.0), texels)
            ref = orc.vertices(sd, nthreads=NTHREADS, want=("pos", "nrm"))
            o = p.vertex_offset * 3
            out = dict(pos=pos[o:o + N * N * 3].cpu().numpy().reshape(-1, 3),
                       nrm=nrm[o:o + N * N * 3].cpu().numpy().reshape(-1, 3))
            edge = near_texel_edge(ref["pos"], 0.02, (0.3, 0.0), t, 512, 512)
            res = compare(out, ref, cams[i][4:7], edge)
            print(f, i, N, s["W"], res, flush=True)
            assert res["n_near"] > 0
            assert res["pos_max"] <= POS_TOL and res["nrm_max"] <= NRM_TOL, (f, i)
            assert res["n_excluded"] <= 0.02 * res["n_near"] + 4
    batch.close()
