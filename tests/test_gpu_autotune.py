"""GPU tests of the autotuners (wt_autotune_forward, wt_autotune_batch) and of their
table's save / load round trip."""
import pytest

import wavetool_amd as wt
from wavetool_amd import scenes

pytestmark = pytest.mark.gpu

# the batched kernels' segment shapes (kBatchShape): (wave_split, steps, lanes_per_row, wave_chunk)
BATCH_SHAPES = [(2, 8, 64, 64), (2, 8, 32, 64), (2, 8, 16, 64)]


def say(msg):
    print("[case] " + msg, flush=True)


def specs_of(N, W, count=3):
    specs = scenes.batched_patches(seed=8, count=count)
    return [dict(s, N=N, W=W) for s in specs]


def test_batch_autotune_and_table_round_trip(torch_cuda, tmp_path):
    ctx = wt.Context()
    say("autotune_batch N=64 W=16, autotune N=96 W=16")
    shape, ms = ctx.autotune_batch(64, 16)
    assert shape in (0, 1, 2) and ms > 0.0
    cfg, ms2 = ctx.autotune(96, 16)
    assert ms2 > 0.0
    # a batch in the tuned bucket (sides 33..64, 9..16 waves) takes the tuned shape
    patches, _, _ = wt.batch_inputs(specs_of(50, 12))
    b = wt.Batch(ctx, patches)
    assert all(b.patch_launch_config(i) == BATCH_SHAPES[shape] for i in range(len(patches)))
    b.close()
    path = tmp_path / "tune.txt"
    ctx.autotune_save(path)
    assert "\nb 6 4 %d\n" % shape in path.read_text()
    ctx.close()

    fresh = wt.Context()
    fresh.autotune_load(path)
    assert fresh.get_launch_config(96, 16) == cfg
    patches, _, _ = wt.batch_inputs(specs_of(64, 16))
    b = wt.Batch(fresh, patches)
    assert all(b.patch_launch_config(i) == BATCH_SHAPES[shape] for i in range(len(patches)))
    b.close()
    for bad in ("b 6 4 3\n", "b 6 x\n", "1 2 3\n"):
        p = tmp_path / "bad.txt"
        p.write_text(bad)
        with pytest.raises(wt.WtError):
            fresh.autotune_load(p)
    fresh.close()


def test_tuned_batch_matches_plain_launches(torch_cuda, wtctx):
    """Whatever shape the tuner picks, the batch still equals the plain kernel run on
    each patch with that shape, bit for bit."""
    torch = torch_cuda
    import numpy as np
    specs = specs_of(40, 10, count=4)
    say("tuned batch vs plain: sides 40, 10 waves")
    ctx = wt.Context()
    ctx.autotune_batch(40, 10)
    patches, waves_np, V = wt.batch_inputs(specs)
    batch = wt.Batch(ctx, patches)
    w = torch.from_numpy(waves_np).cuda()
    hm = torch.from_numpy(scenes.value_noise(3, 64)).cuda()
    pos = torch.empty(V * 3, device="cuda")
    nrm = torch.empty_like(pos)
    batch.displace(w, pos, nrm, heightmap=hm)
    torch.cuda.synchronize()
    for i, (s, p) in enumerate(zip(specs, patches)):
        ctx.set_launch_config(*batch.patch_launch_config(i))
        params = wt.water_params(np.ctypeslib.as_array(p.corners).copy(), s["camera"][4:7], s["N"], s["time"], 0.02,
                                 0.05, 0.0, (0.3, 0.0))
        P = torch.empty(s["N"] ** 2 * 3, device="cuda")
        Nn = torch.empty_like(P)
        ctx.displace(params, w[p.wave_offset:p.wave_offset + p.wave_count], P, Nn, heightmap=hm,
                     wave_count=p.wave_count)
        torch.cuda.synchronize()
        lo, hi = 3 * p.vertex_offset, 3 * (p.vertex_offset + s["N"] ** 2)
        assert torch.equal(P, pos[lo:hi]) and torch.equal(Nn, nrm[lo:hi]), i
    batch.close()
    ctx.close()


def test_forward_autotune_overlapped_and_chunks(torch_cuda):
    """The overlapped autotuner (two and three frames in flight) returns a supported shape
    and a per-frame time, stores it in the table, and a 32 / 128 wave chunk is accepted only
    on the one-row shapes (two wavefronts per segment, 64 lanes per row)."""
    ctx = wt.Context()
    say("autotune overlapped N=256 W=40")
    for nfl in (2, 3):
        cfg, ms = ctx.autotune(256, 40, True, frames_in_flight=nfl)
        assert ms > 0.0 and cfg["wave_chunk"] in (32, 64, 128)
        assert ctx.get_launch_config(256, 40) == cfg
        if cfg["wave_chunk"] != 64:
            assert cfg["wave_split"] == 2 and cfg["lanes_per_row"] == 64
    with pytest.raises(wt.WtError):
        ctx.autotune(256, 40, True, frames_in_flight=5)
    ctx.set_launch_config(2, 8, 64, 128)
    ctx.set_launch_config(2, 16, 64, 32)
    for bad in ((2, 8, 32, 128), (1, 8, 64, 32), (2, 8, 64, 96)):
        with pytest.raises(wt.WtError):
            ctx.set_launch_config(*bad)
    ctx.close()
