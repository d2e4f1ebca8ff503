"""GPU tests of the device projector (wt_projector_corners_batched, wt_batch_set_cameras):
the reference's per-frame projector (src/render-engine.cpp:797-1042) run on the device,
one thread per camera, against the library's host projector and the oracle's; and a
batch re-aimed on the device against a batch built from host corners."""
import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from wavetool_amd import scenes

pytestmark = pytest.mark.gpu

OK, NO_HIT, INVALID = wt.WT_OK, wt.WT_ERR_NO_INTERSECTION, wt.WT_ERR_INVALID_ARG


def say(msg):
    print("[case] " + msg, flush=True)


def cam(**kw):
    return scenes.camera_tuple(**kw)


def camera_set():
    """Config-2 cameras, random cameras over every branch, the named special cases and
    cameras the host projector rejects, each with its slab half-height."""
    cams = [s["camera"] for s in scenes.batched_patches(seed=2, count=512)]
    amps = [2.0] * len(cams)
    rng = np.random.default_rng(21)
    for _ in range(256):
        cams.append(cam(height=rng.uniform(-20, 60), pitch_deg=rng.uniform(-89, 89), yaw_deg=rng.uniform(-720, 720),
                        fov_deg=rng.uniform(20, 100), aspect=rng.uniform(0.5, 2.5), far=rng.uniform(50, 2000)))
        amps.append(float(rng.choice([0.0, 0.5, 2.0, 5.0])))
    special = [
        (cam(fov_deg=72.0, aspect=1.0, far=100.0, height=4.0, pitch_deg=0.0, z=70.0), 1.26),  # reference default
        (cam(), 2.0),  # the plain-grid scenes
        (cam(height=-3.0, pitch_deg=10.0), 2.0),  # underwater, looking up
        (cam(height=0.5, pitch_deg=-80.0), 2.0),  # inside the slab, looking down
        (cam(height=10.0, pitch_deg=0.0), 2.0),  # horizon: forward y clamped to -0.001
        (cam(height=10.0, pitch_deg=-120.0, yaw_deg=725.0), 2.0),  # pitch clamped, yaw wrapped
        (cam(height=100.0, pitch_deg=80.0, far=50.0), 2.0),  # misses the slab
        (cam(fov_deg=0.0), 2.0),  # rejected: fov
        (cam(near=10.0, far=5.0), 2.0),  # rejected: far < near
        (cam(), -1.0),  # rejected: negative amplitude
        (cam(height=float("nan")), 2.0),  # rejected: NaN position
    ]
    for c, a in special:
        cams.append(c)
        amps.append(a)
    return cams, amps


def host_corners(c, a):
    try:
        return OK, wt.projector_corners(c, a)
    except wt.WtError as e:
        return e.status, None


def test_device_projector_matches_host_and_oracle(torch_cuda, wtctx):
    torch = torch_cuda
    cams, amps = camera_set()
    n = len(cams)
    say("device projector: %d cameras" % n)
    d_cams = torch.from_numpy(wt.camera_rows(cams)).cuda()
    d_amps = torch.tensor(amps, dtype=torch.float32, device="cuda")
    corners = torch.zeros((n, 4, 4), dtype=torch.float64, device="cuda")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    wt.projector_corners_batched(wtctx, d_cams, d_amps, corners, status)
    torch.cuda.synchronize()
    got, st = corners.cpu().numpy(), status.cpu().numpy()
    worst_host = worst_oracle = 0.0
    counts = {}
    for i, (c, a) in enumerate(zip(cams, amps)):
        hs, ref = host_corners(c, a)
        counts[hs] = counts.get(hs, 0) + 1
        assert st[i] == hs, (i, c, a, st[i], hs)
        if hs != OK:
            assert np.isnan(got[i]).all()
            continue
        scale = max(1.0, np.abs(ref).max())
        worst_host = max(worst_host, np.abs(got[i] - ref).max() / scale)
        os_, oref, _ = orc.projector(c, a)
        assert os_ == OK
        worst_oracle = max(worst_oracle, np.abs(got[i] - oref).max() / max(1.0, np.abs(oref).max()))
        assert np.abs(got[i][:, 1]).max() <= 1e-9 * scale and (got[i][:, 3] == 1.0).all()
    say("status counts %s; max relative difference vs host %.3g, vs oracle %.3g" % (counts, worst_host, worst_oracle))
    assert counts.get(OK, 0) >= 600 and counts.get(NO_HIT, 0) >= 1 and counts.get(INVALID, 0) == 4
    # same steps in both builds with contraction off: only tan / sin / cos / acos rounding differs
    assert worst_host <= 1e-9
    assert worst_oracle <= 1e-8


def test_device_projector_rejects_short_arrays(torch_cuda, wtctx):
    torch = torch_cuda
    d_cams = torch.from_numpy(wt.camera_rows([cam(), cam()])).cuda()
    d_amps = torch.tensor([2.0], dtype=torch.float32, device="cuda")  # one short
    corners = torch.zeros((2, 4, 4), dtype=torch.float64, device="cuda")
    status = torch.zeros(2, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        wt.projector_corners_batched(wtctx, d_cams, d_amps, corners, status)
    # host memory is refused by the ABI itself
    with pytest.raises(wt.WtError) as e:
        wt.projector_corners_batched(wtctx, d_cams.cpu(), d_amps.repeat(2).cpu(), corners, status)
    assert e.value.status == INVALID


def small_specs(seed, count):
    specs = scenes.batched_patches(seed=seed, count=count)
    for s in specs:
        s["N"] = max(2, s["N"] // 8)
    return specs


def moved(specs, seed):
    """The same patches seen from other cameras (new height, pitch and yaw)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in specs:
        c = cam(height=rng.uniform(2.0, 50.0), pitch_deg=rng.uniform(-35.0, -5.0), yaw_deg=rng.uniform(0.0, 360.0),
                x=rng.uniform(-100, 100), z=rng.uniform(-100, 100))
        out.append(dict(s, camera=c))
    return out


def run(torch, ctx, batch, w, hm, V, fill=123.0, time_offset=0.0):
    pos = torch.full((V * 3,), fill, device="cuda")
    nrm = torch.full((V * 3,), fill, device="cuda")
    batch.displace(w, pos, nrm, heightmap=hm, time_offset=time_offset)
    torch.cuda.synchronize()
    return pos.cpu().numpy(), nrm.cpu().numpy()


def test_batch_set_cameras_matches_host_corners(torch_cuda, wtctx):
    """A batch built from cameras A and re-aimed on the device with cameras B gives the
    outputs of a batch built from B's host corners; a patch whose camera misses the
    slab keeps its old outputs and comes back when it gets corners again."""
    torch = torch_cuda
    specs_a = small_specs(9, 16)
    specs_b = moved(specs_a, 3)
    miss = 5
    say("set_cameras: %d patches, sides %s, patch %d misses the slab" % (len(specs_a), [s["N"] for s in specs_a],
                                                                         miss))
    patches_a, waves, V = wt.batch_inputs(specs_a)
    patches_b, _, _ = wt.batch_inputs(specs_b)
    w = torch.from_numpy(waves).cuda()
    hm = torch.from_numpy(scenes.value_noise(3, 128)).cuda()
    ref_batch = wt.Batch(wtctx, patches_b)
    ref_pos, ref_nrm = run(torch, wtctx, ref_batch, w, hm, V)
    ref_batch.close()

    batch = wt.Batch(wtctx, patches_a)
    a_pos, _ = run(torch, wtctx, batch, w, hm, V)
    cams_b = [s["camera"] for s in specs_b]
    cams_miss = list(cams_b)
    cams_miss[miss] = cam(height=100.0, pitch_deg=80.0, far=50.0)
    amps = torch.full((len(specs_a),), 2.0, dtype=torch.float32, device="cuda")
    status = torch.full((len(specs_a),), -1, dtype=torch.int32, device="cuda")
    batch.set_cameras(torch.from_numpy(wt.camera_rows(cams_miss)).cuda(), amps, status)
    pos, nrm = run(torch, wtctx, batch, w, hm, V)
    st = status.cpu().numpy()
    assert st[miss] == NO_HIT and (np.delete(st, miss) == OK).all()
    lo, hi = 3 * patches_a[miss].vertex_offset, 3 * (patches_a[miss].vertex_offset + specs_a[miss]["N"] ** 2)
    assert (pos[lo:hi] == 123.0).all() and (nrm[lo:hi] == 123.0).all()  # skipped: left untouched
    keep = np.ones(pos.size, bool)
    keep[lo:hi] = False
    dp, dn = np.abs(pos - ref_pos)[keep].max(), np.abs(nrm - ref_nrm)[keep].max()
    say("re-aimed vs host-built batch: pos %.3g m, nrm %.3g" % (dp, dn))
    assert dp <= 1e-5 and dn <= 1e-5
    assert np.abs(a_pos - ref_pos).max() > 1.0  # the cameras really moved the grids

    batch.set_cameras(torch.from_numpy(wt.camera_rows(cams_b)).cuda(), amps, status)
    pos, nrm = run(torch, wtctx, batch, w, hm, V)
    assert (status.cpu().numpy() == OK).all()
    assert np.abs(pos - ref_pos).max() <= 1e-5 and np.abs(nrm - ref_nrm).max() <= 1e-5
    batch.close()


def test_set_cameras_in_graph(torch_cuda, wtctx):
    """A captured frame loop that re-aims every patch each frame replays like the eager
    sequence, on every replay."""
    torch = torch_cuda
    specs = small_specs(12, 10)
    frames = 3
    cams = [[s["camera"] for s in moved(specs, 40 + f)] for f in range(frames)]
    patches, waves, V = wt.batch_inputs(specs)
    w = torch.from_numpy(waves).cuda()
    hm = torch.from_numpy(scenes.value_noise(4, 64)).cuda()
    d_cams = [torch.from_numpy(wt.camera_rows(c)).cuda() for c in cams]
    amps = torch.full((len(specs),), 2.0, dtype=torch.float32, device="cuda")
    outs = [(torch.zeros(V * 3, device="cuda"), torch.zeros(V * 3, device="cuda")) for _ in range(frames)]
    batch = wt.Batch(wtctx, patches)
    stream = torch.cuda.Stream()

    def frame_loop():
        for f in range(frames):
            batch.set_cameras(d_cams[f], amps, stream=stream)
            batch.displace(w, outs[f][0], outs[f][1], heightmap=hm, time_offset=0.1 * f, stream=stream)

    say("graph: %d frames x %d patches" % (frames, len(specs)))
    torch.cuda.synchronize()  # inputs were written on the default stream
    frame_loop()
    torch.cuda.synchronize()
    eager = [(p.cpu().numpy(), n.cpu().numpy()) for p, n in outs]
    assert all(np.isfinite(p).all() for p, _ in eager)
    torch.cuda.synchronize()
    wtctx.graph_begin(stream)
    frame_loop()
    graph = wtctx.graph_end(stream)
    for rep in range(3):
        for p, n in outs:
            p.zero_()
            n.zero_()
        torch.cuda.synchronize()
        graph.launch(stream)
        stream.synchronize()
        for f, (p, n) in enumerate(outs):
            assert np.array_equal(p.cpu().numpy(), eager[f][0]) and np.array_equal(n.cpu().numpy(), eager[f][1]), (rep, f)
    graph.close()
    batch.close()
