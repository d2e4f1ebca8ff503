"""Benchmark of the water-displacement stage on one MI355X.

Headline workload (one step): the plain forward pass over the 1024 x 1024 projected
grid with 64 Gerstner waves and the 512 x 512 R32F value-noise heightmap, 240 frames
at t = 10 + k/60 s (the per-frame vertex-buffer update of the reference's render loop).
By default two frames are in flight: frame k runs on stream k mod 2 and writes
vertex-buffer set k mod 2 (double-buffered vertex buffers), so one frame's epilogue and
stores overlap the next frame's wave sums; --frames-in-flight 1 runs them one after
another on one stream, and --loop graph replays the step as one captured HIP graph.
Inputs are resident in HBM before timing starts.

Prints one JSON line: vertex-waves/s (V * W * frames / wall time), the roofline of the
forward kernel (algorithmic 26 FLOP per vertex-wave over the launch period from HIP
events around the step's launches, against the 157.3 TFLOP/s FP32 peak, with each
launch's own serial duration beside it), the CPU baseline
(the oracle's FP32 instantiation over std::threads on a bounded frame sample of the
same workload), and `extras` for the other workloads: the batched forward over 512
uneven patches, the varyings pass, the fragment stage on a 3840 x 2160 frame, one
heightfield fit, and 256 uneven fits batched against the same fits launched one by one.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "wave-tool_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Gerstner vertex-waves/s at 1 GPU; batched adjoint fits/s vs plain @1e-4 m"
PEAK_FP32_TFLOPS = 157.3
ADAM = dict(lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the CPU baselines: every CPU this process may run on (its
    affinity set), capped by OMP_NUM_THREADS where the host sets it to the share of
    the CPUs that belongs to this job."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def host_cpu():
    """The host CPU model and its logical CPU count, for the CPU baselines' records."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return "%s, %d logical CPUs visible" % (model, os.cpu_count() or 1)


def cpu_baseline(corners, cam_pos, waves, texels, cfg, max_frames=240, budget_s=10.0):
    """FP32 oracle instantiation on cpu_threads() host threads, frames of the same
    workload (t = 10 + k/60) until about 10 s of wall time or the step's 240 frames."""
    import oracle as orc
    threads = cpu_threads()
    N, W = cfg["N"], cfg["W"]
    frames = 0
    t_total = 0.0
    while frames < max_frames and t_total < budget_s:
        t = cfg["t0"] + frames * cfg["dt"]
        sd = orc.SceneData(corners, cam_pos, N, waves, t, cfg["sample_scale"], cfg["displacement_scale"], 0.0,
                           cfg["scroll"], texels)
        t0 = time.perf_counter()
        orc.vertices(sd, precision="f32", nthreads=threads, want=("pos", "nrm"))
        t_total += time.perf_counter() - t0
        frames += 1
    vw = N * N * W * frames
    return dict(value=vw / t_total, unit="vertex-waves/s", cores=threads, kind="port", host=host_cpu(),
                sample="%d frame(s) of the same workload (%dx%d grid, %d waves, heightmap), FP32 oracle, "
                       "%d threads, %.2f s wall" % (frames, N, N, W, threads, t_total))


def load_traffic():
    """HBM bytes per forward launch from the committed FETCH_SIZE / WRITE_SIZE passes,
    each scaled by the same counter's reading for a plain copy kernel of known size.
    The write part (24 B per vertex) does not depend on the launch shape; the fetch part
    is the heightmap, once per XCD L2."""
    path = os.path.join(ROOT, "profiles", "water_forward_1024sq_w64_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("hbm_bytes_per_launch"), rec.get("launch_shape")


def timed(torch, stream, fn, reps=1):
    """Wall time and HIP-event time (on `stream`) of reps calls of fn."""
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, e0.elapsed_time(e1) * 1e-3


# --------------------------------------------------------------------- plain forward, 1024x1024 x 64 waves
def bench_plain(torch, wt, scenes, args):
    dev = torch.device("cuda:0")
    cfg = scenes.scene_1024_w64_heightmap()
    N, W, F = cfg["N"], cfg["W"], cfg["frames"]
    cam = cfg["camera"]
    cam_pos = cam[4:7]
    corners = wt.projector_corners(cam, cfg["amp"])
    waves_np = scenes.gen_waves(cfg["seed"], W)
    texels = scenes.value_noise(cfg["heightmap"]["seed"], cfg["heightmap"]["size"])
    waves = torch.from_numpy(waves_np).to(dev)
    hm = torch.from_numpy(texels).to(dev)
    V = N * N
    nfl = max(1, args.frames_in_flight)
    # one position / normal buffer set per frame in flight (double-buffered vertex
    # buffers: frame k writes set k mod nfl)
    bufs = [(torch.empty(V * 3, dtype=torch.float32, device=dev), torch.empty(V * 3, dtype=torch.float32, device=dev))
            for _ in range(nfl)]
    ctx = wt.Context()
    if args.launch_shape:
        v = [int(x) for x in args.launch_shape.split(",")]
        chunk = v[3] if len(v) > 3 else 64
        ctx.set_launch_config(v[0], v[1], v[2], chunk)
        log("launch shape forced to split %d, steps %d, lanes per row %d, wave chunk %d" % (v[0], v[1], v[2], chunk))
    elif not args.no_autotune:
        shape, ms = ctx.autotune(N, W, True, frames_in_flight=nfl)
        log("autotuned launch shape %s (%.3f ms per frame, %d in flight)" % (shape, ms, nfl))
    shape = ctx.get_launch_config(N, W)
    params = [wt.water_params(corners, cam_pos, N, cfg["t0"] + k * cfg["dt"], cfg["sample_scale"],
                              cfg["displacement_scale"], 0.0, cfg["scroll"]) for k in range(F)]
    streams = [torch.cuda.Stream() for _ in range(nfl)]
    main = streams[0]

    def frame_loop(n_in_flight):
        """The step's F frames, frame k on stream k mod n_in_flight writing buffer set k mod
        n_in_flight; the other streams fork from and join back to the first."""
        ss = streams[:n_in_flight]
        if n_in_flight > 1:
            fork = torch.cuda.Event()
            fork.record(main)
            for s in ss[1:]:
                s.wait_event(fork)
        for k in range(F):
            ctx.displace(params[k], waves, *bufs[k % n_in_flight], heightmap=hm, stream=ss[k % n_in_flight])
        for s in ss[1:]:
            join = torch.cuda.Event()
            join.record(s)
            main.wait_event(join)

    graph = None
    if args.loop == "graph":
        torch.cuda.synchronize()
        ctx.graph_begin(main)
        frame_loop(nfl)
        graph = ctx.graph_end(main)
        step = lambda: graph.launch(main)  # noqa: E731
    else:
        step = lambda: frame_loop(nfl)  # noqa: E731
    for _ in range(args.warmup):
        step()
    wall, ev = timed(torch, main, step, args.steps)
    if graph is not None:
        graph.close()
    for P, Nn in bufs:
        assert torch.isfinite(P).all() and torch.isfinite(Nn).all()
    # the same frames one at a time on one stream: each launch's own duration
    for _ in range(args.warmup):
        frame_loop(1)
    _, ev1 = timed(torch, main, lambda: frame_loop(1), args.steps)
    vw_per_step = V * W * F
    launch_s = ev / (args.steps * F)
    serial_s = ev1 / (args.steps * F)
    flop_per_launch = wt.WT_FLOP_PER_VERTEX_WAVE * V * W
    achieved = flop_per_launch / launch_s / 1e12
    traffic, traffic_shape = load_traffic()
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / PEAK_FP32_TFLOPS, 4), traffic=traffic, traffic_launch_shape=traffic_shape,
                    launch_us=round(launch_s * 1e6, 3), frames_in_flight=nfl,
                    serial_launch_us=round(serial_s * 1e6, 3),
                    serial_frac=round(flop_per_launch / serial_s / 1e12 / PEAK_FP32_TFLOPS, 4),
                    hbm_write_GBs=round(wt.WT_BYTES_PER_VERTEX_PN * V / launch_s / 1e9, 1),
                    note="achieved = 26 FLOP x V x W per launch / launch period (HIP events around the step's "
                         "launches: elapsed / frames; with %d frames in flight a launch overlaps the next one's, "
                         "so each launch's own duration, serial_launch_us, is measured beside it on one stream); "
                         "FP32 VALU kernel, peak = MI355X FP32 157.3 TFLOP/s" % nfl)
    out = dict(metric=METRIC, value=vw_per_step * args.steps / wall, unit="vertex-waves/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=wall * 1e3 / args.steps, higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload="plain forward: 1024x1024 projected grid, 64 Gerstner waves, "
                                    "512x512 R32F value-noise heightmap, 240 frames per step, %d frame(s) in flight "
                                    "(frame k on stream k mod %d writing vertex-buffer set k mod %d), %s launches"
                                    % (nfl, nfl, nfl, "graph-captured" if args.loop == "graph" else "eager"),
                           grid=N, waves=W, frames_per_step=F, frames_in_flight=nfl, loop=args.loop,
                           launch_shape=shape),
               roofline=roofline)
    ctx.close()
    return out, (corners, cam_pos, waves_np, texels, cfg)


# --------------------------------------------------------------------- batched forward, 512 patches
def bench_batched_forward(torch, wt, scenes, frames=120, one_by_one=True):
    dev = torch.device("cuda:0")
    specs = scenes.batched_patches(seed=2, count=512)
    patches, waves_np, V = wt.batch_inputs(specs)
    vw_frame = sum(s["N"] * s["N"] * s["W"] for s in specs)
    ctx = wt.Context()
    # segment shapes from the padded-vertex model (wt_batch_create's default), not the
    # per-bucket autotuner (wt_autotune_batch), which tunes on synthetic same-size batches
    batch = wt.Batch(ctx, patches)
    shapes = [batch.patch_launch_config(i) for i in range(len(patches))]
    waves = torch.from_numpy(waves_np).to(dev)
    hm = torch.from_numpy(scenes.value_noise(1, 512)).to(dev)
    pos = torch.empty(V * 3, device=dev)
    nrm = torch.empty_like(pos)
    # each frame re-aims every patch from its camera on the device (the reference runs its
    # projector every frame), then one launch per segment shape displaces all patches
    cams = torch.from_numpy(wt.camera_rows([s["camera"] for s in specs])).to(dev)
    amps = torch.full((len(specs),), 2.0, dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    for f in range(frames):
        batch.set_cameras(cams, amps, stream=stream)
        batch.displace(waves, pos, nrm, heightmap=hm, time_offset=f / 60.0, stream=stream)
    graph = ctx.graph_end(stream)
    graph.launch(stream)
    wall, _ = timed(torch, stream, lambda: graph.launch(stream))
    graph.close()
    # the device projector alone: one launch per frame over the 512 cameras
    ctx.graph_begin(stream)
    for f in range(frames):
        batch.set_cameras(cams, amps, stream=stream)
    jgraph = ctx.graph_end(stream)
    jgraph.launch(stream)
    jwall, _ = timed(torch, stream, lambda: jgraph.launch(stream))
    jgraph.close()
    if not one_by_one:
        batch.close()
        ctx.close()
        return dict(workload="batched forward: 512 patches, %d frames in one HIP graph (device projector and one "
                             "launch per segment shape per frame)" % frames,
                    value=vw_frame * frames / wall, unit="vertex-waves/s", ms_per_frame=wall * 1e3 / frames)
    # the same frame as one plain launch per patch (no graph), for the batching gain
    params = [wt.water_params(wt.np_corners(p), s["camera"][4:7], s["N"], s["time"], 0.02, 0.05, 0.0, (0.3, 0.0))
              for s, p in zip(specs, patches)]
    views = [(waves[p.wave_offset:p.wave_offset + p.wave_count], pos[3 * p.vertex_offset:],
              nrm[3 * p.vertex_offset:], p.wave_count) for p in patches]

    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    for prm, (w, P, Nn, W) in zip(params, views):
        ctx.displace(prm, w, P, Nn, heightmap=hm, stream=stream, wave_count=W)
    pgraph = ctx.graph_end(stream)
    pgraph.launch(stream)
    pwall, _ = timed(torch, stream, lambda: pgraph.launch(stream))
    pgraph.close()
    batch.close()
    ctx.close()
    batched = vw_frame * frames / wall
    return dict(workload="batched forward: 512 patches (side 32-1024 log-uniform, 4-128 waves, own camera and t in "
                         "[0, 3600) s), %d frames in one HIP graph; per frame the device projector re-aims "
                         "every patch, then one launch per segment shape displaces them" % frames,
                value=batched, unit="vertex-waves/s", vertex_waves_per_frame=vw_frame,
                vertices=V, ms_per_frame=wall * 1e3 / frames,
                patches_per_shape={"%dx%d" % (sh[2] * sh[1], 64 // sh[2]): shapes.count(sh) for sh in set(shapes)},
                device_projector=dict(us_per_frame=jwall * 1e6 / frames, cameras=len(specs),
                                      note="wt_batch_set_cameras alone, %d launches in one HIP graph" % frames),
                plain_per_patch=dict(value=vw_frame / pwall, unit="vertex-waves/s", ms_per_frame=pwall * 1e3,
                                     note="one wt_water_displace launch per patch (autotune-free default shapes), "
                                          "512 launches of one frame captured in a HIP graph"),
                speedup_vs_plain=round(batched / (vw_frame / pwall), 2))


# --------------------------------------------------------------------- varyings
def bench_varyings(torch, wt, scenes, launches=100):
    """The varyings pass (clip position, height-0 NDC, view vector, yaw-space normal,
    heightmap sample) over the 1024 x 1024 grid: an HBM-bound pass, 24 B read and 52 B
    written per vertex, timed as a captured graph of `launches` launches."""
    dev = torch.device("cuda:0")
    cfg = scenes.scene_1024_w64_heightmap()
    N, W = cfg["N"], cfg["W"]
    V = N * N
    cam = cfg["camera"]
    ctx = wt.Context()
    corners = wt.projector_corners(cam, cfg["amp"])
    waves = torch.from_numpy(scenes.gen_waves(cfg["seed"], W)).to(dev)
    hm = torch.from_numpy(scenes.value_noise(cfg["heightmap"]["seed"], cfg["heightmap"]["size"])).to(dev)
    params = wt.water_params(corners, cam[4:7], N, cfg["t0"], cfg["sample_scale"], cfg["displacement_scale"], 0.0,
                             cfg["scroll"])
    pos = torch.empty(V * 3, device=dev)
    nrm = torch.empty_like(pos)
    ctx.displace(params, waves, pos, nrm, heightmap=hm)
    view, proj, _ = wt.camera_matrices(cam)
    view_yaw, _, _ = wt.camera_matrices(cam[:8] + (0.0,))
    outs = dict(clip=torch.empty(V * 4, device=dev), ndc0=torch.empty(V * 2, device=dev),
                view=torch.empty(V * 3, device=dev), nview=torch.empty(V * 3, device=dev),
                sample=torch.empty(V, device=dev))
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    for _ in range(launches):
        ctx.varyings(params, proj @ view, view_yaw, pos, nrm, V, heightmap=hm, stream=stream, **outs)
    graph = ctx.graph_end(stream)
    graph.launch(stream)
    _, ev = timed(torch, stream, lambda: graph.launch(stream))
    graph.close()
    ctx.close()
    t = ev / launches
    traffic = 76 * V
    return dict(workload="varyings of the 1024 x 1024 grid: clip position, height-0 NDC xy, view vector, "
                         "yaw-only view-space normal, heightmap sample",
                value=V / t, unit="vertices/s", launch_us=round(t * 1e6, 3),
                roofline=dict(bound="hbm", achieved=round(traffic / t / 1e9, 1), peak=8000.0, unit="GB/s",
                              frac=round(traffic / t / 8e12, 4), bytes_per_vertex=76))


def bench_shade(torch, wt, scenes, launches=50, cpu=True):
    """The fragment stage (water-grid.frag as a deferred pass) on a 3840 x 2160 frame of
    synthetic sea with full-resolution local reflection / refraction targets and 2048^2
    skybox faces (CUBEMAP_LENGTH of src/render-engine.h:191), timed as a captured graph of
    `launches` launches. HBM-bound: per covered pixel the 48 B fragment record, a 4 B
    depth texel, one 4 B texel's worth of each of the three filtered lookups and the
    16 B colour (80 B), plus the 1 B coverage byte of every pixel."""
    import oracle as orc
    dev = torch.device("cuda:0")
    W, H, L = 3840, 2160, 2048
    sc = scenes.shading_scene(W, H, seed=11, sky_length=L)
    ctx = wt.Context()
    t = {k: torch.from_numpy(sc[k]).to(dev) for k in ("depth", "reflections", "refractions", "skybox", "fragments",
                                                         "coverage")}
    colour = torch.empty((H, W, 4), device=dev)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    for _ in range(launches):
        ctx.shade(sc["uniforms"], t["depth"], t["reflections"], t["refractions"], t["skybox"], t["fragments"], colour,
                  coverage=t["coverage"], stream=stream)
    graph = ctx.graph_end(stream)
    graph.launch(stream)
    _, ev = timed(torch, stream, lambda: graph.launch(stream))
    graph.close()
    ctx.close()
    tl = ev / launches
    covered = int(sc["coverage"].sum())
    traffic = 80 * covered + W * H
    out = dict(workload="water fragment stage, 3840 x 2160 frame (%d covered pixels), local targets 3840 x 2160, "
                        "skybox 6 x 2048^2 RGBA8" % covered,
               value=covered / tl, unit="fragments/s", launch_us=round(tl * 1e6, 3),
               roofline=dict(bound="hbm", achieved=round(traffic / tl / 1e9, 1), peak=8000.0, unit="GB/s",
                             frac=round(traffic / tl / 8e12, 4), bytes_per_covered_pixel=80))
    if cpu:
        threads = cpu_threads()
        sd = orc.ShadeData(sc["uniforms"], sc["depth"], sc["reflections"], sc["refractions"], sc["skybox"])
        buf = np.empty((H, W, 4), dtype=np.float32)
        frames, tc = 0, 0.0
        while frames < 50 and tc < 3.0:
            t0 = time.perf_counter()
            orc.water_shade(sd, sc["fragments"], sc["coverage"], precision="f32", nthreads=threads, out=buf)
            tc += time.perf_counter() - t0
            frames += 1
        out["cpu_baseline"] = dict(value=covered * frames / tc, unit="fragments/s", cores=threads, kind="port", host=host_cpu(),
                                   sample="%d whole frame(s), FP32 oracle, %d threads, %.2f s wall" % (frames, threads,
                                                                                                      tc))
    return out


# --------------------------------------------------------------------- wave fits, one and batched
def make_fits(torch, wt, scenes, ctx, specs, sigma=0.01):
    """Device targets (synthesized from the hidden waves) and perturbed initial params."""
    dev = torch.device("cuda:0")
    problems, keep = [], []
    for sp in specs:
        prob = scenes.fit_problem(sp["seed"], n=sp["n"], S=sp["S"], W_hidden=sp["W"], W_fit=sp["W"])
        hidden = torch.from_numpy(prob["hidden"].astype("float32")).to(dev)
        target = torch.empty((sp["S"], sp["n"] * sp["n"], 3), device=dev)
        hp = wt.fit_problem(sp["n"], sp["S"], sp["W"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"],
                            None, hidden)
        wt.synthesize(ctx, hp, sigma, sp["seed"], target)
        params = torch.from_numpy(prob["init"].astype("float32")).to(dev)
        problems.append(wt.fit_problem(sp["n"], sp["S"], sp["W"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"],
                                       prob["dt"], target, params))
        keep.append((target, params, prob))
    torch.cuda.synchronize()
    return problems, keep


# loss of the hidden waves themselves: 1/2 of the mean squared noise over the three
# displacement components, 1.5 sigma^2 at sigma = 0.01 m
NOISE_FLOOR = 1.5 * 0.01 ** 2


def fit_losses(torch, fit, specs, stream):
    """Each problem's loss at its current parameters (one gradient evaluation)."""
    g = torch.empty(sum(s["W"] * 5 for s in specs), device="cuda")
    l = torch.empty(len(specs), device="cuda")
    fit.loss_grad(g, l, stream=stream)
    stream.synchronize()
    return l.cpu().numpy().astype(np.float64)


def bench_fit_single(torch, wt, scenes, iterations=500, cpu=True):
    ctx = wt.Context()
    spec = dict(n=512, S=16, W=32, seed=3)
    problems, keep = make_fits(torch, wt, scenes, ctx, [spec])
    fit = wt.Fit(ctx, problems)
    losses = torch.zeros(1, device="cuda")
    stream = torch.cuda.Stream()
    loss0 = fit_losses(torch, fit, [spec], stream)
    init = keep[0][1].clone()
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    fit.reset(stream=stream)
    fit.iterate(iterations, losses=losses, stream=stream, **ADAM)
    graph = ctx.graph_end(stream)
    graph.launch(stream)  # warm-up replay, then the timed one from the same start
    torch.cuda.synchronize()
    keep[0][1].copy_(init)
    wall, _ = timed(torch, stream, lambda: graph.launch(stream))
    graph.close()
    loss1 = fit_losses(torch, fit, [spec], stream)
    vws = spec["n"] ** 2 * spec["S"] * spec["W"]
    out = dict(workload="single fit: 512x512 grid at 0.5 m, 16 snapshots, 32 waves, %d Adam iterations "
                        "(one HIP graph)" % iterations,
               value=1.0 / wall, unit="fits/s", ms_per_iteration=wall * 1e3 / iterations,
               loss_initial=float(loss0[0]), loss_final=float(loss1[0]), noise_floor=NOISE_FLOOR,
               adjoint_algorithmic_TFLOPs=round(wt._abi.WT_ADJOINT_FLOP_PER_VERTEX_WAVE_SNAPSHOT * vws * iterations
                                                / wall / 1e12, 2))
    if cpu:
        import oracle as orc
        target_np = keep[0][0].cpu().numpy()
        prob = keep[0][2]
        fd = orc.FitData(prob["n"], prob["S"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"],
                         target_np)
        t0 = time.perf_counter()
        orc.fit_loss_grad(fd, prob["init"], nthreads=cpu_threads())
        it_s = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=1.0 / (it_s * iterations), unit="fits/s", cores=cpu_threads(), kind="port", host=host_cpu(),
                                   sample="one FP64 oracle adjoint iteration (%.2f s wall), extrapolated to %d"
                                          % (it_s, iterations))
    fit.close()
    ctx.close()
    return out


def bench_fit_batched(torch, wt, scenes, iterations=300, count=256, one_by_one=True):
    ctx = wt.Context()
    specs = scenes.fit_batch_specs(seed=4, count=count)
    problems, keep = make_fits(torch, wt, scenes, ctx, specs)
    init = [k[1].clone() for k in keep]
    stream = torch.cuda.Stream()

    def restart():
        torch.cuda.synchronize()
        for k, p0 in zip(keep, init):
            k[1].copy_(p0)
        torch.cuda.synchronize()

    # batched: one problem table, the iterations captured in one graph
    fit = wt.Fit(ctx, problems)
    loss0 = fit_losses(torch, fit, specs, stream)
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    fit.reset(stream=stream)
    fit.iterate(iterations, stream=stream, **ADAM)
    graph = ctx.graph_end(stream)
    graph.launch(stream)  # warm-up replay
    restart()
    wall_b, _ = timed(torch, stream, lambda: graph.launch(stream))
    graph.close()
    loss1 = fit_losses(torch, fit, specs, stream)
    batched_params = [k[1].clone() for k in keep]
    fit.close()
    if not one_by_one:
        ctx.close()
        return dict(workload="batched fits: %d fits, %d Adam iterations, one HIP graph" % (count, iterations),
                    value=count / wall_b, unit="fits/s", loss_final_median=float(np.median(loss1)))
    # one by one: the same problems, each its own single-problem fit, each fit's iterations
    # captured in its own graph (so the comparison is batching, not launch overhead),
    # the graphs replayed in sequence
    singles = [wt.Fit(ctx, [p]) for p in problems]
    graphs = []
    torch.cuda.synchronize()
    for f in singles:
        ctx.graph_begin(stream)
        f.reset(stream=stream)
        f.iterate(iterations, stream=stream, **ADAM)
        graphs.append(ctx.graph_end(stream))
    for g in graphs:  # warm-up replay
        g.launch(stream)
    restart()

    def run_plain():
        for g in graphs:
            g.launch(stream)
    wall_p, _ = timed(torch, stream, run_plain)
    same = all(torch.equal(k[1], b) for k, b in zip(keep, batched_params))
    for g in graphs:
        g.close()
    # the same one-by-one fits with eager launches (what a caller without graphs sees)
    restart()

    def run_eager():
        for f in singles:
            f.reset(stream=stream)
            f.iterate(iterations, stream=stream, **ADAM)
    wall_e, _ = timed(torch, stream, run_eager)
    for f in singles:
        f.close()
    vws = sum(s["n"] ** 2 * s["S"] * s["W"] for s in specs)
    ctx.close()
    near = loss1 <= 2.0 * NOISE_FLOOR
    return dict(workload="batched fits: %d fits (side 64-512 at 0.5 m, 8-64 waves, 4-32 snapshots), %d Adam iterations; "
                         "batched = one problem table, one HIP graph" % (count, iterations),
                value=count / wall_b, unit="fits/s",
                plain_one_by_one=dict(value=count / wall_p, unit="fits/s",
                                      note="each problem its own single-problem fit, its iterations captured in its "
                                           "own HIP graph, the graphs replayed in sequence"),
                plain_one_by_one_eager=dict(value=count / wall_e, unit="fits/s",
                                            note="the same, eager launches"),
                speedup_vs_plain=round(wall_p / wall_b, 2), speedup_vs_eager=round(wall_e / wall_b, 2),
                results_identical=bool(same),
                loss_initial_median=float(np.median(loss0)), loss_final_median=float(np.median(loss1)),
                loss_final_max=float(loss1.max()), noise_floor=NOISE_FLOOR,
                share_within_2x_noise_floor=round(float(near.mean()), 4),
                share_loss_reduced_10x=round(float((loss1 <= 0.1 * loss0).mean()), 4),
                vertex_wave_snapshots_per_iteration=vws)


def run_extra(name, torch, wt, scenes, args):
    """One of the secondary workloads, with a one-line log of its figure."""
    cpu = not args.no_cpu_baseline
    if name == "batched_forward":
        r = bench_batched_forward(torch, wt, scenes, frames=args.frames, one_by_one=not args.no_one_by_one)
        log("batched forward %.4g vertex-waves/s" % r["value"])
    elif name == "varyings":
        r = bench_varyings(torch, wt, scenes)
        log("varyings %.4g vertices/s, %.1f GB/s" % (r["value"], r["roofline"]["achieved"]))
    elif name == "shade":
        r = bench_shade(torch, wt, scenes, cpu=cpu)
        log("fragment stage %.4g fragments/s, %.1f GB/s" % (r["value"], r["roofline"]["achieved"]))
    elif name == "fit_single":
        r = bench_fit_single(torch, wt, scenes, iterations=args.fit_iterations or 500, cpu=cpu)
        log("single fit %.4g fits/s" % r["value"])
    else:
        r = bench_fit_batched(torch, wt, scenes, iterations=args.fit_iterations or 300,
                              one_by_one=not args.no_one_by_one)
        log("batched fits %.4g fits/s (%sx one by one)" % (r["value"], r.get("speedup_vs_plain")))
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 warm-up and 50 timed steps (about 0.35 s of frames): with 2 and 5 the same frames
    # read up to 4% slower on some devices, before the clocks had settled
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch-shape", default="",
                    help="SPLIT,STEPS,LANES[,CHUNK]: fixed forward launch shape, no autotune")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--frames-in-flight", type=int, default=2,
                    help="frames of the step whose launches may overlap (one stream and buffer set each)")
    ap.add_argument("--loop", choices=("eager", "graph"), default="eager",
                    help="the step's launches issued directly, or captured once as one HIP graph and replayed")
    ap.add_argument("--shade-only", action="store_true",
                    help="run only the fragment-stage workload and print its line (for profiling)")
    ap.add_argument("--only", choices=("batched_forward", "varyings", "shade", "fit_single", "fit_batched"),
                    default="", help="run only this secondary workload and print its line (for profiling)")
    ap.add_argument("--frames", type=int, default=120, help="frames of the batched-forward workload")
    ap.add_argument("--fit-iterations", type=int, default=0,
                    help="Adam iterations of the fit workloads (default: 500 single, 300 batched)")
    ap.add_argument("--no-one-by-one", action="store_true",
                    help="skip the one-by-one comparators of the batched workloads (for counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    import wavetool_amd as wt
    from wavetool_amd import scenes

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.shade_only:
        args.only = "shade"
    if args.only:
        print(json.dumps(run_extra(args.only, torch, wt, scenes, args)), flush=True)
        return
    out, inputs = bench_plain(torch, wt, scenes, args)
    log("plain forward %.4g vertex-waves/s, %.2f us per frame" % (out["value"], out["roofline"]["launch_us"]))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(*inputs)
    if not args.no_extras:
        extras = {}
        for name in ("batched_forward", "varyings", "shade", "fit_single", "fit_batched"):
            extras[name] = run_extra(name, torch, wt, scenes, args)
        out["extras"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
