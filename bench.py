"""Benchmark of the water-displacement stage on one MI355X.

Workload (one step): the plain forward pass over the 1024 x 1024 projected grid with
64 Gerstner waves and the 512 x 512 R32F value-noise heightmap, 240 frames at
t = 10 + k/60 s, captured as one HIP graph of 240 kernel launches that re-fill the
same position / normal buffers (the per-frame vertex-buffer update of the
reference's render loop). Inputs are resident in HBM before timing starts.

Prints one JSON line: vertex-waves/s (V * W * frames / wall time), the roofline of
the forward kernel (algorithmic 26 FLOP per vertex-wave / average launch duration
from HIP events on the launch stream, against the 157.3 TFLOP/s FP32 peak), and the
CPU baseline (the oracle's FP32 instantiation over std::threads on a bounded frame
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "wave-tool_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Gerstner vertex-waves/s at 1 GPU; batched adjoint fits/s vs plain @1e-4 m"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_baseline(corners, cam_pos, waves, texels, cfg, max_frames=16, budget_s=2.5):
    """FP32 oracle instantiation, all host threads (up to 16), frames of the same
    workload until the time budget or max_frames is reached."""
    import oracle as orc
    threads = min(16, os.cpu_count() or 1)
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
    return dict(value=vw / t_total, unit="vertex-waves/s", cores=threads, kind="port",
                sample="%d frame(s) of the same workload (%dx%d grid, %d waves, heightmap), FP32 oracle, "
                       "%d threads, %.2f s wall" % (frames, N, N, W, threads, t_total))


def load_traffic():
    path = os.path.join(ROOT, "profiles", "forward_config1_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import wavetool_amd as wt
    from wavetool_amd import scenes

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    cfg = scenes.plain_config(1)
    N, W, F = cfg["N"], cfg["W"], cfg["frames"]
    cam = cfg["camera"]
    cam_pos = cam[4:7]
    corners = wt.projector_corners(cam, cfg["amp"])
    waves_np = scenes.gen_waves(cfg["seed"], W)
    texels = scenes.value_noise(cfg["heightmap"]["seed"], cfg["heightmap"]["size"])
    waves = torch.from_numpy(waves_np).to(dev)
    hm = torch.from_numpy(texels).to(dev)
    V = N * N
    pos = torch.empty(V * 3, dtype=torch.float32, device=dev)
    nrm = torch.empty_like(pos)
    ctx = wt.Context()
    shape = None
    if not args.no_autotune:
        shape, ms = ctx.autotune(N, W, True)
        log("autotuned launch shape %s (%.3f ms per frame)" % (shape, ms))
    shape = ctx.get_launch_config(N, W)
    params = [wt.water_params(corners, cam_pos, N, cfg["t0"] + k * cfg["dt"], cfg["sample_scale"],
                              cfg["displacement_scale"], 0.0, cfg["scroll"]) for k in range(F)]
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.graph_begin(stream)
    for k in range(F):
        ctx.displace(params[k], waves, pos, nrm, heightmap=hm, stream=stream)
    graph = ctx.graph_end(stream)
    for _ in range(args.warmup):
        graph.launch(stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        graph.launch(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    graph.close()
    assert torch.isfinite(pos).all() and torch.isfinite(nrm).all()

    vw_per_step = V * W * F
    ms_per_step = wall * 1e3 / args.steps
    value = vw_per_step * args.steps / wall
    launch_s = ev_ms * 1e-3 / (args.steps * F)
    flop_per_launch = wt.WT_FLOP_PER_VERTEX_WAVE * V * W
    achieved_tflops = flop_per_launch / launch_s / 1e12
    traffic = load_traffic()
    roofline = dict(bound="mfma", engine="valu", achieved=round(achieved_tflops, 3),
                    peak=PEAK_FP32_TFLOPS, unit="TFLOP/s", frac=round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
                    traffic=traffic, launch_us=round(launch_s * 1e6, 3),
                    hbm_write_GBs=round(wt.WT_BYTES_PER_VERTEX_PN * V / launch_s / 1e9, 1),
                    note="achieved = 26 FLOP x V x W per launch / mean launch time (HIP events on the launch "
                         "stream); peak = MI355X FP32 vector")
    out = dict(metric=METRIC, value=value, unit="vertex-waves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
               data="synthetic",
               config=dict(workload="config1: plain forward, 1024x1024 projected grid, 64 Gerstner waves, "
                                    "512x512 R32F value-noise heightmap, 240 frames per step (one HIP graph)",
                           grid=N, waves=W, frames_per_step=F, launch_shape=shape),
               roofline=roofline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(corners, cam_pos, waves_np, texels, cfg)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
