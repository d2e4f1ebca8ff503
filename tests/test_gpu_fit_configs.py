"""GPU parity of the heightfield fit at the benchmarked sizes.

- One 512 x 512 x 16 snapshot x 32 wave problem (the single-fit workload): loss and
  gradient against the oracle's FP64 adjoint, and 12 gradient components against FP64
  central differences of the oracle's loss at h = 1e-3 relative.
- The 256-problem batch (sides 64-512, 8-64 waves, 4-32 snapshots): every problem's
  gradient from the one batched launch at iteration 0, 36 of them (the largest by work,
  the most waves, the most snapshots and the rest drawn at random) against the FP64
  adjoint problem by problem.

Targets are synthesized on the device (hidden waves plus N(0, 0.01 m) noise) and copied
back, so the oracle sees the same data. Tolerance: 2e-4 of the problem's largest
gradient component (FP32 accumulation over up to 10^9 terms); the loss to 1e-5 relative.
"""
import os

import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from wavetool_amd import scenes

pytestmark = pytest.mark.gpu
NTHREADS = min(16, os.cpu_count() or 1)
GRAD_RTOL = 2e-4
LOSS_RTOL = 1e-5


def say(msg):
    print("[case] " + msg, flush=True)


def device_fits(torch, ctx, specs, sigma=0.01):
    """Device targets synthesized from each problem's hidden waves (noise keyed by its
    seed) and perturbed initial parameters, as the benchmark builds them."""
    dev = torch.device("cuda")
    problems, keep = [], []
    for sp in specs:
        prob = scenes.fit_problem(sp["seed"], n=sp["n"], S=sp["S"], W_hidden=sp["W"], W_fit=sp["W"])
        hidden = torch.from_numpy(prob["hidden"].astype(np.float32)).to(dev)
        target = torch.empty((sp["S"], sp["n"] * sp["n"], 3), device=dev)
        hp = wt.fit_problem(sp["n"], sp["S"], sp["W"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"],
                            None, hidden)
        wt.synthesize(ctx, hp, sigma, sp["seed"], target)
        params = torch.from_numpy(prob["init"].astype(np.float32)).to(dev)
        problems.append(wt.fit_problem(sp["n"], sp["S"], sp["W"], prob["spacing"], prob["x0"], prob["z0"],
                                       prob["t0"], prob["dt"], target, params))
        keep.append((target, params, prob))
    torch.cuda.synchronize()
    return problems, keep


def oracle_data(target, prob):
    return orc.FitData(prob["n"], prob["S"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"],
                       target.cpu().numpy())


def test_single_fit_512sq_s16_w32_gradient(torch_cuda, wtctx):
    torch = torch_cuda
    spec = dict(n=512, S=16, W=32, seed=3)
    say("single fit 512x512 x 16 snapshots x 32 waves: gradient vs FP64 adjoint and central differences")
    problems, keep = device_fits(torch, wtctx, [spec])
    target, params, prob = keep[0]
    fit = wt.Fit(wtctx, problems)
    grad = torch.zeros(spec["W"] * 5, device="cuda")
    loss = torch.zeros(1, device="cuda")
    fit.loss_grad(grad, loss)
    torch.cuda.synchronize()
    fit.close()
    p = params.cpu().numpy().astype(np.float64)
    fd = oracle_data(target, prob)
    L_ref, g_ref = orc.fit_loss_grad(fd, p, nthreads=NTHREADS)
    g = grad.cpu().numpy().astype(np.float64).reshape(-1, 5)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max() / scale
    print("loss gpu %.9g oracle %.9g; gradient error %.3e of max" % (loss.item(), L_ref, err), flush=True)
    assert abs(loss.item() - L_ref) <= LOSS_RTOL * L_ref
    assert err <= GRAD_RTOL
    # central differences of the FP64 loss, h = 1e-3 relative; the step's own truncation
    # error is estimated from the half step and allowed on top of the tolerance
    rng = np.random.default_rng(31)
    comps = [(int(i), int(k)) for i, k in zip(rng.integers(0, spec["W"], 10), rng.integers(0, 5, 10))]
    comps += [(int(np.argmax(np.abs(g_ref).max(axis=1))), int(np.argmax(np.abs(g_ref).max(axis=0)))),
              (int(np.argmin(p[:, 1])), 1)]  # the largest component and the shortest wave's wavelength

    def central(i, k, h):
        pp, pm = p.copy(), p.copy()
        pp[i, k] += h
        pm[i, k] -= h
        return (orc.fit_loss_grad(fd, pp, nthreads=NTHREADS, grad=False)[0] -
                orc.fit_loss_grad(fd, pm, nthreads=NTHREADS, grad=False)[0]) / (2.0 * h)

    for i, k in comps:
        h = 1e-3 * max(abs(p[i, k]), 1e-2)
        d1, d2 = central(i, k, h), central(i, k, 0.5 * h)
        trunc = abs(d1 - d2) * 4.0 / 3.0
        print("  dL/dp[%d,%d]: gpu %.8g  FD %.8g  (truncation ~%.2e)" % (i, k, g[i, k], d1, trunc), flush=True)
        assert abs(g[i, k] - d1) <= GRAD_RTOL * scale + trunc


def test_batched_256_fits_gradients_vs_oracle(torch_cuda, wtctx):
    torch = torch_cuda
    specs = scenes.fit_batch_specs(seed=4, count=256)
    say("256-fit batch: gradients of one batched launch vs the FP64 adjoint")
    problems, keep = device_fits(torch, wtctx, specs)
    total = sum(s["W"] * 5 for s in specs)
    grad = torch.zeros(total, device="cuda")
    losses = torch.zeros(len(specs), device="cuda")
    fit = wt.Fit(wtctx, problems)
    fit.loss_grad(grad, losses)
    torch.cuda.synchronize()
    fit.close()
    g_all = grad.cpu().numpy().astype(np.float64)
    l_all = losses.cpu().numpy().astype(np.float64)
    assert np.isfinite(g_all).all() and np.isfinite(l_all).all()
    offs = np.concatenate([[0], np.cumsum([s["W"] * 5 for s in specs])])
    work = [s["n"] ** 2 * s["S"] * s["W"] for s in specs]
    order = np.argsort(work)[::-1]
    pick = list(order[:8])
    pick += [int(np.argmax([s["W"] for s in specs])), int(np.argmax([s["S"] for s in specs])),
             int(np.argmin(work))]
    rng = np.random.default_rng(44)
    rest = [i for i in range(len(specs)) if i not in pick]
    pick += [int(i) for i in rng.choice(rest, size=36 - len(set(pick)), replace=False)]
    pick = sorted(set(int(i) for i in pick))
    assert len(pick) >= 32
    worst = 0.0
    for i in pick:
        s = specs[i]
        target, params, prob = keep[i]
        L_ref, g_ref = orc.fit_loss_grad(oracle_data(target, prob), params.cpu().numpy().astype(np.float64),
                                         nthreads=NTHREADS)
        g = g_all[offs[i]:offs[i + 1]].reshape(-1, 5)
        err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
        worst = max(worst, err)
        print("  problem %3d n=%3d S=%2d W=%2d: loss %.7g / %.7g, gradient error %.2e" % (
            i, s["n"], s["S"], s["W"], l_all[i], L_ref, err), flush=True)
        assert abs(l_all[i] - L_ref) <= LOSS_RTOL * L_ref, i
        assert err <= GRAD_RTOL, i
    print("worst gradient error over %d problems: %.2e" % (len(pick), worst), flush=True)
