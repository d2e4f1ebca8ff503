"""GPU parity of the heightfield-fit adjoint, Adam and target synthesis.

Gradients are compared with the oracle's FP64 adjoint (itself pinned to central finite
differences in tests/test_oracle_fit.py) with a tolerance relative to the largest
gradient component of the problem: 2e-4 (FP32 accumulation over up to 10^6 terms).
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


def say(msg):
    print("[case] " + msg, flush=True)


def make_problem(torch, seed, n, S, W, W_hidden=None, target_np=None):
    prob = scenes.fit_problem(seed, n=n, S=S, W_hidden=W_hidden or W, W_fit=W)
    tgt = scenes.fit_target(prob) if target_np is None else target_np
    dev = torch.device("cuda")
    t_target = torch.from_numpy(np.ascontiguousarray(tgt, dtype=np.float32)).to(dev)
    t_params = torch.from_numpy(prob["init"].astype(np.float32)).to(dev)
    fp = wt.fit_problem(n, S, W, prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], t_target, t_params)
    return prob, tgt, t_target, t_params, fp


def oracle_grad(prob, tgt, params):
    fd = orc.FitData(prob["n"], prob["S"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], tgt)
    return orc.fit_loss_grad(fd, params, nthreads=NTHREADS)


@pytest.mark.parametrize("n,S,W", [(16, 2, 3), (37, 5, 8), (20, 8, 5), (19, 17, 4), (64, 16, 32), (130, 32, 7),
                                   (1, 1, 1)])
def test_gradient_matches_fp64_adjoint(torch_cuda, wtctx, n, S, W):
    torch = torch_cuda
    say("adjoint n=%d S=%d W=%d" % (n, S, W))
    prob, tgt, t_target, t_params, fp = make_problem(torch, 100 + n, n, S, W)
    fit = wt.Fit(wtctx, [fp])
    grad = torch.zeros(W * 5, device="cuda")
    loss = torch.zeros(1, device="cuda")
    fit.loss_grad(grad, loss)
    torch.cuda.synchronize()
    params_used = t_params.cpu().numpy().astype(np.float64)
    L_ref, g_ref = oracle_grad(prob, tgt, params_used)
    g = grad.cpu().numpy().reshape(W, 5)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max() / scale
    print("loss gpu %.9g ref %.9g ; grad rel err %.3e" % (loss.item(), L_ref, err), flush=True)
    assert abs(loss.item() - L_ref) <= 1e-5 * max(L_ref, 1e-12) + 1e-9
    assert err <= GRAD_RTOL
    fit.close()


def test_batched_matches_oracle_and_single(torch_cuda, wtctx):
    """One batched launch over uneven problems: every problem's loss and gradient against
    the FP64 adjoint, and bit-identical to that problem's own single-problem launch (same
    kernels, same fixed reduction order)."""
    torch = torch_cuda
    say("batched vs oracle and single")
    specs = [(23, 4, 8), (64, 9, 12), (5, 1, 3), (100, 16, 20), (33, 32, 64)]
    probs = [make_problem(torch, 300 + i, *s) for i, s in enumerate(specs)]
    total = sum(s[2] * 5 for s in specs)
    gb = torch.zeros(total, device="cuda")
    lb = torch.zeros(len(specs), device="cuda")
    fitb = wt.Fit(wtctx, [p[4] for p in probs])
    fitb.loss_grad(gb, lb)
    off = 0
    torch.cuda.synchronize()
    for i, (p, s) in enumerate(zip(probs, specs)):
        prob, tgt, t_target, t_params, fp = p
        L_ref, g_ref = oracle_grad(prob, tgt, t_params.cpu().numpy().astype(np.float64))
        g = gb[off:off + s[2] * 5].cpu().numpy().reshape(-1, 5)
        assert np.abs(g - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max(), i
        assert abs(lb[i].item() - L_ref) <= 1e-5 * L_ref, i
        g1 = torch.zeros(s[2] * 5, device="cuda")
        l1 = torch.zeros(1, device="cuda")
        f1 = wt.Fit(wtctx, [fp])
        f1.loss_grad(g1, l1)
        torch.cuda.synchronize()
        assert torch.equal(g1, gb[off:off + s[2] * 5]), i
        assert torch.equal(l1[0], lb[i]), i
        off += s[2] * 5
        f1.close()
    fitb.close()


def test_synthesize_matches_fp64_model(torch_cuda, wtctx):
    """Noise-free target from the device generator equals the FP64 model within FP32
    rounding; with noise, the residual has the requested standard deviation."""
    torch = torch_cuda
    say("synthesize")
    n, S, W = 48, 6, 16
    prob = scenes.fit_problem(7, n=n, S=S, W_hidden=W, W_fit=W, noise=0.0)
    ref = scenes.fit_target(prob)
    dev = torch.device("cuda")
    hidden = torch.from_numpy(prob["hidden"].astype(np.float32)).to(dev)
    out = torch.zeros((S, n * n, 3), device=dev)
    fp = wt.fit_problem(n, S, W, prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], None, hidden)
    wt.synthesize(wtctx, fp, 0.0, 1, out)
    torch.cuda.synchronize()
    # the FP64 model evaluated with the FP32-rounded hidden parameters
    prob32 = dict(prob)
    prob32["hidden"] = prob["hidden"].astype(np.float32).astype(np.float64)
    ref32 = scenes.fit_target(prob32)
    assert np.abs(out.cpu().numpy() - ref32).max() <= 2e-5
    wt.synthesize(wtctx, fp, 0.01, 99, out)
    torch.cuda.synchronize()
    res = out.cpu().numpy() - ref32
    assert abs(res.std() - 0.01) < 0.0005 and abs(res.mean()) < 0.0005


def test_adam_tracks_fp64_reference(torch_cuda, wtctx):
    """20 Adam steps on the device against the same Adam recurrence on the host fed
    with the oracle's FP64 gradient."""
    torch = torch_cuda
    say("adam vs host reference")
    n, S, W = 24, 4, 4
    prob, tgt, t_target, t_params, fp = make_problem(torch, 41, n, S, W)
    fit = wt.Fit(wtctx, [fp])
    steps, lr, b1, b2, eps = 20, 1e-2, 0.9, 0.999, 1e-8
    p = prob["init"].astype(np.float32).astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, steps + 1):
        _, g = oracle_grad(prob, tgt, p)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    fit.iterate(steps, lr, b1, b2, eps)
    torch.cuda.synchronize()
    q = t_params.cpu().numpy().astype(np.float64)
    # Adam normalises each step to about lr, so FP32-vs-FP64 gradient sign flips on
    # near-zero components can move a parameter by up to lr per step; the loss at the
    # two endpoints must agree closely
    assert np.abs(q - p).max() <= 0.25 * lr * steps
    l_gpu, _ = oracle_grad(prob, tgt, q)
    l_ref, _ = oracle_grad(prob, tgt, p)
    assert abs(l_gpu - l_ref) <= 0.02 * l_ref
    fit.close()


def test_fit_reduces_loss_and_graph_replay(torch_cuda, wtctx):
    """A captured block of 25 Adam iterations replayed 4 times equals 100 eager
    iterations from the same start, and the fit drives the loss toward the noise floor."""
    torch = torch_cuda
    say("fit convergence + graph replay")
    n, S, W = 64, 8, 6
    prob, tgt, t_target, t_params, fp = make_problem(torch, 17, n, S, W)
    start = t_params.clone()
    fit = wt.Fit(wtctx, [fp])
    losses = torch.zeros(1, device="cuda")
    fit.iterate(1, losses=losses)
    torch.cuda.synchronize()
    first = losses.item()
    fit.iterate(99, losses=losses)
    torch.cuda.synchronize()
    eager = t_params.clone()
    last = losses.item()
    print("loss %.6g -> %.6g" % (first, last), flush=True)
    assert last < 0.5 * first
    # graph replay from the same start
    t_params.copy_(start)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    fit.reset(stream=s)
    s.synchronize()
    wtctx.graph_begin(s)
    fit.iterate(25, stream=s)
    g = wtctx.graph_end(s)
    for _ in range(4):
        g.launch(s)
        s.synchronize()
        _ = float(t_params.sum().item())  # eager work between replays
    torch.cuda.synchronize()
    assert torch.equal(t_params, eager)
    g.close()
    fit.close()


def test_autograd_wrapper(torch_cuda, wtctx):
    """torch.autograd.Function over the C-ABI adjoint (test harness only): backward
    returns the device gradient, and a torch optimiser step on it lowers the loss."""
    torch = torch_cuda
    say("autograd wrapper")

    class AdjointLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx_, params, fit, grad_buf, loss_buf):
            fit.loss_grad(grad_buf, loss_buf)
            ctx_.save_for_backward(grad_buf.clone())
            return loss_buf[0].clone()

        @staticmethod
        def backward(ctx_, gout):
            (g,) = ctx_.saved_tensors
            return g.view(-1, 5) * gout, None, None, None

    n, S, W = 32, 4, 5
    prob, tgt, t_target, t_params, fp = make_problem(torch, 23, n, S, W)
    fit = wt.Fit(wtctx, [fp])
    gbuf = torch.zeros(W * 5, device="cuda")
    lbuf = torch.zeros(1, device="cuda")
    t_params.requires_grad_(True)
    opt = torch.optim.Adam([t_params], lr=1e-2)
    l0 = None
    for it in range(30):
        opt.zero_grad()
        loss = AdjointLoss.apply(t_params, fit, gbuf, lbuf)
        loss.backward()
        if it == 0:
            l0 = loss.item()
            _, g_ref = oracle_grad(prob, tgt, t_params.detach().cpu().numpy().astype(np.float64))
            assert np.abs(t_params.grad.cpu().numpy() - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max()
        opt.step()
    assert loss.item() < l0
    fit.close()


def test_phase_plan_reports_recompute(torch_cuda):
    fp = wt.FitProblem()
    fp.grid_side, fp.snapshots, fp.wave_count = 512, 32, 64
    b, recompute = wt.phase_plan(fp, 1 << 30)
    assert b == 512 * 512 * 32 * 64 * 8 and recompute
