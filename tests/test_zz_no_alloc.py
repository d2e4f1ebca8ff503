"""wt_fit_loss_grad_static (the fit gradient in a fixed device array, no allocation)
against wt_fit_loss_grad on the same problems, through the C ABI.

Both run the same adjoint and reduction kernels over the same block layout, so the
gradients and losses must be bit-identical. Two calls in a row, the second on a smaller
problem, check that nothing of the first call's contents leaks into the second; a
problem beyond the fixed array is refused before any launch."""
import ctypes

import numpy as np
import pytest

import wavetool_amd as wt
from wavetool_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print("[case] " + msg, flush=True)


def problem(torch, seed, n, S, W):
    prob = scenes.fit_problem(seed, n=n, S=S, W_hidden=W, W_fit=W)
    dev = torch.device("cuda")
    hidden = torch.from_numpy(prob["hidden"].astype(np.float32)).to(dev)
    target = torch.empty((S, n * n, 3), device=dev)
    hp = wt.fit_problem(n, S, W, prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], None, hidden)
    params = torch.from_numpy(prob["init"].astype(np.float32)).to(dev)
    fp = wt.fit_problem(n, S, W, prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], target, params)
    return fp, hp, target, params


def test_static_scratch_gradient_equals_fit_object(torch_cuda, wtctx):
    torch = torch_cuda
    # the larger first, then a smaller one in a different snapshot class and wave count
    for k, (n, S, W) in enumerate([(512, 16, 32), (97, 6, 11)]):
        say("fixed-scratch fit gradient, call %d: n=%d S=%d W=%d" % (k + 1, n, S, W))
        fp, hp, target, params = problem(torch, 900 + k, n, S, W)
        wt.synthesize(wtctx, hp, 0.01, 900 + k, target)
        g_ref = torch.full((W * 5,), float("nan"), device="cuda")
        l_ref = torch.full((1,), float("nan"), device="cuda")
        fit = wt.Fit(wtctx, [fp])
        fit.loss_grad(g_ref, l_ref)
        g = torch.full((W * 5,), float("nan"), device="cuda")
        loss = torch.full((1,), float("nan"), device="cuda")
        wt.fit_loss_grad_static(wtctx, fp, g, loss)
        torch.cuda.synchronize()
        fit.close()
        assert torch.isfinite(g_ref).all() and torch.isfinite(l_ref).all()
        assert torch.equal(g, g_ref), (g - g_ref).abs().max().item()
        assert torch.equal(loss, l_ref)


def test_static_scratch_refuses_a_problem_beyond_its_size(torch_cuda, wtctx):
    torch = torch_cuda
    say("fixed-scratch fit gradient: size limit")
    dev = torch.device("cuda")
    W = 128
    params = torch.zeros((W, 5), device=dev)
    target = torch.zeros(3, device=dev)  # never read: the call is refused first
    fp = wt.fit_problem(2048, 16, W, 0.5, 0.0, 0.0, 0.0, 0.25, target, params)
    g = torch.zeros(W * 5, device=dev)
    st = wt.lib().wt_fit_loss_grad_static(wtctx.handle, ctypes.byref(fp), ctypes.c_void_p(g.data_ptr()), None, None)
    assert st == wt.WT_ERR_INVALID_ARG
    assert b"fixed scratch" in wt.lib().wt_last_error(wtctx.handle)
    assert wt.lib().wt_fit_loss_grad_static(None, ctypes.byref(fp), None, None, None) == wt.WT_ERR_INVALID_ARG
