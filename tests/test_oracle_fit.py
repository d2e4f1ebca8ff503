"""Pins the oracle's heightfield-fit loss and FP64 adjoint (no reference counterpart)
by central finite differences, the zero-gradient point and the noise floor."""
import numpy as np
import pytest

import oracle as orc
from wavetool_amd import scenes


def data(prob, tgt):
    return orc.FitData(prob["n"], prob["S"], prob["spacing"], prob["x0"], prob["z0"], prob["t0"], prob["dt"], tgt)


@pytest.mark.parametrize("seed,n,S,W", [(5, 12, 3, 3), (6, 9, 5, 6)])
def test_adjoint_matches_central_differences(seed, n, S, W):
    """h = 1e-3 relative, as the fit workload specifies."""
    prob = scenes.fit_problem(seed, n=n, S=S, W_hidden=W, W_fit=W)
    fd = data(prob, scenes.fit_target(prob))
    p = prob["init"]
    _, g = orc.fit_loss_grad(fd, p)
    num = np.zeros_like(p)
    for i in range(W):
        for k in range(5):
            h = 1e-3 * max(abs(p[i, k]), 1e-2)
            pp, pm = p.copy(), p.copy()
            pp[i, k] += h
            pm[i, k] -= h
            num[i, k] = (orc.fit_loss_grad(fd, pp, grad=False)[0] - orc.fit_loss_grad(fd, pm, grad=False)[0]) / (2 * h)
    assert np.abs(g - num).max() <= 1e-4 * np.abs(g).max()


def test_gradient_vanishes_at_noise_free_truth():
    prob = scenes.fit_problem(8, n=10, S=4, W_hidden=4, W_fit=4, noise=0.0)
    fd = data(prob, scenes.fit_target(prob))
    L, g = orc.fit_loss_grad(fd, prob["hidden"])
    # the float32 target leaves residuals of ~1e-8 m
    assert L < 1e-14 and np.abs(g).max() < 1e-7


def test_loss_at_truth_is_the_noise_variance():
    prob = scenes.fit_problem(9, n=40, S=6, W_hidden=3, W_fit=3, noise=0.01)
    fd = data(prob, scenes.fit_target(prob))
    L, _ = orc.fit_loss_grad(fd, prob["hidden"], grad=False)
    assert abs(L - 0.5 * 3 * 1e-4) < 0.1 * 0.5 * 3 * 1e-4


def test_thread_split_is_exact_order_independent_enough():
    prob = scenes.fit_problem(10, n=20, S=3, W_hidden=5, W_fit=5)
    fd = data(prob, scenes.fit_target(prob))
    L1, g1 = orc.fit_loss_grad(fd, prob["init"], nthreads=1)
    L4, g4 = orc.fit_loss_grad(fd, prob["init"], nthreads=4)
    assert abs(L1 - L4) < 1e-15 * max(1, L1) * 1e3 and np.abs(g1 - g4).max() < 1e-12


def test_fp32_instantiation_close_to_fp64():
    prob = scenes.fit_problem(12, n=16, S=4, W_hidden=4, W_fit=4)
    fd = data(prob, scenes.fit_target(prob))
    L64, g64 = orc.fit_loss_grad(fd, prob["init"])
    L32, g32 = orc.fit_loss_grad(fd, prob["init"], precision="f32")
    assert abs(L32 - L64) < 1e-4 * L64 and np.abs(g32 - g64).max() < 1e-3 * np.abs(g64).max()
