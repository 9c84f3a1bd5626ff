"""CPU tests of the surrogate network's oracle (FP64 backprop pinned by finite
differences) and of its BF16 rounding."""
import numpy as np

import mlp_oracle as M


def test_bf16_round_to_nearest_even():
    assert M.bf16([1.0])[0] == 1.0
    assert M.bf16([1.0 + 2 ** -8])[0] == 1.0                  # tie -> even
    assert M.bf16([1.0 + 3 * 2 ** -8])[0] == 1.0 + 2 ** -6    # tie -> even (upwards)
    assert M.bf16([1.0 + 2 ** -8 + 2 ** -12])[0] == 1.0 + 2 ** -7
    assert M.bf16([-2.5])[0] == -2.5
    x = np.random.default_rng(0).normal(size=1000)
    assert np.max(np.abs(M.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_oracle_gradient_matches_finite_differences():
    nin, H, nout = 6, 16, 2
    rng = np.random.default_rng(1)
    n = 20
    theta = rng.normal(scale=0.4, size=nin * H + H + H * H + H + H * nout + nout)
    X = rng.normal(size=(n, nin))
    T = rng.normal(size=(n, nout))
    loss, g = M.loss_and_grad(theta, X, T, nin, H, nout)
    h = 1e-6
    for idx in rng.choice(len(theta), size=60, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[idx] += h
        tm[idx] -= h
        fd = (M.loss_and_grad(tp, X, T, nin, H, nout)[0] - M.loss_and_grad(tm, X, T, nin, H, nout)[0]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-7 + 1e-5 * abs(fd), (idx, fd, g[idx])


def test_adam_reference_step():
    theta = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    t1, m, v = M.adam_step(theta, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # the first bias-corrected Adam step moves each parameter by lr * sign(g)
    assert np.allclose(t1, theta - 0.1 * np.sign(g), atol=1e-6)
