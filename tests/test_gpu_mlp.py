"""GPU tests of the surrogate network: BF16-MFMA gradients and Adam against the CPU
oracle (which rounds the same operands to bfloat16), and a fit of a smooth function."""
import numpy as np
import pytest

import mlp_oracle as M
from dpd_amd.mlp import Surrogate

pytestmark = pytest.mark.gpu


def _setup(nin=6, H=256, nout=2, n=200, seed=0):
    rng = np.random.default_rng(seed)
    net = Surrogate(nin, H, nout, seed=seed + 1)
    net.set_scaling(np.zeros(nin), np.ones(nin), np.zeros(nout), np.ones(nout))
    X = rng.normal(size=(n, nin))
    T = rng.normal(size=(n, nout))
    return net, X, T


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_gpu_gradients_match_oracle():
    net, X, T = _setup()
    theta = net.params().astype(np.float64)
    loss, g = net.gradients(X, T)
    lr, gr = M.loss_and_grad(theta, X, T, 6, 256, 2, bf=True)
    print(f"loss {loss:.6e} oracle {lr:.6e}", flush=True)
    assert abs(loss - lr) / lr < 1e-4
    W1, b1, W2, b2, W3, b3 = M.unpack(g, 6, 256, 2)
    R = M.unpack(gr, 6, 256, 2)
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), (W1, b1, W2, b2, W3, b3), R):
        err = _rel(a, b)
        print(f"  grad {name}: rel {err:.2e}", flush=True)
        assert err < 2e-3, name
    # plain FP64 gradients differ only by the BF16 rounding of the products
    _, g64 = M.loss_and_grad(theta, X, T, 6, 256, 2, bf=False)
    assert _rel(g, g64) < 5e-2


def test_gpu_adam_matches_oracle():
    net, X, T = _setup(seed=3)
    theta = net.params().astype(np.float64)
    _, g = net.gradients(X, T)
    net.train(X, T, iters=1, lr=1e-3)
    after = net.params().astype(np.float64)
    ref, _, _ = M.adam_step(theta, g.astype(np.float64), np.zeros_like(theta), np.zeros_like(theta), 1, 1e-3)
    assert np.max(np.abs(after - ref)) < 1e-6


def test_surrogate_fits_smooth_function():
    """An equation-of-state-like target: held-out error well under 2%."""
    rng = np.random.default_rng(5)
    X = rng.uniform([20, 20, 20, 20, 0.05, 64], [90, 90, 90, 90, 0.4, 256], size=(400, 6))
    rho = 3.0
    P = rho + 0.101 * rho ** 2 * (X[:, 0] * (1 - X[:, 4]) + X[:, 1] * X[:, 4]) / 1.0 + 0.002 * X[:, 5]
    A = 1.2 + 0.01 * (X[:, 2] - 50) - 0.3 * X[:, 4]
    Y = np.stack([P, A], axis=1)
    net = Surrogate(6, 256, 2, seed=7)
    net.fit_scaling(X[:320], Y[:320])
    losses = net.train(X[:320], Y[:320], iters=1500, lr=2e-3)
    assert losses[-1] < 0.01 * losses[0]
    pred = net.predict(X[320:])
    err = np.abs(pred[:, 0] - Y[320:, 0]) / Y[320:, 0]
    print(f"held-out pressure error: mean {err.mean():.2e} max {err.max():.2e}", flush=True)
    assert err.mean() < 0.01
