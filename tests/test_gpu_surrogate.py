"""The learned part end to end: batched replica runs on the GPU produce (pressure, area
per lipid) for random parameter sets in boxes of side L ~ U{8..24}; the BF16 surrogate
(7 -> 256 -> 256 -> 2: the six run parameters and the box side) trained on 80% of them
must predict the held-out pressure within 2% (mean, 95th percentile, and for 95% of the
held-out replicas; the largest error is reported); screening a large sweep with it costs a
tiny fraction of simulating the sweep, and the simulation checks the accepted candidates."""
import time

import numpy as np
import pytest

from dpd_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dataset():
    # 2048 replicas of 6000 steps: ~32M beads advanced together, ~50 s on one MI355X
    rng = np.random.default_rng(2026)
    n = 2048
    params = S.sample_params(n, rng)
    sizes = S.sample_sizes(n, rng)
    seeds = rng.integers(1, 2 ** 31, size=n)
    res = S.run_replicas(params, sizes, seeds, steps=6000, discard=2000)
    print(f"replicas: {n} boxes, L {sizes.min()}..{sizes.max()}, {res['nbeads']} beads, "
          f"{res['bead_steps_per_s']:.3e} bead-steps/s, wall {res['wall']:.1f} s", flush=True)
    return params, sizes, res


def _split(params, sizes, res):
    X = S.features(params, sizes)
    Y = np.stack([res["P"], res["apl"]], axis=1)
    ok = np.isfinite(Y).all(axis=1)
    X, Y = X[ok], Y[ok]
    return X, Y, int(0.8 * len(X))


def test_replica_pressures_are_converged(dataset):
    params, sizes, res = dataset
    rel_se = res["P_se"] / res["P"]
    print(f"pressure: {res['P'].min():.2f}..{res['P'].max():.2f}, relative s.e. median "
          f"{np.median(rel_se):.2e} max {rel_se.max():.2e}", flush=True)
    assert np.isfinite(res["P"]).all()
    assert np.median(rel_se) < 0.002 and rel_se.max() < 0.01
    assert np.all(np.abs(res["kT"] - 1.0) < 0.05)


def test_surrogate_predicts_held_out_pressure(dataset):
    params, sizes, res = dataset
    X, Y, ntr = _split(params, sizes, res)
    net, losses = S.train(X[:ntr], Y[:ntr], iters=3000, lr=2e-3)
    pred = net.predict(X[ntr:])
    errP = np.abs(pred[:, 0] - Y[ntr:, 0]) / Y[ntr:, 0]
    errA = np.abs(pred[:, 1] - Y[ntr:, 1]) / Y[ntr:, 1]
    # the method without the network: predict the training mean
    base = np.abs(Y[:ntr, 0].mean() - Y[ntr:, 0]) / Y[ntr:, 0]
    p95 = np.percentile(errP, 95)
    within = np.mean(errP < 0.02)
    print(f"held-out pressure error ({len(errP)} replicas): mean {errP.mean():.4f} median {np.median(errP):.4f} "
          f"95th percentile {p95:.4f} max {errP.max():.4f}, within 2%: {within:.3f} (training-mean predictor: "
          f"mean {base.mean():.4f}); area per lipid: mean {errA.mean():.4f} 95th percentile "
          f"{np.percentile(errA, 95):.4f}", flush=True)
    assert losses[-1] < 0.05 * losses[0]
    assert errP.mean() < 0.02 and p95 < 0.02 and within >= 0.95
    assert errP.mean() < 0.2 * base.mean()
    assert errA.mean() < 0.1


def test_screening_gain_and_simulation_checks(dataset):
    params, sizes, res = dataset
    X, Y, _ = _split(params, sizes, res)
    net, _ = S.train(X, Y, iters=3000, lr=2e-3)
    rng = np.random.default_rng(99)
    cand_p = S.sample_params(20000, rng)
    cand_L = S.sample_sizes(20000, rng)
    cand = S.features(cand_p, cand_L)
    target = float(np.median(res["P"]))
    t0 = time.perf_counter()
    keep, pred = S.screen(net, cand, target, tol=0.02)
    t_pred = time.perf_counter() - t0
    assert len(keep) > 0
    checked = keep[:32]
    sim = S.run_replicas(cand_p[checked], cand_L[checked], rng.integers(1, 2 ** 31, size=len(checked)),
                         steps=3000, discard=1500)
    err = np.abs(pred[checked, 0] - sim["P"]) / sim["P"]
    t_sim_per = sim["wall"] / len(checked)
    t_pred_per = t_pred / len(cand)
    print(f"screened {len(cand)} candidates in {t_pred * 1e3:.1f} ms, accepted {len(keep)}; checked {len(checked)}: "
          f"mean error {err.mean():.4f}, max {err.max():.4f}, within 2%: {np.mean(err < 0.02):.2f}; cost per "
          f"candidate {t_pred_per * 1e6:.2f} us vs {t_sim_per * 1e3:.1f} ms simulated (x{t_sim_per / t_pred_per:.0f})",
          flush=True)
    assert err.mean() < 0.02 and np.mean(err < 0.02) >= 0.9
    assert t_sim_per / t_pred_per > 1000
