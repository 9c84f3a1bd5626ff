"""GPU tests of the command hooks that act on the engine state between steps:
ZeroTotalMom (CSimBox::ZeroTotalMom, src/SimBox.cpp:1919-1947), RenormaliseMomenta
(CSimBox::UpdateRenormalisedMom, src/SimBox.cpp:3177-3215) and ToggleDPDBeadForces
(CCNTCell::ToggleDPDBeadForces, src/CNTCell.cpp:223-227), checked through the invariants
they establish."""
import sys

import numpy as np
import pytest

import dpd_amd as D
from helpers import random_state

pytestmark = pytest.mark.gpu
A2 = np.array([[25.0, 40.0], [40.0, 25.0]])
G2 = np.full((2, 2), 4.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def drifting(cells, seed, drift=(0.3, -0.2, 0.5), scale=1.3):
    ids, t, x, v = random_state(cells, 3 * int(np.prod(cells)), 2, seed=seed)
    return ids, t, x, scale * v + np.array(drift)


def test_zero_total_momentum():
    cells = (10, 11, 12)
    ids, t, x, v = drifting(cells, 1)
    e = D.Engine(cells, A2, G2, seed=2)
    e.upload(ids, t, x, v)
    e.step(5)
    _, _, v0, _ = e.download()
    e.zero_total_momentum()
    _, _, v1, _ = e.download()
    mean = v0.mean(axis=0)
    say(f"  mean velocity before {mean}, after {v1.mean(axis=0)}")
    assert np.abs(v1.sum(axis=0)).max() < 1e-5 * np.abs(v0).sum()
    # every bead loses the same vector: the mean
    assert np.allclose(v1, v0 - mean, rtol=0, atol=2e-6)


def test_renormalise_momenta():
    cells = (12, 12, 12)
    ids, t, x, v = drifting(cells, 3)
    kT = 1.7
    e = D.Engine(cells, A2, G2, seed=4, kT=kT)
    e.upload(ids, t, x, v)
    e.step(5)
    _, _, v0, _ = e.download()
    e.renormalise_momenta()
    _, _, v1, _ = e.download()
    got = (v1 ** 2).sum() / (3 * len(v1))
    say(f"  sum v^2 / 3N before {(v0 ** 2).sum() / (3 * len(v0)):.5f}, after {got:.7f} (kT {kT})")
    assert abs(got - kT) < 1e-6 * kT
    # a uniform scale: directions kept
    ratio = np.linalg.norm(v1, axis=1) / np.linalg.norm(v0, axis=1)
    assert np.ptp(ratio) < 1e-5 * ratio.mean()


def test_momentum_commands_per_box_of_a_batch():
    specs = [((8, 8, 8), 5), ((9, 8, 10), 6)]
    boxes = [{"cells": c, "a": A2, "gamma": G2, "seed": s, "kT": 1.0 + 0.5 * i} for i, (c, s) in enumerate(specs)]
    e = D.Batch(boxes)
    for i, (c, _) in enumerate(specs):
        ids, t, x, v = drifting(c, 10 + i)
        e.upload_box(i, ids, t, x, v)
    e.commit()
    e.step(3)
    before = [e.download_box(i)[2] for i in range(2)]
    e.zero_total_momentum(box=1)
    after = [e.download_box(i)[2] for i in range(2)]
    assert np.array_equal(before[0], after[0])             # box 0 untouched
    assert np.abs(after[1].mean(axis=0)).max() < 1e-6
    e.renormalise_momenta()                                 # every box to its own kT
    for i in range(2):
        v = e.download_box(i)[2]
        assert abs((v ** 2).sum() / (3 * len(v)) - boxes[i]["kT"]) < 1e-6 * boxes[i]["kT"]


def test_toggle_bead_forces():
    """Forces off: half-step constants zero, so a step is x + dt v (unwrapped) with the
    velocities unchanged; forces back on restore the dynamics."""
    cells = (10, 10, 10)
    ids, t, x, v = random_state(cells, 3000, 2, seed=7)
    dt = 0.02
    e = D.Engine(cells, A2, G2, seed=8, dt=dt)
    e.upload(ids, t, x, v)
    e.step(3)
    x0, u0, v0, _ = e.download()
    e.toggle(D.TOGGLE_FORCES, False)
    e.step(4)
    x1, u1, v1, _ = e.download()
    assert np.array_equal(v1, v0)
    assert np.allclose(u1, u0 + 4 * dt * v0, rtol=0, atol=2e-5)
    L = np.array(cells, float)
    assert np.allclose(np.mod(x1 - u1 + L / 2, L) - L / 2, 0.0, atol=2e-5)   # x is u wrapped into the box
    e.toggle(D.TOGGLE_FORCES, True)
    e.step(2)
    _, _, v2, _ = e.download()
    say(f"  |dv| after forces back on {np.abs(v2 - v1).mean():.3f}")
    assert np.abs(v2 - v1).mean() > 1e-2
