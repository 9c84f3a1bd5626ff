"""Pair-kernel selection on the GPU: the automatic choice by bead count, dpd_autotune (times
every kernel, restores the state after each trial, keeps the fastest), and that the chosen
kernel still reproduces the oracle's forces."""
import sys

import numpy as np
import pytest

import dpd_amd as D
import oracle as O
from helpers import force_errors, random_state

pytestmark = pytest.mark.gpu
A2 = np.array([[25.0, 40.0], [40.0, 25.0]])
G2 = np.full((2, 2), 4.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_automatic_choice_by_bead_count():
    cells = (10, 10, 10)
    ids, t, x, v = random_state(cells, 3000, 2, seed=8)
    e = D.Engine(cells, A2, G2, seed=3, pair_kernel=D.PAIR_AUTO)
    e.upload(ids, t, x, v)
    assert e.pair_kernel == D.PAIR_HITLIST
    e.set_pair_kernel(D.PAIR_FULLSHELL)
    assert e.pair_kernel == D.PAIR_FULLSHELL
    e.set_pair_kernel(D.PAIR_AUTO)
    assert e.pair_kernel == D.PAIR_HITLIST
    with pytest.raises(D.DpdError):
        e.set_pair_kernel(D.PAIR_KINDS + 1)
    # the retired kernels are refused by name
    for k in (D.PAIR_BRICK, D.PAIR_BRICK512, D.PAIR_STREAM, D.PAIR_STREAM1024, D.PAIR_COMPACT, D.PAIR_CELLS):
        with pytest.raises(D.DpdError) as err:
            e.set_pair_kernel(k)
        assert err.value.code == 8 and "retired" in str(err.value)
    assert e.pair_kernel == D.PAIR_HITLIST


def test_autotune_restores_state_and_keeps_parity():
    cells = (12, 12, 12)
    ids, t, x, v = random_state(cells, 5184, 2, seed=9)
    box = O.Box(cells, A2, G2, seed=4, rng_mode=O.RNG_PHILOX)
    s = O.System(box, ids, t, x, v)
    s.run(20)
    x, u, v, f = s.state()
    e = D.Engine(cells, A2, G2, seed=4, pair_kernel=D.PAIR_FULLSHELL)
    e.upload(ids, t, x, v, f=f, unpbc=u, step0=20)
    before = e.download()
    chosen, ms = e.autotune(nsteps=3)
    say(f"  autotune chose {chosen}: " + ", ".join(f"{k}: {m:.3f} ms" for k, m in ms.items()))
    assert 1 <= chosen <= D.PAIR_KINDS and e.pair_kernel == chosen
    avail = {k: m for k, m in ms.items() if k in (D.PAIR_FULLSHELL, D.PAIR_HITLIST)}
    assert all(m > 0 for m in avail.values()) and ms[chosen] == min(avail.values())
    assert all(m == 0 for k, m in ms.items() if k not in avail)
    assert e.step_count == 20
    after = e.download()
    for b, a in zip(before, after):
        assert np.array_equal(b, a)
    # the tuned engine still steps like the oracle
    e.step(1)
    s.run(1, step0=20)
    X, U, V, F = e.download()
    norm, _ = force_errors(F, s.state()[3])
    say(f"  force normwise after autotune {norm:.2e}")
    assert norm < 1e-5


def test_pair_timing_every_launch_or_sampled():
    cells = (12, 12, 12)
    ids, t, x, v = random_state(cells, 5184, 2, seed=10)
    e = D.Engine(cells, A2, G2, seed=4, pair_kernel=D.PAIR_HITLIST)
    e.upload(ids, t, x, v)
    e.step(5)
    e.step(20, timing=True)
    n, ms = e.last_timing()
    assert n == 20 and 0 < ms[1] < ms[0] and abs(ms[2] - (ms[0] - ms[1])) < 1e-9
    e.step(40, timing="sampled")
    n2, ms2 = e.last_timing()
    assert n2 == 2                                  # steps 15 and 31 of the call
    say(f"  pair per launch: all {ms[1] / n * 1e3:.1f} us, sampled {ms2[1] / n2 * 1e3:.1f} us")
    assert 0.5 < (ms2[1] / n2) / (ms[1] / n) < 2.0


@pytest.mark.parametrize("kernel", [D.PAIR_HITLIST, D.PAIR_FULLSHELL])
def test_pair_launch_timing_and_diagnostics_leave_the_state(kernel):
    """dpd_time_pair (back-to-back launches between two events) and dpd_compute_forces after
    steps are diagnostics: the trajectory continues exactly as in an engine that ran neither.
    With chains, so that the hit-list launch carries the bonded terms."""
    from helpers import chain_state
    cells = (12, 12, 12)
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, 40, 16, 4544, seed=12)
    k, l0 = np.full(len(h), 128.0), np.full(len(h), 0.5)
    mod, ph = np.full(len(fi), 20.0), np.zeros(len(fi))
    A3 = np.array([[25.0, 25.0, 25.0], [25.0, 10.0, 25.0], [25.0, 25.0, 25.0]])
    G3 = np.full((3, 3), 4.5)
    runs = []
    for diag in (False, True):
        e = D.Engine(cells, A3, G3, seed=6, pair_kernel=kernel)
        e.upload(ids, t, x, v, unpbc=u)
        e.upload_topology(h, tl, k, l0, fi, se, mod, ph)
        e.step(7)
        if diag:
            ms = e.time_pair(5)
            say(f"  kernel {kernel}: {ms * 1e3:.1f} us per pair launch")
            assert ms > 0
            e.compute_forces(step=3)
            e.compute_forces(step=7, thermostat=False)
        e.step(5, observe=True)
        runs.append((e.download(), e.observables()))
    for a, b in zip(runs[0][0], runs[1][0]):
        assert np.array_equal(a, b)
    # kT is a sum of block sums added with FP64 atomics, in any order: equal to rounding
    assert abs(runs[0][1]["kT"] - runs[1][1]["kT"]) <= 1e-12 * runs[0][1]["kT"]


SCHEDULES = list(range(D.schedule_count()))
A3 = np.array([[25.0, 30.0, 35.0], [30.0, 10.0, 25.0], [35.0, 25.0, 40.0]])
G3 = np.full((3, 3), 4.5)


@pytest.mark.parametrize("sched", SCHEDULES, ids=[D.schedule_name(k) for k in SCHEDULES])
def test_every_schedule_matches_the_oracle(sched):
    """Every schedule of the pair launch against the FP64 oracle: forces, virial and energies
    of one box with chains and bond pairs; five steps of it (the hit-list launch then carries
    the bonded terms); and forces of a batch of uneven boxes down to three cells per axis,
    where every neighbour row wraps through a periodic face."""
    from helpers import chain_state
    cells = (12, 11, 10)
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, 30, 16, 3 * 1320 - 480, seed=31)
    k, l0 = np.full(len(h), 128.0), np.full(len(h), 0.5)
    mod, ph = np.full(len(fi), 20.0), np.full(len(fi), 0.3)
    box = O.Box(cells, A3, G3, seed=6, rng_mode=O.RNG_PHILOX)
    topo = O.Topo(h, tl, k, l0, fi, se, mod, ph)
    f0, _, _ = O.forces(box, ids, t, x, v, step=0, unpbc=u, topo=topo)
    e = D.Engine(cells, A3, G3, seed=6, pair_kernel=D.PAIR_HITLIST)
    e.upload(ids, t, x, v, f=f0, unpbc=u, step0=0)
    e.upload_topology(h, tl, k, l0, fi, se, mod, ph)
    e.set_schedule(sched)
    assert e.schedule == sched
    say(f"schedule {sched} {D.schedule_name(sched)}")
    F, vir, pe = e.compute_forces(step=5)
    fr, virr, per = O.forces(box, ids, t, x, v, step=5, unpbc=u, topo=topo)
    norm, per_bead = force_errors(F, fr)
    say(f"  one box: force normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < 1e-5
    assert np.allclose(vir, virr, rtol=1e-5, atol=1e-6 * np.abs(virr).max())
    assert np.allclose(pe, per, rtol=1e-5)
    s = O.System(box, ids, t, x, v, f0, unpbc=u, topo=topo)
    e.step(5)
    s.run(5, step0=0)
    X, U, V, F = e.download()
    xr, ur, vr, fr = s.state()
    norm, _ = force_errors(F, fr)
    say(f"  five steps: force normwise {norm:.2e}, |u - u_ref| max {np.abs(U - ur).max():.1e}")
    assert norm < 1e-5 and np.abs(U - ur).max() < 1e-5
    # a batch (each hit-list workgroup takes slots of one box)
    rng = np.random.default_rng(7)
    specs = [((3, 3, 3), 2), ((3, 4, 5), 3), ((9, 10, 11), 3), ((5, 3, 4), 1), ((13, 8, 9), 2)]
    boxes = []
    for i, (bc, nt) in enumerate(specs):
        a = rng.uniform(20, 90, size=(nt, nt))
        boxes.append({"cells": bc, "a": 0.5 * (a + a.T), "gamma": np.full((nt, nt), 4.5), "seed": 50 + i,
                      "ntypes": nt})
    b = D.Batch(boxes, pair_kernel=D.PAIR_HITLIST)
    states = []
    for i, bx in enumerate(boxes):
        n = 3 * int(np.prod(bx["cells"]))
        bids, bt, bxx, bv = random_state(bx["cells"], n, bx["ntypes"], seed=60 + i)
        b.upload_box(i, bids, bt, bxx, bv)
        states.append((bids, bt, bxx, bv))
    b.commit(step0=0)
    b.set_schedule(sched)
    f, _, _ = b.compute_forces(step=9)
    off = 0
    for i, bx in enumerate(boxes):
        bids, bt, bxx, bv = states[i]
        n = len(bids)
        ob = O.Box(bx["cells"], bx["a"], bx["gamma"], seed=bx["seed"])
        fr, _, _ = O.forces(ob, bids, bt, bxx, bv, step=9)
        norm, _ = force_errors(f[off:off + n], fr)
        assert norm < 1e-5, (i, bx["cells"], norm)
        off += n
    say(f"  batch of {len(boxes)} boxes: forces within 1e-5 of the oracle")


def test_autotune_schedules_and_the_tuning_file(tmp_path):
    """dpd_autotune_schedules times every schedule, restores the state after each trial and
    keeps the fastest; the choice, written to a tuning file under the engine's shape, is what
    dpd_apply_tuning sets on a fresh engine of that shape, and nothing on another shape."""
    from dpd_amd import tune as T
    cells = (14, 14, 14)
    ids, t, x, v = random_state(cells, 3 * 14 ** 3, 2, seed=13)
    e = D.Engine(cells, A2, G2, seed=4, pair_kernel=D.PAIR_AUTO)
    e.upload(ids, t, x, v)
    e.step(10)
    before = e.download()
    path = tmp_path / "tune.json"
    chosen, ms = T.tune(e, 1, len(ids), 2, False, nsteps=3, path=str(path))
    say(f"  chose {D.schedule_name(chosen)}: " + ", ".join(f"{k} {m:.4f}" for k, m in ms.items()))
    assert all(m > 0 for m in ms.values()) and ms[D.schedule_name(chosen)] == min(ms.values())
    assert e.schedule == chosen
    for p, q in zip(before, e.download()):
        assert np.array_equal(p, q)
    rec = T.load(str(path))
    assert len(rec) == 1 and rec[0]["schedule"] == chosen and rec[0]["beads_lo"] <= len(ids) < rec[0]["beads_hi"]
    f = D.Engine(cells, A2, G2, seed=5, pair_kernel=D.PAIR_AUTO)
    f.upload(ids, t, x, v)
    other = (chosen + 1) % len(SCHEDULES)
    f.set_schedule(other)
    assert T.apply(f, str(path)) == chosen and f.schedule == chosen
    small = D.Engine((6, 6, 6), A2, G2, seed=5, pair_kernel=D.PAIR_AUTO)
    sids, st, sx, sv = random_state((6, 6, 6), 648, 2, seed=14)
    small.upload(sids, st, sx, sv)
    s0 = small.schedule
    assert T.apply(small, str(path)) is None and small.schedule == s0
