"""GPU tests of batched replicas (many boxes advanced by the same kernels) and of
hipGraph replay, against the FP64 oracle box by box."""
import sys

import numpy as np
import pytest

import dpd_amd as D
import oracle as O
from helpers import chain_state, force_errors, random_state, rel_norm

pytestmark = pytest.mark.gpu

KERNELS = [pytest.param(D.PAIR_FULLSHELL, id="fullshell"), pytest.param(D.PAIR_HITLIST, id="hitlist")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_boxes(rng):
    specs = [((8, 8, 8), 2), ((9, 10, 11), 3), ((12, 12, 12), 2), ((6, 7, 5), 1), ((13, 8, 9), 3)]
    boxes = []
    for i, (cells, nt) in enumerate(specs):
        a = rng.uniform(20, 90, size=(nt, nt))
        a = 0.5 * (a + a.T)
        g = np.full((nt, nt), 4.5)
        boxes.append({"cells": cells, "a": a, "gamma": g, "seed": 100 + i, "dt": 0.02 if i % 2 else 0.015,
                      "ntypes": nt})
    return boxes


@pytest.mark.parametrize("kernel", KERNELS)
def test_batch_forces_match_oracle_per_box(kernel):
    rng = np.random.default_rng(3)
    boxes = make_boxes(rng)
    say("batch forces", kernel, [b["cells"] for b in boxes])
    e = D.Batch(boxes, pair_kernel=kernel)
    states = []
    for i, b in enumerate(boxes):
        n = 3 * int(np.prod(b["cells"]))
        ids, t, x, v = random_state(b["cells"], n, b["ntypes"], seed=10 + i)
        e.upload_box(i, ids, t, x, v)
        states.append((ids, t, x, v))
    e.commit(step0=0)
    f, vir, pe = e.compute_forces(step=9)
    off = 0
    for i, b in enumerate(boxes):
        ids, t, x, v = states[i]
        n = len(ids)
        box = O.Box(b["cells"], b["a"], b["gamma"], dt=b["dt"], seed=b["seed"])
        fr, _, _ = O.forces(box, ids, t, x, v, step=9)
        norm, per_bead = force_errors(f[off:off + n], fr)
        say(f"  box {i} normwise {norm:.2e} per-bead {per_bead:.2e}")
        assert norm < 1e-5
        off += n


@pytest.mark.parametrize("kernel", KERNELS)
def test_batch_steps_match_oracle_per_box(kernel):
    rng = np.random.default_rng(4)
    boxes = make_boxes(rng)
    e = D.Batch(boxes, pair_kernel=kernel)
    refs = []
    for i, b in enumerate(boxes):
        cells = b["cells"]
        nch = int(np.prod(cells)) // 40
        ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, nch, 8, 3 * int(np.prod(cells)) - 8 * nch, seed=20 + i)
        t = np.minimum(t, b["ntypes"] - 1).astype(np.int32)
        k = np.full(len(h), 128.0)
        l0 = np.full(len(h), 0.5)
        mod = np.full(len(fi), 20.0)
        ph = np.zeros(len(fi))
        box = O.Box(cells, b["a"], b["gamma"], dt=b["dt"], seed=b["seed"])
        topo = O.Topo(h, tl, k, l0, fi, se, mod, ph)
        f0, _, _ = O.forces(box, ids, t, x, v, step=0, unpbc=u, topo=topo)
        e.upload_box(i, ids, t, x, v, f=f0, unpbc=u)
        e.upload_box_topology(i, h, tl, k, l0, fi, se, mod, ph)
        s = O.System(box, ids, t, x, v, f0, unpbc=u, topo=topo)
        refs.append(s)
    e.commit(step0=0)
    e.step(3, observe=True)
    for i, s in enumerate(refs):
        s.run(3, step0=0)
        xr, ur, vr, fr = s.state()
        X, U, V, F = e.download_box(i)
        say(f"  box {i}: u {rel_norm(U, ur):.2e} v {rel_norm(V, vr):.2e} f {rel_norm(F, fr):.2e}")
        assert rel_norm(U, ur) < 1e-6
        assert rel_norm(V, vr) < 1e-4
        assert rel_norm(F, fr) < 1e-4
        ob, obr = e.observables(i), s.observables()
        assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < 1e-4
        assert abs(ob["pressure"] - obr["pressure"]) / obr["pressure"] < 1e-4
        assert ob["nbeads"] == len(ur)


@pytest.mark.parametrize("kernel", KERNELS)
def test_graph_replays_match_eager(kernel):
    """Each replay of the captured step graph, with eager work and synchronisation in
    between, against the same steps run eagerly."""
    cells = (12, 12, 12)
    n = 3 * 12 ** 3
    ids, t, x, v = random_state(cells, n, 2, seed=6)
    a = np.array([[25.0, 40.0], [40.0, 25.0]])
    g = np.full((2, 2), 4.5)
    s = O.System(O.Box(cells, a, g, seed=2), ids, t, x, v)
    s.run(20)
    x, u, v, f = s.state()
    eager = D.Engine(cells, a, g, seed=2, pair_kernel=kernel)
    graph = D.Engine(cells, a, g, seed=2, pair_kernel=kernel)
    for e in (eager, graph):
        e.upload(ids, t, x, v, f=f, unpbc=u, step0=20)
    for rep in range(3):
        eager.step(16, observe=True)
        graph.step(16, observe=True, graph=True)
        Xe, Ue, Ve, Fe = eager.download()
        Xg, Ug, Vg, Fg = graph.download()
        say(f"  replay {rep}: u {rel_norm(Ug, Ue):.2e} v {rel_norm(Vg, Ve):.2e} f {rel_norm(Fg, Fe):.2e}")
        assert eager.step_count == graph.step_count == 20 + 16 * (rep + 1)
        assert rel_norm(Ug, Ue) < 1e-6
        assert rel_norm(Vg, Ve) < 1e-4
        assert rel_norm(Fg, Fe) < 1e-4
        oe, og = eager.observables(), graph.observables()
        assert abs(oe["kT"] - og["kT"]) / oe["kT"] < 1e-4


@pytest.mark.parametrize("kernel", [pytest.param(D.PAIR_HITLIST, id="hitlist"),
                                    pytest.param(D.PAIR_FULLSHELL, id="fullshell")])
def test_graph_replays_with_bonds_match_eager(kernel):
    """Chains and angles under graph replay (with the hit-list kernel the bond and angle
    terms run in the pair launch): every replay against the same steps run eagerly."""
    cells = (12, 12, 12)
    nch = 60
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, nch, 8, 3 * 12 ** 3 - 8 * nch, seed=33)
    t = np.minimum(t, 1).astype(np.int32)
    a = np.array([[25.0, 40.0], [40.0, 25.0]])
    g = np.full((2, 2), 4.5)
    k = np.full(len(h), 128.0)
    l0 = np.full(len(h), 0.5)
    mod = np.full(len(fi), 20.0)
    ph = np.zeros(len(fi))
    eager = D.Engine(cells, a, g, seed=5, pair_kernel=kernel)
    graph = D.Engine(cells, a, g, seed=5, pair_kernel=kernel)
    for e in (eager, graph):
        e.upload(ids, t, x, v, unpbc=u)
        e.upload_topology(h, tl, k, l0, fi, se, mod, ph)
    for rep in range(3):
        eager.step(12, observe=True)
        graph.step(12, observe=True, graph=True)
        Xe, Ue, Ve, Fe = eager.download()
        Xg, Ug, Vg, Fg = graph.download()
        say(f"  replay {rep}: u {rel_norm(Ug, Ue):.2e} v {rel_norm(Vg, Ve):.2e} f {rel_norm(Fg, Fe):.2e}")
        assert rel_norm(Ug, Ue) < 1e-6 and rel_norm(Vg, Ve) < 1e-4 and rel_norm(Fg, Fe) < 1e-4
        oe, og = eager.observables(), graph.observables()
        assert abs(oe["pressure"] - og["pressure"]) / oe["pressure"] < 1e-4
        assert abs(oe["pe"][1] - og["pe"][1]) <= 1e-4 * abs(oe["pe"][1])
