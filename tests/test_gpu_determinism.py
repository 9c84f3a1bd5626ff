"""Run-to-run determinism of the step: the counting sort orders the beads of every cell by
bead index, and every sum a bead's force takes is made in that order, so two runs of 200
steps from the same state -- eagerly and as replayed graphs -- end with bitwise identical
positions, velocities and forces.  The FP64 observables are block sums added with atomics,
whose order varies: they agree to rounding."""
import sys

import numpy as np
import pytest

import dpd_amd as D
from dpd_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(wl, kernel, graph, nsteps=200):
    e = D.Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed, pair_kernel=kernel)
    e.upload(wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u)
    if wl.bonds is not None:
        h, t, k, l0 = wl.bonds
        fi, se, mo, ph = wl.pairs if wl.pairs is not None else ((), (), (), ())
        e.upload_topology(h, t, k, l0, fi, se, mo, ph)
    e.step(nsteps, observe=True, graph=graph)
    out = e.download()
    ob = e.observables()
    e.close()
    return out, ob


@pytest.mark.parametrize("kernel", [pytest.param(D.PAIR_HITLIST, id="hitlist"),
                                    pytest.param(D.PAIR_FULLSHELL, id="fullshell")])
@pytest.mark.parametrize("name", ["bilayer", "condensate"])
def test_two_runs_are_bitwise_identical(name, kernel):
    wl = W.bilayer() if name == "bilayer" else W.condensate()
    say(f"determinism: {name} ({wl.n} beads, {wl.nbonds} bonds, {wl.nangles} angles), kernel {kernel}")
    (a, oa) = run(wl, kernel, graph=False)
    (b, ob) = run(wl, kernel, graph=False)
    (c, oc) = run(wl, kernel, graph=True)
    for what, p, q, r in zip("xuvf", a, b, c):
        say(f"  {what}: eager/eager max diff {np.abs(p - q).max():.1e}, eager/graph {np.abs(p - r).max():.1e}")
        assert np.array_equal(p, q) and np.array_equal(p, r)
    for o in (ob, oc):
        for k in ("kT", "pressure"):
            assert abs(o[k] - oa[k]) <= 1e-12 * abs(oa[k])
        assert np.allclose(o["stress"], oa["stress"], rtol=1e-11, atol=1e-12)
