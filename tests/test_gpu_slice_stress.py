"""GPU tests of dpd_slice_stress, the z-slice stress profile (CMonitor::AddBeadStress /
AddBondStress, src/Monitor.cpp:1900-2104, :2283-2518), against the oracle's restatement
of the reference's half-shell traversal, and at larger sizes through the sum rule: summed
over slices, the pair and bond tables are the virial the step's observables report."""
import sys

import numpy as np
import pytest

import dpd_amd as D
import oracle as O
from helpers import chain_state

pytestmark = pytest.mark.gpu
A3 = np.array([[25.0, 40.0, 30.0], [40.0, 25.0, 35.0], [30.0, 35.0, 25.0]])
G3 = np.full((3, 3), 4.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sym(t):
    """pair table [ns, nt*nt, 9] -> [ns, nt, nt, 9] with (ti, tj) + (tj, ti): the reference
    files a pair under the type of the bead its traversal meets first, so only this sum is
    fixed for ti != tj."""
    nt = int(round(np.sqrt(t.shape[1])))
    t = t.reshape(t.shape[0], nt, nt, 9)
    return t + t.swapaxes(1, 2)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def chains(cells, seed):
    n = 3 * int(np.prod(cells))
    nch = n // 60
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, nch, 10, n - 10 * nch, seed=seed)
    # bonds are built at length 0.5; unstretched lengths away from it give every bond a force
    k = np.full(len(h), 128.0)
    l0 = np.random.default_rng(seed).uniform(0.3, 0.8, size=len(h))
    return ids, t, x, u, v, (h, tl, k, l0, fi, se, np.full(len(fi), 20.0), np.zeros(len(fi)))


# (nslices, which path): 10 slices of a 3-type box keep the table in LDS (3.2 KB); 150
# slices (48.6 KB) accumulate with global atomics; 37 slices of width 0.30 make most pairs
# span several slices
@pytest.mark.parametrize("nslices", [1, 10, 37, 150])
def test_tables_match_oracle(nslices):
    cells = (8, 9, 11)
    ids, t, x, u, v, topo = chains(cells, seed=31)
    box = O.Box(cells, A3, G3, seed=77)
    step = 13
    pr, br = O.slice_stress(box, ids, t, x, v, step, nslices, unpbc=u, topo=O.Topo(*topo))
    e = D.Engine(cells, A3, G3, seed=77, pair_kernel=D.PAIR_FULLSHELL)
    e.upload(ids, t, x, v, unpbc=u, step0=step)
    e.upload_topology(*topo)
    pg, bg = e.slice_stress(nslices)
    ep, eb = rel(sym(pg), sym(pr)), rel(bg, br)
    say(f"  {nslices} slices: pair {ep:.2e} bond {eb:.2e}")
    assert pg.shape == pr.shape and bg.shape == br.shape
    assert ep < 1e-5
    assert eb < 1e-5
    # the reference's ordering convention for bonds is exact: (head type, tail type)
    nt = 3
    for ti in range(nt):
        for tj in range(nt):
            if not np.any(br[:, ti * nt + tj]):
                assert not np.any(bg[:, ti * nt + tj])
    # a pair-only call gives the same pair table (FP64 atomics sum in any order)
    p2, b2 = e.slice_stress(nslices, bonds=False)
    assert b2 is None and np.allclose(p2, pg, rtol=1e-12, atol=1e-9)


def test_tables_match_oracle_for_one_box_of_a_batch():
    """Box 1 of a two-box batch: its slots and bonds are found inside the batch's arrays."""
    specs = [((7, 8, 9), 5), ((9, 8, 10), 6)]
    boxes = [{"cells": c, "a": A3, "gamma": G3, "seed": s} for c, s in specs]
    e = D.Batch(boxes, pair_kernel=D.PAIR_FULLSHELL)
    states = []
    for i, (cells, _) in enumerate(specs):
        ids, t, x, u, v, topo = chains(cells, seed=40 + i)
        e.upload_box(i, ids, t, x, v, unpbc=u)
        e.upload_box_topology(i, *topo)
        states.append((ids, t, x, u, v, topo))
    e.commit(step0=5)
    for i, (cells, seed) in enumerate(specs):
        ids, t, x, u, v, topo = states[i]
        pr, br = O.slice_stress(O.Box(cells, A3, G3, seed=seed), ids, t, x, v, 5, 12, unpbc=u,
                                topo=O.Topo(*topo))
        pg, bg = e.slice_stress(12, box=i)
        ep, eb = rel(sym(pg), sym(pr)), rel(bg, br)
        say(f"  box {i}: pair {ep:.2e} bond {eb:.2e}")
        assert ep < 1e-5 and eb < 1e-5


@pytest.mark.parametrize("cells,nslices", [((16, 16, 24), 24), ((48, 48, 32), 64)])
def test_sum_rule_after_steps(cells, nslices):
    """After dpd_step the tables hold the last step's forces: summed over slices and type
    pairs they equal V * stress of the step's observables (pairs and bonds)."""
    ids, t, x, u, v, topo = chains(cells, seed=50)
    e = D.Engine(cells, A3, G3, seed=9)
    e.upload(ids, t, x, v, unpbc=u, step0=0)
    e.upload_topology(*topo)
    e.step(6, observe=True)
    pg, bg = e.slice_stress(nslices)
    ob = e.observables()
    V = float(np.prod(cells))
    tot = pg.sum(axis=(0, 1)) + bg.sum(axis=(0, 1))
    err = np.linalg.norm(tot - V * ob["stress"]) / np.linalg.norm(V * ob["stress"])
    say(f"  {len(ids)} beads, {nslices} slices: sum rule {err:.2e}")
    assert err < 1e-5
    # a homogeneous melt has a flat normal-stress profile
    pzz = (pg[..., 8] + bg[..., 8]).sum(axis=1)
    say(f"  P_zz per slice: min {pzz.min():.1f} max {pzz.max():.1f}")
    assert np.all(np.abs(pzz - pzz.mean()) < 0.2 * pzz.mean())


@pytest.mark.parametrize("nslices", [1, 8, 33])
def test_bondpair_tables_match_oracle(nslices):
    """Bond-pair tables (CMonitor::AddBondPairStress) of the reference's branched-lipid
    snapshot with bond pairs at phi0 = 0, 30 and 60 degrees, against the oracle."""
    import os
    from dpd_amd import host as H
    from helpers import GOLDEN, load_beads
    inp = H.Input(os.path.join(GOLDEN, "lipidphi8", "case.dmpci"))
    (h, tl, k, l0), (fi, se, mo, ph) = inp.topology()
    d = load_beads("lipidphi8", "in")
    L = np.array(inp.cells, float)
    x = np.mod(d["u"], L)
    x[x >= L] = 0.0
    box = O.Box(inp.cells, inp.a, inp.gamma, seed=3)
    e = D.Engine(inp.cells, inp.a, inp.gamma, seed=3, pair_kernel=D.PAIR_HITLIST)
    e.upload(d["id"], d["type"], x, d["v"], unpbc=d["u"])
    e.upload_topology(h, tl, k, l0, fi, se, mo, ph)
    got = e.slice_stress_bondpairs(nslices)
    # the oracle at the positions the engine holds (FP32 offsets in their cells)
    X, U = e.download()[:2]
    want = O.slice_stress_bondpairs(box, d["id"], d["type"], X, nslices, unpbc=U,
                                    topo=O.Topo(h, tl, k, l0, fi, se, mo, ph))
    err = rel(got, want)
    say(f"  {len(fi)} bond pairs, {nslices} slices: {err:.2e}")
    assert got.shape == want.shape
    assert err < 1e-9
    assert np.array_equal(np.abs(got).sum(axis=(0, 2)) > 0, np.abs(want).sum(axis=(0, 2)) > 0)


def test_bondpair_sum_rule_on_the_bilayer():
    """The 48x48x32 bilayer (221,175 beads, 22k bond pairs) after ten steps: the bond-pair
    table summed over slices equals sum_i F_i (x) u_i of the bond pairs alone (times
    0.99999999), and its P_zz profile is confined to the lipid layer at the mid-plane."""
    from dpd_amd import workloads as W
    wl = W.bilayer()
    e = D.Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed)
    e.upload(wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u)
    h, tl, k, l0 = wl.bonds
    fi, se, mo, ph = wl.pairs
    e.upload_topology(h, tl, k, l0, fi, se, mo, ph)
    e.step(10)
    X, U, V, F = e.download()
    tab = e.slice_stress_bondpairs(32)
    z = np.zeros_like(wl.a)
    f, _, _ = O.forces(O.Box(wl.cells, z, z), wl.ids, wl.types, X, np.zeros_like(X), step=1, unpbc=U,
                       topo=O.Topo(h, tl, np.zeros(len(h)), l0, fi, se, mo, ph))
    vir = np.einsum("ia,ib->ab", f, U).ravel()
    err = np.linalg.norm(tab.sum(axis=(0, 1)) - 0.99999999 * vir) / np.linalg.norm(vir)
    pzz = tab[..., 8].sum(axis=1)
    say(f"  bilayer bond pairs: sum rule {err:.2e}; P_zz by slice {np.round(pzz, 1)}")
    assert err < 1e-9
    mid = np.abs(np.arange(32) + 0.5 - 16) < 6
    assert np.abs(pzz[~mid]).max() < 1e-3 * np.abs(pzz[mid]).max()
