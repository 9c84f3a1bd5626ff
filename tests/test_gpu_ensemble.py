"""Statistical parity of thermostatted runs: the GPU engine (counter-based pair random
stream) against several independent runs of the reference (serial LCG stream) on the
same control files.  The GPU side runs 8 replicas from independent starting states in
one batched engine.  The bilayer runs of the reference start from its own lamella builder
and the GPU replicas from the package's restatement of it, which gives the same states for
the same seeds (the starting structure alone moves 4000-step bilayer averages by ~0.2% in
pressure, so the starts must agree).  Criteria: kT within 0.5% and
pressure within 1% of the reference ensemble mean, and both consistent with it within 4
standard errors of the run-to-run spread; g(r) and bilayer density profiles within their
combined error bars."""
import json
import os

import numpy as np
import pytest

from dpd_amd import analysis as A
from dpd_amd import host as H
from helpers import GOLDEN

pytestmark = pytest.mark.gpu
ENS = os.path.join(GOLDEN, "ensemble")
REPLICAS = 8


def _run(name, discard, snap_from, snap_every, seeds, unwrapped=False):
    inp = H.Input(os.path.join(ENS, name + ".dmpci"))
    e, meta = H.replica_batch([inp] * REPLICAS, seeds, pair_kernel=0)
    steps = inp.counts["total_time"]
    hist = [[] for _ in range(REPLICAS)]
    snaps = [[] for _ in range(REPLICAS)]
    for s in range(10, steps + 1, 10):
        e.step(10, observe=True, graph=True)
        for b in range(REPLICAS):
            o = e.observables(b)
            if s > discard:
                hist[b].append((o["kT"], o["pressure"]))
        if s >= snap_from and s % snap_every == 0:
            for b in range(REPLICAS):
                X, U = e.download_box(b)[:2]
                snaps[b].append((X, U) if unwrapped else X)
    assert e.observables(0)["jumps"] == 0
    runs = [np.array(h) for h in hist]
    return runs, snaps, meta[0][0], meta[0][2]


def _ensemble(values):
    v = np.asarray(values, float)
    return v.mean(), v.std(ddof=1) / np.sqrt(len(v))


def _close(name, gpu_runs, ref_runs, rel, nsig=4.0):
    m, se = _ensemble(gpu_runs)
    mr, ser = _ensemble(ref_runs)
    print(f"  {name}: gpu {m:.5f} +- {se:.5f} ({len(gpu_runs)} runs)  reference {mr:.5f} +- {ser:.5f} "
          f"({len(ref_runs)} runs)  rel diff {abs(m - mr) / abs(mr):.2e}", flush=True)
    assert abs(m - mr) / abs(mr) < rel
    assert abs(m - mr) < nsig * np.hypot(se, ser)


def test_water_ensemble():
    ref = json.load(open(os.path.join(ENS, "water20.json")))
    runs, snaps, types, L = _run("water20", ref["discard"], 1250, 250, [4242 + 104729 * r for r in range(REPLICAS)])
    _close("kT", [A.block_average(r[:, 0])[0] for r in runs], [r["kT"][0] for r in ref["runs"]], 0.005)
    _close("P", [A.block_average(r[:, 1])[0] for r in runs], [r["P"][0] for r in ref["runs"]], 0.01)
    gs = np.array([A.rdf(x, L)[1] for s in snaps for x in s])
    g, gse = gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))
    r = np.array(ref["rdf_r"])
    sel = r > 0.3
    dg = np.abs(g - np.array(ref["rdf_g"]))[sel]
    z = dg / np.hypot(gse, np.array(ref["rdf_se"]))[sel]
    print(f"  g(r): max |dg| {dg.max():.4f}, mean |dg|/sigma {z.mean():.2f}, max {z.max():.2f}", flush=True)
    assert dg.max() < 0.02
    assert np.mean(z) < 2.0


def test_bilayer_ensemble():
    ref = json.load(open(os.path.join(ENS, "bilayer16.json")))
    # the reference runs started from its lamella builder with these seeds, which this
    # build's builder reproduces: the first replicas start from the same states, the rest
    # from further seeds
    seeds = list(ref["seeds"]) + [9090 + 104729 * r for r in range(REPLICAS - len(ref["seeds"]))]
    runs, snaps, types, L = _run("bilayer16", ref["discard"], 2250, 250, seeds)
    _close("kT", [A.block_average(r[:, 0])[0] for r in runs], [r["kT"][0] for r in ref["runs"]], 0.005)
    _close("P", [A.block_average(r[:, 1])[0] for r in runs], [r["P"][0] for r in ref["runs"]], 0.01)
    profs = np.array([A.density_profile(x, types, L, centre_type=1, dz=0.5, ntypes=3)[1] for s in snaps for x in s])
    p, pse = profs.mean(0), profs.std(0, ddof=1) / np.sqrt(len(profs))
    pr, prse = np.array(ref["profile"]), np.array(ref["profile_se"])
    dev = np.abs(p - pr) / (np.hypot(pse, prse) + 0.02)
    print(f"  density profiles: mean dev/sigma {dev.mean():.2f}, max {dev.max():.2f}", flush=True)
    for t, name in enumerate(ref["types"]):
        # beads of each type in the profile window agree
        assert abs(p[t].sum() - pr[t].sum()) / pr[t].sum() < 0.02, name
    assert np.mean(dev) < 2.0
    nlip = ref["nlipids"]
    apl = [np.mean([A.area_per_lipid(x, types, L, 0, 1, nlip) for x in s]) for s in snaps]
    _close("area per lipid", apl, [r["apl"] for r in ref["runs"]], 0.03)


def test_binary_ensemble():
    """Water/oil 50/50 at a_WO = 80 from random starts: kT, P and the W-W / W-O g(r) while
    the mixture demixes (snapshots from step 2250 on)."""
    ref = json.load(open(os.path.join(ENS, "binary20.json")))
    runs, snaps, types, L = _run("binary20", ref["discard"], ref["snap_from"], ref["snap_every"],
                                 [5151 + 104729 * r for r in range(REPLICAS)])
    _close("kT", [A.block_average(r[:, 0])[0] for r in runs], [r["kT"][0] for r in ref["runs"]], 0.005)
    _close("P", [A.block_average(r[:, 1])[0] for r in runs], [r["P"][0] for r in ref["runs"]], 0.01)
    r = np.array(ref["rdf_r"])
    sel = r > 0.3
    for pair, key in (((0, 0), "g_ww"), ((0, 1), "g_wo")):
        gs = np.array([A.rdf(x, L, types=types, pair=pair)[1] for s in snaps for x in s])
        g, gse = gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))
        dg = np.abs(g - np.array(ref[key]))[sel]
        # floor: at short range g_WO is zero on both sides, with zero spread
        z = dg / (np.hypot(gse, np.array(ref[key + "_se"]))[sel] + 0.005)
        print(f"  {key}: max |dg| {dg.max():.4f}, mean |dg|/sigma {z.mean():.2f}, max {z.max():.2f}", flush=True)
        assert dg.max() < 0.05
        assert np.mean(z) < 2.0


def test_condensate_ensemble():
    """The condensate composition (16-mers (S P P P)x4 at 20%, a_SS = 10, Hookean bonds) on a
    20^3 box, both sides from the package's random states: kT, P, mean bond length and the
    sticker-sticker g(r)."""
    ref = json.load(open(os.path.join(ENS, "condensate20.json")))
    seeds = list(ref["seeds"]) + [6161 + 104729 * r for r in range(REPLICAS - len(ref["seeds"]))]
    runs, snaps, types, L = _run("condensate20", ref["discard"], ref["snap_from"], ref["snap_every"], seeds,
                                 unwrapped=True)
    _close("kT", [A.block_average(r[:, 0])[0] for r in runs], [r["kT"][0] for r in ref["runs"]], 0.005)
    _close("P", [A.block_average(r[:, 1])[0] for r in runs], [r["P"][0] for r in ref["runs"]], 0.01)
    (bh, bt, _, _), _ = H.Input(os.path.join(ENS, "condensate20.dmpci")).topology()
    blen = [np.mean([np.linalg.norm(u[bh] - u[bt], axis=1).mean() for _, u in s]) for s in snaps]
    _close("bond length", blen, [r["bond_length"] for r in ref["runs"]], 0.005)
    r = np.array(ref["rdf_r"])
    sel = r > 0.3
    gs = np.array([A.rdf(x, L, types=types, pair=(1, 1))[1] for s in snaps for x, _ in s])
    g, gse = gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))
    dg = np.abs(g - np.array(ref["g_ss"]))[sel]
    z = dg / (np.hypot(gse, np.array(ref["g_ss_se"]))[sel] + 0.005)
    print(f"  g_ss: max |dg| {dg.max():.4f}, mean |dg|/sigma {z.mean():.2f}, max {z.max():.2f}", flush=True)
    assert np.mean(z) < 2.0
