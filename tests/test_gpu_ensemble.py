"""Statistical parity of thermostatted runs: the GPU engine (counter-based pair random
stream) against long runs of the reference (serial LCG stream) on the same control
files.  Criteria: kT within 0.5%, pressure within 1%, g(r) and bilayer density profiles
within their combined block-averaged error bars."""
import json
import os

import numpy as np
import pytest

from dpd_amd import analysis as A
from dpd_amd import host as H
from helpers import GOLDEN

pytestmark = pytest.mark.gpu
ENS = os.path.join(GOLDEN, "ensemble")


def _run(name, snap_from, snap_every, seed=12345):
    inp = H.Input(os.path.join(ENS, name + ".dmpci"))
    c = inp.counts
    ids, types, pid, pt = inp.beads()
    x, u, v = inp.build_state(seed=seed)
    e = inp.engine(pair_kernel=0)
    e.upload(ids, types, x, v, unpbc=u)
    (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
    if len(h):
        e.upload_topology(h, t, k, l0, fi, se, mo, ph)
    L = np.array(inp.cells, float)
    hist, snaps = [], []
    steps = c["total_time"]
    for s in range(10, steps + 1, 10):
        e.step(10, observe=True)
        o = e.observables()
        hist.append((s, o["kT"], o["pressure"]))
        if s >= snap_from and s % snap_every == 0:
            snaps.append(e.download()[0])
    assert o["jumps"] == 0
    return np.array(hist), snaps, types, L


def _close(name, gpu, ref, rel, nsig=4.0):
    m, se = gpu
    mr, ser = ref
    print(f"  {name}: gpu {m:.5f} +- {se:.5f}  reference {mr:.5f} +- {ser:.5f}", flush=True)
    assert abs(m - mr) / abs(mr) < rel
    assert abs(m - mr) < nsig * np.hypot(se, ser) + 1e-12 or abs(m - mr) / abs(mr) < 0.25 * rel


def test_water_ensemble():
    ref = json.load(open(os.path.join(ENS, "water20.json")))
    hist, snaps, types, L = _run("water20", 1250, 250)
    h = hist[hist[:, 0] > ref["discard"]]
    _close("kT", A.block_average(h[:, 1]), ref["kT"], 0.005)
    _close("P", A.block_average(h[:, 2]), ref["P"], 0.01)
    gs = np.array([A.rdf(x, L)[1] for x in snaps])
    g, gse = gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))
    r = np.array(ref["rdf_r"])
    sel = r > 0.3
    z = np.abs(g - np.array(ref["rdf_g"]))[sel] / np.hypot(gse, np.array(ref["rdf_se"]))[sel]
    print(f"  g(r): max |dg|/sigma {z.max():.2f}, max |dg| {np.abs(g - ref['rdf_g'])[sel].max():.4f}", flush=True)
    assert np.abs(g - np.array(ref["rdf_g"]))[sel].max() < 0.02
    assert np.mean(z) < 2.0


def test_bilayer_ensemble():
    ref = json.load(open(os.path.join(ENS, "bilayer16.json")))
    hist, snaps, types, L = _run("bilayer16", 2250, 250)
    h = hist[hist[:, 0] > ref["discard"]]
    _close("kT", A.block_average(h[:, 1]), ref["kT"], 0.005)
    _close("P", A.block_average(h[:, 2]), ref["P"], 0.01)
    profs = np.array([A.density_profile(x, types, L, centre_type=1, dz=0.5, ntypes=3)[1] for x in snaps])
    p, pse = profs.mean(0), profs.std(0, ddof=1) / np.sqrt(len(profs))
    pr, prse = np.array(ref["profile"]), np.array(ref["profile_se"])
    dev = np.abs(p - pr) / (np.hypot(pse, prse) + 0.02)
    print(f"  density profiles: max dev/sigma {dev.max():.2f}", flush=True)
    for t, name in enumerate(ref["types"]):
        # integrated number in each half-box, and the peak positions, agree
        assert abs(p[t].sum() - pr[t].sum()) / pr[t].sum() < 0.02, name
    assert np.mean(dev) < 2.0
    nlip = ref["nlipids"]
    apl = [A.area_per_lipid(x, types, L, 0, 1, nlip) for x in snaps]
    _close("area per lipid", A.block_average(apl, len(apl)), ref["apl"], 0.03)
