#!/usr/bin/env python3
"""Ensemble statistics of thermostatted runs of the compiled reference (oracle/_ref/dpd),
for the statistical-parity tests of the GPU engine (whose random force is a counter-based
pair stream, so trajectories agree only in distribution).

Run in the development container only:
    make -f oracle/ref.mk -j8 && python tests/golden/make_ensemble_golden.py [case ...]

Cases:
  water20   pure DPD water, 20^3 cells, rho 3 (24,000 beads), a = 25, gamma = 4.5, kT = 1,
            dt = 0.02, lambda = 0.5, 5000 steps from a random start: kT and P every 10
            steps (dmpchs), snapshots every 250 steps for g(r).
  bilayer16 H3(T4)2 lipids (Bond k 128 l0 0.5, BondPair k 20 phi0 0) in water, 16^3
            cells, a_HT 35, a_TW 75, a_HW 35, others 25, started as a lamella by the
            reference's own builder (State lamella; the engine's builder reproduces the
            same states), 4000 steps: kT, P, density profiles along the normal and the
            mesh area per lipid from snapshots every 250 steps.
  binary20  water/oil 50/50 at random (the binary demixing configuration on a 20^3 box,
            24,000 beads), a_WW = a_OO = 25, a_WO = 80, 4000 steps from a random start:
            kT and P, and the W-W and W-O g(r) of snapshots every 250 steps after step 2000
            while the mixture demixes.
  condensate20  the condensate composition on a 20^3 box (24,000 beads): 16-mers
            (S P P P)x4 at 20% of the beads in water, a_SS = 10, other a = 25, Hookean bonds
            k 128 l0 0.5, 4000 steps from the package's random states (its restatement of
            the reference's random builder, src/RandomBuilder.cpp, handed over as State
            restart, as for the bilayer): kT, P, the sticker-sticker g(r) and the mean bond
            length from snapshots every 250 steps after step 2000.
Each case runs NRUNS times from independent starting states (RNGSeed differs per run).
The condensate runs start from the package's random states handed to the reference as
restart snapshots (State restart); the water, binary and bilayer runs from the reference's
own builders, whose states the package reproduces from the same seeds.
Per run, kT and P are block-averaged over the second part of the run (water: after step
1000, bilayer: after step 2000); the fixtures keep every run's mean so that comparisons
use the run-to-run spread, which for a bilayer relaxing over a few thousand steps is
larger than a single run's block error.  g(r), density profiles and the area per lipid
pool the snapshots of all runs.  Results go to tests/golden/ensemble/<case>.json together
with the control file of the first run.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "osprey-dpd_amd"))
from dpd_amd import analysis as A  # noqa: E402
from dpd_amd import host as H  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "dpd")
OUT = os.path.join(HERE, "ensemble")
NRUNS = 4
WATER_SEEDS = [20261202 + 7919 * r for r in range(NRUNS)]
BILAYER_SEEDS = [777 + 7919 * r for r in range(NRUNS)]

WATER = """dpd

Title	" water 20 "
Date    02/12/26
Comment	" ensemble statistics "

State	random

Bead  W
      0.5
      25
      4.5

Polymer	Water    1.0   " (W) "

Box         20 20 20       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		0.02
Time		5000
SamplePeriod     10
AnalysisPeriod	 1000
DensityPeriod    1000
DisplayPeriod    5000
RestartPeriod    250
Grid		1  1  1
"""

BINARY_SEEDS = [3131 + 7919 * r for r in range(NRUNS)]

BINARY = """dpd

Title	" binary 20 "
Date    02/12/26
Comment	" ensemble statistics "

State	random

Bead  W
      0.5
      25
      4.5

Bead  O
      0.5
      80  25
      4.5 4.5

Polymer	Water    0.5   " (W) "
Polymer	Oil      0.5   " (O) "

Box         20 20 20       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		0.02
Time		4000
SamplePeriod     10
AnalysisPeriod	 1000
DensityPeriod    1000
DisplayPeriod    4000
RestartPeriod    250
Grid		1  1  1
"""

CONDENSATE_SEEDS = [4545 + 7919 * r for r in range(NRUNS)]

CONDENSATE = """dpd

Title	" condensate 20 "
Date    02/12/26
Comment	" ensemble statistics "

State	random

Bead  W
      0.5
      25
      4.5

Bead  S
      0.5
      25  10
      4.5 4.5

Bead  P
      0.5
      25  25  25
      4.5 4.5 4.5

Bond    S P  128  0.5
Bond    P P  128  0.5

Polymer	Water    0.984615   " (W) "
Polymer	Chain    0.015385   " (S P P P S P P P S P P P S P P P) "

Box         20 20 20       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		0.02
Time		4000
SamplePeriod     10
AnalysisPeriod	 1000
DensityPeriod    1000
DisplayPeriod    4000
RestartPeriod    250
Grid		1  1  1
"""

BILAYER = """dpd

Title	" bilayer 16 "
Date    02/12/26
Comment	" ensemble statistics "

State	lamella
        Polymer			Lipid
        Normal			0 0 1
        Centre			0.5
        Thickness		5.0
        Linearise		1
        UpperFraction	0.5
        Polymerise		0

Bead  H
      0.5
      25
      4.5

Bead  T
      0.5
      35  25
      4.5 4.5

Bead  W
      0.5
      35  75  25
      4.5 4.5 4.5

Bond    H H  128  0.5
Bond    H T  128  0.5
Bond    T T  128  0.5

BondPair  H T T  20.0  0.0
BondPair  T T T  20.0  0.0

Polymer	Water    0.95333   " (W) "
Polymer	Lipid    0.04667   " (H H (* (T T T T)) H T T T T) "

Box         16 16 16       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		0.02
Time		4000
SamplePeriod     10
AnalysisPeriod	 1000
DensityPeriod    1000
DisplayPeriod    4000
RestartPeriod    250
Grid		1  1  1
"""


def run(work, run_id, text):
    with open(os.path.join(work, "dmpci." + run_id), "w") as fh:
        fh.write(text)
    subprocess.run([REF, run_id], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                   timeout=3600, check=True)


def history(path):
    rows = [l.split() for l in open(path) if l.strip()]
    return np.array([[float(r[0]), float(r[1]), float(r[2])] for r in rows])


def snapshot(path):
    rows = []
    with open(path) as fh:
        fh.readline()
        for line in fh:
            t = line.split()
            if not t or t[0] == "inclusive":
                break
            rows.append([float(v) for v in t[:8]])
    d = np.array(rows)
    return d[:, 3].astype(int), d[:, 5:8]


def stats(series, nblocks=10):
    m, se = A.block_average(series, nblocks)
    return [m, se]


def water_run(work, r):
    rid = f"w{r}"
    run(work, rid, WATER.format(seed=WATER_SEEDS[r]))
    h = history(os.path.join(work, "dmpchs." + rid))
    h = h[h[:, 0] > 1000]
    L = np.array([20.0, 20.0, 20.0])
    gs = []
    for t in range(1250, 5001, 250):
        _, x = snapshot(os.path.join(work, f"dmpcrs.{rid}.con.{t}.dat"))
        r_, g = A.rdf(x, L)
        gs.append(g)
    return {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "r": r_, "g": gs}


def binary_run(work, r):
    rid = f"o{r}"
    run(work, rid, BINARY.format(seed=BINARY_SEEDS[r]))
    h = history(os.path.join(work, "dmpchs." + rid))
    h = h[h[:, 0] > 1000]
    L = np.array([20.0, 20.0, 20.0])
    gww, gwo = [], []
    for t in range(2250, 4001, 250):
        types, x = snapshot(os.path.join(work, f"dmpcrs.{rid}.con.{t}.dat"))
        r_, g = A.rdf(x, L, types=types, pair=(0, 0))
        gww.append(g)
        gwo.append(A.rdf(x, L, types=types, pair=(0, 1))[1])
    return {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "r": r_, "gww": gww, "gwo": gwo}


def binary(runs):
    gww, gwwse = pooled([g for r in runs for g in r["gww"]])
    gwo, gwose = pooled([g for r in runs for g in r["gwo"]])
    return {"runs": [{"kT": r["kT"], "P": r["P"]} for r in runs], "seeds": BINARY_SEEDS,
            "rdf_r": runs[0]["r"].tolist(), "g_ww": gww, "g_ww_se": gwwse, "g_wo": gwo, "g_wo_se": gwose,
            "nsnap": sum(len(r["gww"]) for r in runs), "steps": 4000, "discard": 1000,
            "snap_from": 2250, "snap_every": 250}, BINARY.format(seed=BINARY_SEEDS[0])


def condensate_run(work, r):
    rid = f"c{r}"
    text = CONDENSATE.format(seed=CONDENSATE_SEEDS[r])
    own_start(work, f"cown{r}", text, CONDENSATE_SEEDS[r])
    run(work, rid, restart_text(text, f"cown{r}"))
    h = history(os.path.join(work, "dmpchs." + rid))
    h = h[h[:, 0] > 1000]
    L = np.array([20.0, 20.0, 20.0])
    (bh, bt, _, _), _ = H.Input(text=text).topology()
    gss, blen = [], []
    for t in range(2250, 4001, 250):
        types, x = snapshot(os.path.join(work, f"dmpcrs.{rid}.con.{t}.dat"))
        r_, g = A.rdf(x, L, types=types, pair=(1, 1))
        gss.append(g)
        u = unwrapped(os.path.join(work, f"dmpcrs.{rid}.con.{t}.dat"))
        blen.append(float(np.linalg.norm(u[bh] - u[bt], axis=1).mean()))
    return {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "r": r_, "gss": gss, "blen": blen}


def unwrapped(path):
    """Unwrapped positions in bead-id order (ids 1..N are the creation order the topology
    indexes)."""
    rows = []
    with open(path) as fh:
        fh.readline()
        for line in fh:
            t = line.split()
            if not t or t[0] == "inclusive":
                break
            rows.append([float(t[2])] + [float(v) for v in t[8:11]])
    d = np.array(rows)
    return d[np.argsort(d[:, 0]), 1:]


def condensate(runs):
    gss, gsse = pooled([g for r in runs for g in r["gss"]])
    return {"runs": [{"kT": r["kT"], "P": r["P"], "bond_length": float(np.mean(r["blen"]))} for r in runs],
            "seeds": CONDENSATE_SEEDS, "start": "dpd_amd random builder, seed per run",
            "rdf_r": runs[0]["r"].tolist(), "g_ss": gss, "g_ss_se": gsse,
            "nsnap": sum(len(r["gss"]) for r in runs), "steps": 4000, "discard": 1000,
            "snap_from": 2250, "snap_every": 250}, CONDENSATE.format(seed=CONDENSATE_SEEDS[0])


def restart_text(text, run_id):
    """The control file with its State block replaced by a restart from run_id, step 0."""
    out, skip = [], False
    for line in text.split("\n"):
        if line.startswith("State"):
            out.append(f"State\trestart RunId {run_id} StateId 0")
            skip = True
            continue
        if skip and line[:1] in (" ", "\t") and line.strip():
            continue
        skip = False
        out.append(line)
    return "\n".join(out)


def own_start(work, run_id, text, seed):
    """Write the package's starting state for `text` (seed) as the restart snapshot of
    run_id at step 0, next to the control file the reference checks a restart against."""
    inp = H.Input(text=text)
    ids, types, pid, pt = inp.beads()
    x, u, v = inp.build_state(seed=seed)
    L = np.array(inp.cells, float) * np.array(inp.widths)
    H.write_restart(os.path.join(work, f"dmpcrs.{run_id}.con.0.dat"), pid, pt, ids, types, L, x, u, v,
                    np.zeros_like(x))
    with open(os.path.join(work, "dmpci." + run_id), "w") as fh:
        fh.write(text)


def bilayer_run(work, r):
    rid = f"b{r}"
    text = BILAYER.format(seed=BILAYER_SEEDS[r])
    run(work, rid, text)
    h = history(os.path.join(work, "dmpchs." + rid))
    h = h[h[:, 0] > 2000]
    L = np.array([16.0, 16.0, 16.0])
    profs, apls = [], []
    for t in range(2250, 4001, 250):
        types, x = snapshot(os.path.join(work, f"dmpcrs.{rid}.con.{t}.dat"))
        z, p = A.density_profile(x, types, L, centre_type=1, dz=0.5, ntypes=3)
        profs.append(p)
        nlip = int((types == 0).sum() // 3)
        apls.append(A.area_per_lipid(x, types, L, 0, 1, nlip))
    return {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "z": z, "prof": profs, "apl": apls, "nlip": nlip}


def pooled(samples):
    a = np.array(samples)
    return a.mean(0).tolist(), (a.std(0, ddof=1) / np.sqrt(len(a))).tolist()


def water(runs):
    g, gse = pooled([g for r in runs for g in r["g"]])
    return {"runs": [{"kT": r["kT"], "P": r["P"]} for r in runs], "seeds": WATER_SEEDS,
            "rdf_r": runs[0]["r"].tolist(), "rdf_g": g, "rdf_se": gse, "nsnap": sum(len(r["g"]) for r in runs),
            "steps": 5000, "discard": 1000}, WATER.format(seed=WATER_SEEDS[0])


def bilayer(runs):
    p, pse = pooled([p for r in runs for p in r["prof"]])
    apl = [a for r in runs for a in r["apl"]]
    return {"runs": [{"kT": r["kT"], "P": r["P"], "apl": float(np.mean(r["apl"]))} for r in runs],
            "seeds": BILAYER_SEEDS, "start": "the reference's lamella builder (State lamella), seed per run",
            "z": runs[0]["z"].tolist(), "profile": p, "profile_se": pse,
            "apl": [float(np.mean(apl)), float(np.std(apl, ddof=1) / np.sqrt(len(apl)))],
            "nlipids": runs[0]["nlip"], "nsnap": len(apl), "types": ["H", "T", "W"], "steps": 4000,
            "discard": 2000}, BILAYER.format(seed=BILAYER_SEEDS[0])


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    wanted = set(sys.argv[1:]) or {"water20", "bilayer16", "binary20", "condensate20"}
    os.makedirs(OUT, exist_ok=True)
    work = tempfile.mkdtemp(prefix="dpd_ens_")
    try:
        with ThreadPoolExecutor(max_workers=8) as pool:
            jobs = {}
            if "water20" in wanted:
                jobs["water20"] = ([pool.submit(water_run, work, r) for r in range(NRUNS)],)
            if "bilayer16" in wanted:
                jobs["bilayer16"] = ([pool.submit(bilayer_run, work, r) for r in range(NRUNS)],)
            if "binary20" in wanted:
                jobs["binary20"] = ([pool.submit(binary_run, work, r) for r in range(NRUNS)],)
            if "condensate20" in wanted:
                jobs["condensate20"] = ([pool.submit(condensate_run, work, r) for r in range(NRUNS)],)
            make = {"water20": water, "bilayer16": bilayer, "binary20": binary, "condensate20": condensate}
            cases = [(name, make[name](*[[f.result() for f in fs] for fs in futs])) for name, futs in jobs.items()]
        for name, (res, text) in cases:
            with open(os.path.join(OUT, name + ".json"), "w") as fh:
                json.dump(res, fh, indent=1)
            with open(os.path.join(OUT, name + ".dmpci"), "w") as fh:
                fh.write(text)
            print(name, [(r["kT"][0], r["P"][0]) for r in res["runs"]], flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
