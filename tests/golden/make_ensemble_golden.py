#!/usr/bin/env python3
"""Ensemble statistics of thermostatted runs of the compiled reference (oracle/_ref/dpd),
for the statistical-parity tests of the GPU engine (whose random force is a counter-based
pair stream, so trajectories agree only in distribution).

Run in the development container only:
    make -f oracle/ref.mk -j8 && python tests/golden/make_ensemble_golden.py

Cases:
  water20   pure DPD water, 20^3 cells, rho 3 (24,000 beads), a = 25, gamma = 4.5, kT = 1,
            dt = 0.02, lambda = 0.5, 5000 steps from a random start: kT and P every 10
            steps (dmpchs), snapshots every 250 steps for g(r).
  bilayer16 H3(T4)2 lipids (Bond k 128 l0 0.5, BondPair k 20 phi0 0) in water, 16^3
            cells, a_HT 35, a_TW 75, a_HW 35, others 25, started as a lamella
            (State lamella), 4000 steps: kT, P, density profiles along the normal and
            the mesh area per lipid from snapshots every 250 steps.
Statistics use the second half (water: after step 1000, bilayer: after step 2000),
block-averaged; results go to tests/golden/ensemble/<case>.json together with the
control file used.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "osprey-dpd_amd"))
from dpd_amd import analysis as A  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "dpd")
OUT = os.path.join(HERE, "ensemble")

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
RNGSeed		-20261202
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
RNGSeed		-777
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


def water(work):
    run(work, "w", WATER)
    h = history(os.path.join(work, "dmpchs.w"))
    h = h[h[:, 0] > 1000]
    L = np.array([20.0, 20.0, 20.0])
    gs = []
    for t in range(1250, 5001, 250):
        _, x = snapshot(os.path.join(work, f"dmpcrs.w.con.{t}.dat"))
        r, g = A.rdf(x, L)
        gs.append(g)
    gs = np.array(gs)
    res = {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "rdf_r": r.tolist(), "rdf_g": gs.mean(0).tolist(),
           "rdf_se": (gs.std(0, ddof=1) / np.sqrt(len(gs))).tolist(), "nsnap": len(gs),
           "steps": 5000, "discard": 1000}
    return res, WATER


def bilayer(work):
    run(work, "b", BILAYER)
    h = history(os.path.join(work, "dmpchs.b"))
    h = h[h[:, 0] > 2000]
    L = np.array([16.0, 16.0, 16.0])
    profs, apls = [], []
    nlip = None
    for t in range(2250, 4001, 250):
        types, x = snapshot(os.path.join(work, f"dmpcrs.b.con.{t}.dat"))
        z, p = A.density_profile(x, types, L, centre_type=1, dz=0.5, ntypes=3)
        profs.append(p)
        nlip = int((types == 0).sum() // 3)
        apls.append(A.area_per_lipid(x, types, L, 0, 1, nlip))
    profs = np.array(profs)
    res = {"kT": stats(h[:, 1]), "P": stats(h[:, 2]), "z": z.tolist(), "profile": profs.mean(0).tolist(),
           "profile_se": (profs.std(0, ddof=1) / np.sqrt(len(profs))).tolist(),
           "apl": stats(apls, nblocks=len(apls)), "nlipids": nlip, "nsnap": len(profs),
           "types": ["H", "T", "W"], "steps": 4000, "discard": 2000}
    return res, BILAYER


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    os.makedirs(OUT, exist_ok=True)
    work = tempfile.mkdtemp(prefix="dpd_ens_")
    try:
        for name, fn in (("water20", water), ("bilayer16", bilayer)):
            res, text = fn(work)
            with open(os.path.join(OUT, name + ".json"), "w") as fh:
                json.dump(res, fh, indent=1)
            with open(os.path.join(OUT, name + ".dmpci"), "w") as fh:
                fh.write(text)
            print(name, {k: res[k] for k in ("kT", "P")}, flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
