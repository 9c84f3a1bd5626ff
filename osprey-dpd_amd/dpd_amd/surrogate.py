"""Batched replica runs and the surrogate that screens parameter sweeps.

A replica is a cubic water/lipid box (H3(T4)2 lipids started as a bilayer at the
mid-plane, State lamella) whose run parameters are six interaction and composition
features -- a_HW, a_TW, a_HT, a_TT (a_HH = a_WW = 25), the lipid bead fraction and the
Hookean bond stiffness -- and the box side L ~ U{8..24}.  Hundreds of replicas of uneven
size advance together in one batched engine; the time-averaged pressure and the mesh area
per lipid of each replica are the targets.  The surrogate (dpd_amd.mlp.Surrogate, BF16
matrix cores) takes the six parameters and the box side: in a periodic box of fixed size
the bilayer's area per lipid is set by the box (about 2 L^2 / lipids, lipids ~ L^3) and the
pressure carries the bilayer's tension over the box height, so a map from the six
parameters alone leaves scatter no training removes.  It is trained on those runs and
screens candidate parameter sets; the simulation then checks every accepted candidate.
"""
import time

import numpy as np

from . import PAIR_AUTO
from . import analysis as A
from .host import Input, replica_batch
from .mlp import Surrogate

PARAMS = ("a_HW", "a_TW", "a_HT", "a_TT", "lipid_fraction", "k_bond")
FEATURES = PARAMS + ("box_side",)
LOW = np.array([20.0, 20.0, 20.0, 20.0, 0.10, 64.0])
HIGH = np.array([90.0, 90.0, 90.0, 90.0, 0.35, 256.0])

TEMPLATE = """dpd

Title	" replica "
Date    02/12/26
Comment	" water/lipid replica "

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
      {aHT}  {aTT}
      4.5 4.5

Bead  W
      0.5
      {aHW}  {aTW}  25
      4.5 4.5 4.5

Bond    H H  {k}  0.5
Bond    H T  {k}  0.5
Bond    T T  {k}  0.5

BondPair  H T T  20.0  0.0
BondPair  T T T  20.0  0.0

Polymer	Water    {fw:.8f}   " (W) "
Polymer	Lipid    {fl:.8f}   " (H H (* (T T T T)) H T T T T) "

Box         {L} {L} {L}       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		{dt}
Time		{steps}
SamplePeriod     10
AnalysisPeriod	 {steps}
DensityPeriod    {steps}
DisplayPeriod    {steps}
RestartPeriod    {steps}
Grid		1  1  1
"""


def sample_params(n, rng):
    return LOW + (HIGH - LOW) * rng.uniform(size=(n, len(LOW)))


def sample_sizes(n, rng, lo=8, hi=24):
    """Box sides L ~ U{lo..hi}"""
    return rng.integers(lo, hi + 1, size=n)


def features(params, sizes):
    """The surrogate's inputs: the six run parameters and the box side"""
    return np.column_stack([np.asarray(params, float), np.asarray(sizes, float)])


def replica_text(p, L, seed, steps=5000, dt=0.02):
    aHW, aTW, aHT, aTT, c, k = p
    fl = c / (11.0 - 10.0 * c)          # lipid number fraction giving bead fraction c
    return TEMPLATE.format(aHT=aHT, aTT=aTT, aHW=aHW, aTW=aTW, k=k, fw=1.0 - fl, fl=fl, L=int(L), seed=int(seed),
                           dt=dt, steps=steps)


def build_batch(params, sizes, seeds, device=0, pair_kernel=PAIR_AUTO, steps=5000):
    """One batched engine holding a replica per parameter row."""
    inputs = [Input(text=replica_text(p, L, s, steps)) for p, L, s in zip(params, sizes, seeds)]
    e, boxes = replica_batch(inputs, [int(s) for s in seeds], device=device, pair_kernel=pair_kernel)
    meta = [{"types": t, "L": L, "nlipids": int((pt == 1).sum() // 11)} for t, pt, L in boxes]
    return e, meta


def run_replicas(params, sizes, seeds, steps=5000, discard=None, sample_every=10, device=0,
                 pair_kernel=PAIR_AUTO, graph=True):
    """Advance all replicas together; returns pressure (mean, s.e.), area per lipid, kT,
    bead-steps/s of the batch and its wall time."""
    discard = steps // 2 if discard is None else discard
    e, meta = build_batch(params, sizes, seeds, device, pair_kernel, steps)
    nb = len(params)
    P = [[] for _ in range(nb)]
    kT = [[] for _ in range(nb)]
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(sample_every, steps - done)
        e.step(k, observe=True, graph=graph)
        done += k
        if done > discard:
            for b in range(nb):
                o = e.observables(b)
                P[b].append(o["pressure"])
                kT[b].append(o["kT"])
    wall = time.perf_counter() - t0
    apl = []
    for b in range(nb):
        x = e.download_box(b)[0]
        m = meta[b]
        apl.append(A.area_per_lipid(x, m["types"], m["L"], 0, 1, m["nlipids"]) if m["nlipids"] else np.nan)
    stats = [A.block_average(p, 5) for p in P]
    return {"P": np.array([s[0] for s in stats]), "P_se": np.array([s[1] for s in stats]),
            "kT": np.array([np.mean(t) for t in kT]), "apl": np.array(apl),
            "bead_steps_per_s": e.n * steps / wall, "wall": wall, "nbeads": e.n}


def train(X, Y, iters=3000, lr=2e-3, seed=1, hidden=256):
    net = Surrogate(X.shape[1], hidden, Y.shape[1], seed=seed)
    net.fit_scaling(X, Y)
    losses = net.train(X, Y, iters=iters, lr=lr)
    return net, losses


def screen(net, candidates, target, tol):
    """Candidates (rows of features()) whose predicted pressure lies within `tol`
    (relative) of `target`."""
    pred = net.predict(candidates)
    keep = np.abs(pred[:, 0] - target) <= tol * target
    return np.where(keep)[0], pred
