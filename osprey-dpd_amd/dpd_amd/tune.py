"""Pair-launch schedules chosen by measurement, kept in a JSON file.

dpd_autotune_schedules (include/dpd_engine.h) times every schedule of the pair launch --
the hit-list kernel at 512 or 256 threads per workgroup, 16 to 32 list entries per thread,
with or without a flush threshold, and the full-shell kernel -- on an engine's current state
and keeps the fastest.  This module records each result under the engine's shape (number of
boxes, beads per box as a power-of-two bucket, bead types, whether it has bonds) and applies
a recorded result to an engine of the same shape, so a run does not pay for the timing again.
The file holds one JSON object per line:

    {"nbox": 1, "beads_lo": 524288, "beads_hi": 1048576, "ntypes": 3, "bonded": 1,
     "schedule": 0, "name": "hitlist-t512-c24", "ms_per_step": {...}}

`PYTHONPATH=osprey-dpd_amd python -m dpd_amd.tune [path]` tunes the benchmark
configurations on device 0 and writes them (the default path is osprey-dpd_amd/tune.json,
which bench.py reads).
"""
import json
import os
import sys

import numpy as np

from . import PKG, Engine, schedule_name

DEFAULT_PATH = os.path.join(PKG, "tune.json")


def shape(nbox, beads, ntypes, bonded):
    """The key of an engine shape: beads per box in [2^k, 2^(k+1))."""
    per = max(1, int(beads) // max(1, int(nbox)))
    lo = 1 << (per.bit_length() - 1)
    return {"nbox": int(nbox) if nbox == 1 else 2, "beads_lo": lo, "beads_hi": 2 * lo, "ntypes": int(ntypes),
            "bonded": int(bool(bonded))}


def _matches(entry, key):
    return all(entry.get(k) == v for k, v in key.items())


def load(path=DEFAULT_PATH):
    if not os.path.exists(path):
        return []
    with open(path) as fh:
        return [json.loads(line) for line in fh if line.strip()]


def save(entries, path=DEFAULT_PATH):
    keys = ("nbox", "beads_lo", "beads_hi", "ntypes", "bonded", "schedule", "name", "ms_per_step", "workload")
    with open(path, "w") as fh:
        for e in sorted(entries, key=lambda e: (e["nbox"], e["beads_lo"], e["ntypes"], e["bonded"])):
            fh.write(json.dumps({k: e[k] for k in keys if k in e}) + "\n")


def lookup(nbox, beads, ntypes, bonded, path=DEFAULT_PATH):
    """The recorded schedule for this shape, or None."""
    key = shape(nbox, beads, ntypes, bonded)
    for e in load(path):
        if _matches(e, key):
            return int(e["schedule"])
    return None


def apply(engine, path=DEFAULT_PATH):
    """Set the schedule recorded for `engine`'s shape (the engine reads the file itself,
    dpd_apply_tuning); returns it (None: no record, or no file)."""
    if not os.path.exists(path):
        return None
    return engine.apply_tuning(path)


def tune(engine, nbox, beads, ntypes, bonded, nsteps=20, path=DEFAULT_PATH, workload=None):
    """Time every schedule on `engine`'s current state, keep the fastest on it and record it
    under the engine's shape (replacing an older record of the same shape)."""
    chosen, ms = engine.autotune_schedules(nsteps)
    key = shape(nbox, beads, ntypes, bonded)
    entries = [e for e in load(path) if not _matches(e, key)]
    entry = dict(key, schedule=int(chosen), name=schedule_name(chosen),
                 ms_per_step={k: round(v, 5) for k, v in ms.items() if v > 0})
    if workload:
        entry["workload"] = workload
    entries.append(entry)
    save(entries, path)
    return chosen, ms


def main(argv):
    """Tune the benchmark configurations: the condensate, water, binary and bilayer boxes and
    the replica batch, each after a relaxation of its synthetic start."""
    path = argv[1] if len(argv) > 1 else DEFAULT_PATH
    from . import surrogate as S
    from . import workloads as W
    for wl in (W.condensate(), W.water(), W.binary(), W.bilayer()):
        e = Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed)
        e.upload(wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u)
        if wl.bonds is not None:
            h, t, k, l0 = wl.bonds
            fi, se, mo, ph = wl.pairs if wl.pairs is not None else ((), (), (), ())
            e.upload_topology(h, t, k, l0, fi, se, mo, ph)
        for _ in range(4):
            e.step(50, graph=True)
            e.renormalise_momenta()
        chosen, ms = tune(e, 1, wl.n, len(wl.a), wl.nbonds + wl.nangles > 0, path=path, workload=wl.name)
        print(f"{wl.name}: {schedule_name(chosen)}  " + ", ".join(f"{k} {v:.4f}" for k, v in ms.items() if v > 0),
              flush=True)
        e.close()
    rng = np.random.default_rng(20261202)
    params = S.sample_params(256, rng)
    sizes = rng.integers(8, 25, size=256)
    seeds = rng.integers(1, 2 ** 31, size=256)
    e, _ = S.build_batch(params, sizes, seeds)
    e.step(100, graph=True)
    chosen, ms = tune(e, 256, e.n, 3, True, path=path, workload="256 replicas")
    print(f"256 replicas: {schedule_name(chosen)}  " + ", ".join(f"{k} {v:.4f}" for k, v in ms.items() if v > 0),
          flush=True)


if __name__ == "__main__":
    main(sys.argv)
