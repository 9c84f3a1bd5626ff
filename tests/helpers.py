"""Shared helpers for the tests: golden-fixture readers and error norms."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

# seeds of the golden cases (|RNGSeed| of tests/golden/<case>/case.dmpci)
SEEDS = {"water6": 4321, "binary6": 977, "chains7": 31337, "lipid8": 8086, "lipidphi8": 5150}


def load_beads(case, name):
    """Bead table of a golden snapshot: dict of polyId, polyType, id, type, x, u, v, f."""
    d = np.loadtxt(os.path.join(GOLDEN, case, name + ".dat.gz"))
    return {"poly": d[:, 0].astype(np.int64), "ptype": d[:, 1].astype(np.int64),
            "id": d[:, 2].astype(np.int32), "type": d[:, 3].astype(np.int32),
            "x": d[:, 4:7].copy(), "u": d[:, 7:10].copy(), "v": d[:, 10:13].copy(),
            "f": d[:, 13:16].copy()}


def load_history(case):
    rows = []
    with open(os.path.join(GOLDEN, case, "th10_history.txt")) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(t) for t in line.split()])
    return np.array(rows)


def restart_clamp(x, L):
    """The reference's restart writer moves coordinates within 0.001 of a box face to
    0.001 (CRestartState::CheckBeadWithinBox, src/RestartState.cpp:492-540)."""
    x = x.copy()
    for d in range(3):
        col = x[:, d]
        col[(col > L[d] - 0.001) | (col < 0.001)] = 0.001
    return x


def six_digit_close(got, ref, rtol=5.01e-6, atol=1e-12):
    """True where `got` agrees with a value printed to 6 significant digits."""
    return np.abs(got - ref) <= rtol * np.abs(ref) + atol


def rel_norm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
