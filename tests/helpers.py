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


def random_state(cells, n, ntypes, seed, widths=(1.0, 1.0, 1.0), kT=1.0):
    """Uniform positions in the box, Maxwell velocities (COM removed), random types."""
    rng = np.random.default_rng(seed)
    L = np.array(cells, dtype=float) * np.array(widths)
    x = rng.uniform(0, 1, size=(n, 3)) * L
    v = rng.normal(0, np.sqrt(kT), size=(n, 3))
    v -= v.mean(axis=0)
    t = rng.integers(0, ntypes, size=n).astype(np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    return ids, t, x, v


def chain_state(cells, nchains, length, nsolvent, seed, bond=0.5):
    """Linear chains as random walks (step `bond`) plus solvent; returns ids, types (0 =
    solvent, 1 = chain end/sticker, 2 = chain interior), wrapped x, unwrapped u, v and the
    topology: bonds (head = earlier bead), bond pairs (first = later bond) as in the
    reference's polymer construction (src/InputData.cpp:1114-1164, :1469-1513)."""
    rng = np.random.default_rng(seed)
    L = np.array(cells, dtype=float)
    u = []
    types = []
    head, tail, first, second = [], [], [], []
    for c in range(nchains):
        start = rng.uniform(0, 1, size=3) * L
        base = len(u)
        p = start.copy()
        for k in range(length):
            if k:
                d = rng.normal(size=3)
                p = p + bond * d / np.linalg.norm(d)
            u.append(p.copy())
            types.append(1 if k % 4 == 0 else 2)
            if k:
                head.append(base + k - 1)
                tail.append(base + k)
            if k >= 2:
                # bonds A = (k-2 -> k-1) index b-1, B = (k-1 -> k) index b: first = B, second = A
                b = len(head) - 1
                first.append(b)
                second.append(b - 1)
    for s in range(nsolvent):
        u.append(rng.uniform(0, 1, size=3) * L)
        types.append(0)
    u = np.array(u)
    x = np.mod(u, L)
    x[x >= L] = 0.0
    n = len(u)
    v = rng.normal(size=(n, 3))
    v -= v.mean(axis=0)
    ids = np.arange(1, n + 1, dtype=np.int32)
    return ids, np.array(types, dtype=np.int32), x, u, v, (head, tail, first, second)


def force_errors(f, ref, floor=1.0):
    """Normwise relative error and per-bead max |dF_i| / max(|F_i|, floor)."""
    norm = rel_norm(f, ref)
    per = np.linalg.norm(f - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), floor)
    return norm, float(per.max())


def clamped_and_bonded(d, L, head=(), tail=(), first=(), second=()):
    """Beads whose wrapped and unwrapped coordinates of a reference restart snapshot
    disagree (its writer moves a coordinate within 0.001 of a face to 0.001 and leaves the
    unwrapped one, src/RestartState.cpp:492-540), together with every bead that shares a bond
    or a bond pair with one: the reference computes pair forces from the one and bonds from
    the other, which no single consistent position reproduces."""
    img = (d["u"] - d["x"]) / np.asarray(L, float)
    bad = ~(np.abs(img - np.round(img)) < 1e-5).all(axis=1)
    head, tail = np.asarray(head, int), np.asarray(tail, int)
    first, second = np.asarray(first, int), np.asarray(second, int)
    out = bad.copy()
    if len(head):
        hit = bad[head] | bad[tail]
        out[head[hit]] = True
        out[tail[hit]] = True
    if len(first):
        a, b, c = head[second], head[first], tail[first]
        hit = bad[a] | bad[b] | bad[c]
        out[a[hit]] = out[b[hit]] = out[c[hit]] = True
    return out


def consistent(d, L):
    """The snapshot with its wrapped coordinates taken from the unwrapped ones."""
    L = np.asarray(L, float)
    x = np.mod(d["u"], L)
    x[x >= L] = 0.0
    return dict(d, x=x)
