"""Synthetic starting states for the benchmark configurations (host-side set-up, not
timed).  Positions uniform at density rho, chains as random walks with the bond rest
length, Maxwell velocities at kT with the centre-of-mass velocity removed."""
import numpy as np


class Workload:
    def __init__(self, name, cells, a, gamma, ids, types, x, u, v, bonds=None, pairs=None,
                 dt=0.02, seed=20261202):
        self.name, self.cells = name, tuple(cells)
        self.a, self.gamma = np.asarray(a, float), np.asarray(gamma, float)
        self.ids, self.types, self.x, self.u, self.v = ids, types, x, u, v
        self.bonds = bonds      # (head, tail, k, l0)
        self.pairs = pairs      # (first, second, modulus, phi0)
        self.dt, self.seed = dt, seed

    @property
    def n(self):
        return len(self.ids)

    @property
    def nbonds(self):
        return 0 if self.bonds is None else len(self.bonds[0])

    @property
    def nangles(self):
        return 0 if self.pairs is None else len(self.pairs[0])


def _finish(rng, L, u, kT=1.0):
    x = np.mod(u, L)
    x[x >= L] = 0.0
    v = rng.normal(0.0, np.sqrt(kT), size=u.shape)
    v -= v.mean(axis=0)
    return x, v


def water(cells=(20, 20, 20), rho=3, seed=20261202):
    """Pure DPD water: a_WW = 25, gamma = 4.5 (one type)."""
    rng = np.random.default_rng(seed)
    L = np.array(cells, float)
    n = int(round(rho * np.prod(L)))
    u = rng.uniform(0, 1, size=(n, 3)) * L
    x, v = _finish(rng, L, u)
    return Workload("water-%dx%dx%d" % tuple(cells), cells, [[25.0]], [[4.5]],
                    np.arange(1, n + 1, dtype=np.int32), np.zeros(n, np.int32), x, u, v, seed=seed)


def binary(cells=(40, 40, 40), rho=3, a_wo=80.0, seed=20261202):
    """Water/oil 50/50 at random: a_WW = a_OO = 25, a_WO = 80, gamma = 4.5."""
    w = water(cells, rho, seed)
    rng = np.random.default_rng(seed + 1)
    w.types = rng.integers(0, 2, size=w.n).astype(np.int32)
    w.a = np.array([[25.0, a_wo], [a_wo, 25.0]])
    w.gamma = np.full((2, 2), 4.5)
    w.name = "binary-%dx%dx%d" % tuple(cells)
    return w


def condensate(cells=(64, 64, 64), rho=3, length=16, polymer_fraction=0.2, k=128.0, l0=0.5,
               a_sticky=10.0, seed=20261202):
    """Linear 16-mers (S P P P) x 4 at 20% of the beads in water; types W=0, S=1, P=2;
    a_SS = 10, all other a = 25; gamma = 4.5; Hookean bonds k = 128, l0 = 0.5."""
    rng = np.random.default_rng(seed)
    L = np.array(cells, float)
    n = int(round(rho * np.prod(L)))
    nch = int(polymer_fraction * n / length)
    nw = n - nch * length
    steps = rng.normal(size=(nch, length - 1, 3))
    steps *= l0 / np.linalg.norm(steps, axis=2, keepdims=True)
    start = rng.uniform(0, 1, size=(nch, 1, 3)) * L
    chain = np.concatenate([start, start + np.cumsum(steps, axis=1)], axis=1).reshape(-1, 3)
    u = np.concatenate([chain, rng.uniform(0, 1, size=(nw, 3)) * L])
    types = np.concatenate([np.tile(np.where(np.arange(length) % 4 == 0, 1, 2), nch),
                            np.zeros(nw, int)]).astype(np.int32)
    base = (np.arange(nch) * length)[:, None]
    head = (base + np.arange(length - 1)[None, :]).ravel().astype(np.int32)
    tail = head + 1
    x, v = _finish(rng, L, u)
    a = np.full((3, 3), 25.0)
    a[1, 1] = a_sticky
    nb = len(head)
    return Workload("condensate-%dx%dx%d" % tuple(cells), cells, a, np.full((3, 3), 4.5),
                    np.arange(1, n + 1, dtype=np.int32), types, x, u, v,
                    bonds=(head, tail, np.full(nb, k), np.full(nb, l0)), seed=seed)


BILAYER_DMPCI = """dpd

Title	" bilayer "
Date    02/12/26
Comment	" H3(T4)2 lipids in water "

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

Polymer	Water    {fw:.6f}   " (W) "
Polymer	Lipid    {fl:.6f}   " (H H (* (T T T T)) H T T T T) "

Box         {nx} {ny} {nz}       1  1  1
Density		3
Temp        1
RNGSeed		-{seed}
Lambda		0.5
Step		0.02
Time		1000
SamplePeriod     10
AnalysisPeriod	 1000
DensityPeriod    1000
DisplayPeriod    1000
RestartPeriod    1000
Grid		1  1  1
"""


def bilayer(cells=(48, 48, 32), lipid_fraction=0.0201, seed=20261202):
    """H3(T4)2 lipid bilayer in water (a_HT 35, a_TW 75, a_HW 35, others 25; Hookean bonds
    k 128, l0 0.5; bond pairs k 20, phi0 0), started as a lamella at the mid-plane; a
    lipid number fraction of 0.0201 gives ~3.7k lipids in the 48x48x32 box."""
    from .host import Input
    text = BILAYER_DMPCI.format(fw=1.0 - lipid_fraction, fl=lipid_fraction, nx=cells[0], ny=cells[1],
                                nz=cells[2], seed=seed)
    inp = Input(text=text)
    ids, types, pid, pt = inp.beads()
    x, u, v = inp.build_state(seed=seed)
    bonds, pairs = inp.topology()
    return Workload("bilayer-%dx%dx%d" % tuple(cells), cells, inp.a, inp.gamma, ids, types, x, u, v,
                    bonds=bonds, pairs=pairs, seed=seed)
