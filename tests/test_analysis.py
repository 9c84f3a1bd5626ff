"""CPU tests of the ensemble estimators in dpd_amd.analysis against closed forms: the mesh
area per lipid of flat and corrugated bilayers (whose mesh areas are known exactly), across the
periodic z face too, and the periodic mid-plane."""
import numpy as np

from dpd_amd import analysis as A


def bilayer(L, heights, per_column=6, nlipids=40, rng=None, z0=None):
    """Tail beads (type 1) in each column of a 4 x 4 mesh at the column's height relative to
    the mid-plane z0, plus head beads (type 0) away from it."""
    rng = rng or np.random.default_rng(0)
    z0 = 0.5 * L[2] if z0 is None else z0
    g = heights.shape[0]
    x = []
    t = []
    for i in range(g):
        for j in range(g):
            for _ in range(per_column):
                px = (i + rng.uniform(0.05, 0.95)) * L[0] / g
                py = (j + rng.uniform(0.05, 0.95)) * L[1] / g
                x.append([px, py, (z0 + heights[i, j]) % L[2]])
                t.append(1)
    for _ in range(20):
        x.append([rng.uniform(0, L[0]), rng.uniform(0, L[1]), (z0 + 2.5) % L[2]])
        t.append(0)
    return np.array(x), np.array(t)


def test_flat_bilayer_area_is_the_box_face():
    L = np.array([12.0, 10.0, 16.0])
    x, t = bilayer(L, np.zeros((4, 4)))
    nl = 40
    assert np.isclose(A.area_per_lipid(x, t, L, 0, 1, nl), L[0] * L[1] / (0.5 * nl), rtol=1e-12)


def test_corrugated_bilayer_area_is_the_tilted_strips():
    """Columns at alternating heights +a, -a along x: every mesh square is a planar strip of
    width dy tilted by the rise 2a over dx, so the mesh area is Lx Ly sqrt(1 + (2a/dx)^2)."""
    L = np.array([12.0, 12.0, 20.0])
    a = 0.4
    h = np.array([[a if i % 2 == 0 else -a for _ in range(4)] for i in range(4)])
    x, t = bilayer(L, h)
    nl = 64
    dx = L[0] / 4
    want = L[0] * L[1] * np.sqrt(1 + (2 * a / dx) ** 2) / (0.5 * nl)
    got = A.area_per_lipid(x, t, L, 0, 1, nl)
    assert np.isclose(got, want, rtol=1e-12)
    # a plausible mistake fails: the projected area alone
    assert not np.isclose(got, L[0] * L[1] / (0.5 * nl), rtol=1e-3)


def test_bilayer_across_the_periodic_face():
    """The same corrugated bilayer with its mid-plane on the z face gives the same area."""
    L = np.array([12.0, 12.0, 20.0])
    a = 0.3
    h = np.array([[a if (i + j) % 2 == 0 else -a for j in range(4)] for i in range(4)])
    x1, t1 = bilayer(L, h, rng=np.random.default_rng(3))
    x2, t2 = bilayer(L, h, rng=np.random.default_rng(3), z0=0.0)
    nl = 50
    assert np.isclose(A.area_per_lipid(x1, t1, L, 0, 1, nl), A.area_per_lipid(x2, t2, L, 0, 1, nl), rtol=1e-9)


def test_midplane_on_the_circle():
    Lz = 20.0
    z = np.concatenate([np.full(10, 19.5), np.full(10, 0.5)])
    assert np.isclose(A.bilayer_midplane(z, Lz) % Lz, 0.0, atol=1e-9) or np.isclose(A.bilayer_midplane(z, Lz), Lz)
    z = np.full(7, 13.25)
    assert np.isclose(A.bilayer_midplane(z, Lz), 13.25)
