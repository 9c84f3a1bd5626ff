"""Ensemble observables for statistical parity: block averages, radial distribution
functions and density profiles along a bilayer normal.

The reference's own RDF process (prPairCorrelationFunction, src/prPairCorrelationFunction.cpp:
211-331) is O(N^2) and unnormalised; these estimators are normalised (g -> 1 at large r)
and use a periodic k-d tree.
"""
import numpy as np
from scipy.spatial import cKDTree


def block_average(series, nblocks=10):
    """Mean and standard error from `nblocks` contiguous blocks (samples are correlated)."""
    s = np.asarray(series, dtype=float)
    nb = max(2, min(nblocks, len(s)))
    blocks = np.array([b.mean() for b in np.array_split(s, nb)])
    return float(s.mean()), float(blocks.std(ddof=1) / np.sqrt(nb))


def rdf(x, L, rmax=2.5, dr=0.05, types=None, pair=None):
    """g(r) of one configuration in a periodic box of lengths L.  pair = (ta, tb) restricts
    to pairs of those types (ordered-pair counts, both directions)."""
    x = np.mod(np.asarray(x, dtype=float), L)
    L = np.asarray(L, dtype=float)
    edges = np.arange(0.0, rmax + 0.5 * dr, dr)
    if pair is None:
        a_idx = b_idx = np.arange(len(x))
    else:
        a_idx = np.where(types == pair[0])[0]
        b_idx = np.where(types == pair[1])[0]
    ta = cKDTree(x[a_idx], boxsize=L)
    tb = ta if pair is None or pair[0] == pair[1] else cKDTree(x[b_idx], boxsize=L)
    pairs = ta.query_pairs(rmax, output_type="ndarray") if tb is ta else None
    if pairs is not None:
        d = x[a_idx[pairs[:, 0]]] - x[a_idx[pairs[:, 1]]]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        hist = 2.0 * np.histogram(r, edges)[0]
        na, nb = len(a_idx), len(a_idx) - 1
    else:
        sp = ta.sparse_distance_matrix(tb, rmax, output_type="ndarray")
        hist = np.histogram(sp["v"], edges)[0].astype(float)
        na, nb = len(a_idx), len(b_idx)
    V = np.prod(L)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    g = hist / (na * nb / V * shell)
    return 0.5 * (edges[1:] + edges[:-1]), g


def bilayer_midplane(z, Lz):
    """Centre of mass of tail beads along z on the circle (periodic)."""
    th = 2 * np.pi * np.asarray(z) / Lz
    ang = np.arctan2(np.sin(th).mean(), np.cos(th).mean())
    return (ang % (2 * np.pi)) * Lz / (2 * np.pi)


def density_profile(x, types, L, centre_type, nbins=None, dz=0.5, ntypes=None):
    """Number density of each type along z, centred on the mid-plane of `centre_type`
    beads; returns bin centres in [-Lz/2, Lz/2) and rho[type][bin]."""
    x = np.asarray(x, dtype=float)
    Lz = float(L[2])
    zc = bilayer_midplane(x[types == centre_type, 2], Lz)
    z = np.mod(x[:, 2] - zc + 0.5 * Lz, Lz) - 0.5 * Lz
    nb = nbins or int(round(Lz / dz))
    edges = np.linspace(-0.5 * Lz, 0.5 * Lz, nb + 1)
    nt = ntypes or int(types.max()) + 1
    area = L[0] * L[1]
    prof = np.array([np.histogram(z[types == t], edges)[0] / (area * (edges[1] - edges[0])) for t in range(nt)])
    return 0.5 * (edges[1:] + edges[:-1]), prof


def area_per_lipid(x, types, L, head_type, tail_type, nlipids, grid=4):
    """Undulation-corrected area per lipid of a bilayer with normal z: the mid-surface
    height h(x, y) is the mean z of tail beads per column of a grid x grid mesh; its area
    (two triangles per mesh square, periodic) over the lipids per leaflet.  Follows the
    mesh construction of the reference's bilayer analysis (src/Bilayer.cpp:1682)."""
    x = np.asarray(x, dtype=float)
    Lz = float(L[2])
    tails = x[types == tail_type]
    zc = bilayer_midplane(tails[:, 2], Lz)
    z = np.mod(tails[:, 2] - zc + 0.5 * Lz, Lz) - 0.5 * Lz
    ix = np.minimum((tails[:, 0] / L[0] * grid).astype(int), grid - 1)
    iy = np.minimum((tails[:, 1] / L[1] * grid).astype(int), grid - 1)
    h = np.zeros((grid, grid))
    for i in range(grid):
        for j in range(grid):
            m = (ix == i) & (iy == j)
            h[i, j] = z[m].mean() if m.any() else 0.0
    dx, dy = L[0] / grid, L[1] / grid
    area = 0.0
    for i in range(grid):
        for j in range(grid):
            p00 = np.array([0.0, 0.0, h[i, j]])
            p10 = np.array([dx, 0.0, h[(i + 1) % grid, j]])
            p01 = np.array([0.0, dy, h[i, (j + 1) % grid]])
            p11 = np.array([dx, dy, h[(i + 1) % grid, (j + 1) % grid]])
            area += 0.5 * np.linalg.norm(np.cross(p10 - p00, p11 - p00))
            area += 0.5 * np.linalg.norm(np.cross(p11 - p00, p01 - p00))
    return area / (0.5 * nlipids)


def bond_type_slices(bond_tab, bond_types, ntypes):
    """The reference's bond slice-stress table, keyed by bond type, from a table keyed by
    (head type, tail type) such as dpd_slice_stress returns.  bond_types lists (head type,
    tail type) per Bond declaration (Input.bond_types()); a bond's type is the declaration
    that names its two bead types in either order (src/InputData.cpp:556-598, :1489-1498),
    so its entry is the sum of the (h, t) and (t, h) entries."""
    bond_tab = np.asarray(bond_tab)
    out = np.zeros((bond_tab.shape[0], len(bond_types), bond_tab.shape[2]))
    for b, (h, t) in enumerate(bond_types):
        out[:, b] = bond_tab[:, h * ntypes + t]
        if h != t:
            out[:, b] += bond_tab[:, t * ntypes + h]
    return out
