"""CPU tests of the dmpci reader and polymer expansion (the product's host layer) against
the counts the compiled reference creates for the same control files."""
import json
import os

import numpy as np
import pytest

from dpd_amd import host as H
from helpers import GOLDEN

PARSER = os.path.join(GOLDEN, "parser")
_counts = json.load(open(os.path.join(PARSER, "counts.json"))) if os.path.exists(
    os.path.join(PARSER, "counts.json")) else {}


@pytest.mark.parametrize("case", ["water6", "binary6", "chains7", "lipid8", "lipidphi8"])
def test_golden_case_counts(case):
    inp = H.Input(os.path.join(GOLDEN, case, "case.dmpci"))
    exp = json.load(open(os.path.join(GOLDEN, case, "counts.json")))
    c = inp.counts
    assert {"polymers": c["npolymers"], "beads": c["nbeads"], "bonds": c["nbonds"],
            "bondpairs": c["npairs"]} == exp


@pytest.mark.parametrize("name", sorted(k for k, v in _counts.items() if v is not None))
def test_reference_example_counts(name):
    """Every example control file the reference ships (incl. the Biology-2023 set)."""
    path = os.path.join(PARSER, name + ".dmpci")
    exp = _counts[name]
    try:
        inp = H.Input(path)
    except H.InputError as e:
        assert "Wall" in str(e) or "Gravity" in str(e), str(e)
        pytest.skip("Wall/Gravity inputs are outside the step engine")
    c = inp.counts
    assert {"polymers": c["npolymers"], "beads": c["nbeads"], "bonds": c["nbonds"],
            "bondpairs": c["npairs"]} == exp


def test_lipid_topology_shape():
    """(H H (* (T T T T)) H T T T T): 11 beads, the branch hangs off bead 2, bond heads are
    the earlier beads, bond pairs store (later bond, earlier bond)."""
    text = open(os.path.join(GOLDEN, "lipid8", "case.dmpci")).read()
    inp = H.Input(text=text)
    ids, t, pid, pt = inp.beads()
    (h, tl, k, l0), (fi, se, mo, ph) = inp.topology()
    lip = np.where(pt == 1)[0]
    first = lip[0]
    sel = (h >= first) & (h < first + 11)
    pairs = sorted(zip(h[sel] - first, tl[sel] - first))
    assert pairs == [(0, 1), (1, 2), (1, 6), (2, 3), (3, 4), (4, 5), (6, 7), (7, 8), (8, 9), (9, 10)]
    bsel = np.where(sel)[0]
    b0 = bsel[0]
    psel = (fi >= b0) & (fi < b0 + 10)
    trip = sorted((h[s] - first, tl[s] - first, tl[f] - first) for f, s in zip(fi[psel], se[psel]))
    # H-T-T and T-T-T triples only (HHT, HHH have no bond-pair type)
    assert trip == [(1, 2, 3), (2, 3, 4), (3, 4, 5), (6, 7, 8), (7, 8, 9), (8, 9, 10)]
    assert np.allclose(mo[psel], 20.0) and np.allclose(ph[psel], 0.0)


def test_polymer_counts_floor_rounding():
    """counts = floor(fraction * rho V / sum(size * fraction)) (src/InitialState.cpp:471)."""
    text = open(os.path.join(GOLDEN, "chains7", "case.dmpci")).read()
    inp = H.Input(text=text)
    rho_v = 3 * 7 ** 3
    nalpha = 0.984615 * 1 + 0.015385 * 16
    assert inp.counts["npolymers"] == int(0.984615 * rho_v / nalpha) + int(0.015385 * rho_v / nalpha)


@pytest.mark.parametrize("bad,msg", [
    ("Title", "Title"), ("Density\t\t3", "Density"), ("Lambda\t\t0.5", "lambda"),
])
def test_malformed_inputs_fail_with_the_item_named(bad, msg):
    text = open(os.path.join(GOLDEN, "water6", "case.dmpci")).read()
    broken = text.replace(bad, "Bogus", 1)
    with pytest.raises(H.InputError) as e:
        H.Input(text=broken)
    assert msg.lower() in str(e.value).lower()


def test_commands_are_read():
    text = open(os.path.join(GOLDEN, "water6", "case.dmpci")).read()
    text += "\nCommand ToggleDPDBeadThermostat 5\nCommand SetTimeStepSize 7 0.01\n"
    inp = H.Input(text=text)
    assert inp.commands() == [("ToggleDPDBeadThermostat", 5, []), ("SetTimeStepSize", 7, ["0.01"])]


def test_random_state_builder():
    inp = H.Input(text=open(os.path.join(GOLDEN, "lipid8", "case.dmpci")).read().replace(
        "restart RunId golden StateId 0", "random"))
    x, u, v = inp.build_state(seed=5)
    L = np.array(inp.cells, float)
    assert (x >= 0).all() and (x < L).all()
    img = (u - x) / L
    assert np.abs(img - np.round(img)).max() < 1e-9
    n = len(x)
    assert abs((v ** 2).sum() / (3 * n) - 1.0) < 1e-12
    assert np.abs(v.sum(axis=0)).max() < 1e-9
    # the reference's bond fix-up: no bond spans more than 1 along any axis
    (h, t, k, l0), _ = inp.topology()
    assert np.abs(u[h] - u[t]).max() <= 1.0 + 1e-12


def test_restart_roundtrip(tmp_path):
    from helpers import load_beads
    d = load_beads("water6", "in")
    L = np.array([6.0, 6.0, 6.0])
    p = tmp_path / "dmpcrs.t.con.0.dat"
    H.write_restart(p, d["poly"], d["ptype"], d["id"], d["type"], L, d["x"], d["u"], d["v"], d["f"])
    r = H.read_restart(p, len(d["id"]))
    assert (r["id"] == d["id"]).all()
    for key in ("u", "v", "f"):
        assert np.allclose(r[key], d[key], rtol=5e-6, atol=0)


def test_lamella_state_builder():
    """State lamella (CLamellaBuilder::Assemble, src/LamellaBuilder.cpp:126-1003): each
    lipid's head (first bead) on a monolayer plane at centre +- Thickness/2, its tail (last
    bead) at centre +- 0.2, the beads between on the line joining them (Linearise 1), the
    water outside the slab, no velocity along the normal.  (The reference's own states are
    matched to six digits in test_osprey_io.)"""
    text = open(os.path.join(GOLDEN, "lipid8", "case.dmpci")).read().replace(
        "restart RunId golden StateId 0",
        "lamella Polymer Lipid Normal 0 0 1 Centre 0.5 Thickness 5.0 Linearise 1 UpperFraction 0.5 Polymerise 0")
    inp = H.Input(text=text)
    assert inp.state == "lamella"
    x, u, v = inp.build_state(seed=9)
    ids, t, pid, pt = inp.beads()
    zc = 4.0
    lip = np.unique(pid[pt == 1])
    first = np.array([np.where(pid == p)[0][0] for p in lip])
    last = np.array([np.where(pid == p)[0][-1] for p in lip])
    upper = u[first, 2] > zc
    assert 0 < upper.sum() < len(lip)
    assert np.allclose(np.abs(u[first, 2] - zc), 2.5)
    assert np.allclose(u[last, 2] - zc, np.where(upper, 0.2, -0.2))
    for f, l in zip(first, last):
        seg = u[f:l + 1]
        assert np.allclose(seg[:, :2], seg[0, :2])                 # one line along the normal
        assert np.all(np.abs(seg[1:-1, 2] - zc) <= 2.5 + 1e-9)
    water = pt == 0
    assert (np.abs(u[water, 2] - zc) >= 2.5 - 1e-9).all()
    assert np.all(v[:, 2] == 0.0)


def test_bond_type_keying_of_the_slice_stress():
    """The reference keys its bond slice-stress table by bond type, the Bond declaration
    whose names match the bond's beads in either order (src/InputData.cpp:556-598,
    :1489-1498; src/Monitor.cpp:2305).  Folding the (head type, tail type) table into
    declarations must equal accumulating each bond directly under its declaration."""
    import oracle as O
    from dpd_amd import analysis as A
    text = open(os.path.join(GOLDEN, "lipid8", "case.dmpci")).read().replace(
        "restart RunId golden StateId 0",
        "lamella Polymer Lipid Normal 0 0 1 Centre 0.5 Thickness 5.0 Linearise 1 UpperFraction 0.5 Polymerise 0")
    inp = H.Input(text=text)
    assert inp.bond_types() == [(0, 0), (0, 1), (1, 1)]
    x, u, v = inp.build_state(seed=4)
    ids, t, pid, pt = inp.beads()
    (h, tl, k, l0), (fi, se, mo, ph) = inp.topology()
    # stretch the bonds away from rest so every one carries a force
    l0 = np.random.default_rng(0).uniform(0.3, 0.8, size=len(h))
    box = O.Box(inp.cells, inp.a, inp.gamma, seed=3)
    nt, ns = len(inp.bead_names), 16
    _, bond = O.slice_stress(box, ids, t, x, v, 0, ns, unpbc=u, topo=O.Topo(h, tl, k, l0, fi, se, mo, ph))
    got = A.bond_type_slices(bond, inp.bond_types(), nt)
    names = inp.bead_names
    decl = {}
    for b, (bh, bt) in enumerate(inp.bond_types()):
        decl[names[bh] + names[bt]] = b
        decl[names[bt] + names[bh]] = b
    want = np.zeros((ns, 3, 9))
    for q in range(len(h)):
        d = u[h[q]] - u[tl[q]]
        L = np.linalg.norm(d)
        F = k[q] * (l0[q] - L) * d / L
        O.slice_add(want, inp.cells[2], decl[names[t[h[q]]] + names[t[tl[q]]]], x[h[q], 2], x[tl[q], 2], F, d)
    assert np.abs(want).max() > 1.0
    assert np.allclose(got, want, rtol=1e-12, atol=1e-9)


def test_run_with_missing_restart_fails_with_the_file_named(tmp_path):
    """dpd_run reads the restart snapshot before it touches a device: a missing file is an
    error naming the file (the reference's CRestartBuilder returns false, src/RestartBuilder.cpp:84-92)."""
    text = open(os.path.join(GOLDEN, "lipid8", "case.dmpci")).read()
    text = text.replace("restart RunId golden StateId 0", "restart RunId nowhere StateId 5")
    with open(os.path.join(tmp_path, "dmpci.m"), "w") as fh:
        fh.write(text)
    with pytest.raises(H.InputError) as err:
        H.run(tmp_path, "m")
    assert "dmpcrs.nowhere.con.5.dat" in str(err.value)
