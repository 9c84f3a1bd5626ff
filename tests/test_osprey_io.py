"""CPU checks of the Osprey file formats against files the compiled reference wrote:
the dmpcis initial-state summary (CInitialState::SerializeAsciiFile,
src/InitialState.cpp:1153-1347) and a restart snapshot read back through dpd_restart_read
(src/RestartState.cpp:100-156, :250-330).  No GPU calls."""
import gzip
import os
import shutil

import numpy as np
import pytest

from dpd_amd import host as H
from helpers import GOLDEN, load_beads

IO = os.path.join(GOLDEN, "io")
CASES = ["water6", "binary6", "chains7", "lipid8", "lipidphi8"]


@pytest.mark.parametrize("case", CASES)
def test_dmpcis_summary_matches_reference(case, tmp_path):
    """The summary written for the case's control file is the reference's, byte for byte
    (counts, bead rows, bond and bond-pair types with phi0 in radians, polymer fractions
    and the bead, bond and bond-pair types of each polymer)."""
    inp = H.Input(os.path.join(GOLDEN, case, "case.dmpci"))
    out = tmp_path / "dmpcis"
    inp.write_summary(out)
    want = open(os.path.join(IO, "dmpcis." + case)).read()
    got = open(out).read()
    assert got == want, "\n".join(f"{a!r} != {b!r}" for a, b in zip(got.splitlines(), want.splitlines()) if a != b)


def test_dmpcis_summary_detects_a_changed_input(tmp_path):
    """A changed parameter shows in the summary (the comparison above is not vacuous)."""
    text = open(os.path.join(GOLDEN, "lipidphi8", "case.dmpci")).read().replace("15.0  30.0", "15.0  35.0")
    H.Input(text=text).write_summary(tmp_path / "s")
    assert open(tmp_path / "s").read() != open(os.path.join(IO, "dmpcis.lipidphi8")).read()


def test_reads_a_restart_file_the_reference_wrote(tmp_path):
    """dpd_restart_read on the reference's own restart file (flag line, 17 columns with the
    bead radius, 6 significant digits) gives the snapshot's beads."""
    path = tmp_path / "dmpcrs.lipidphi8.con.0.dat"
    with gzip.open(os.path.join(IO, "dmpcrs.lipidphi8.con.0.dat.gz"), "rb") as fi, open(path, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    ref = load_beads("lipidphi8", "in")
    n = len(ref["id"])
    r = H.read_restart(path, n)
    for k in ("poly", "ptype", "id", "type"):
        assert np.array_equal(r[k], ref[k]), k
    for k in ("x", "u", "v", "f"):
        assert np.array_equal(r[k], ref[k]), k
    # and the control file's polymers are the ones in the file, bead for bead
    inp = H.Input(os.path.join(GOLDEN, "lipidphi8", "case.dmpci"))
    ids, types, pid, pt = inp.beads()
    assert np.array_equal(ids, r["id"]) and np.array_equal(types, r["type"])
    assert np.array_equal(pid, r["poly"]) and np.array_equal(pt, r["ptype"])
    # a file with fewer beads than asked for is refused
    with pytest.raises(H.InputError):
        H.read_restart(path, n + 1)


@pytest.mark.parametrize("case", ["water6", "chains7", "lipid8"])
def test_random_state_is_the_references(case):
    """State random builds the reference's own starting state: the same uniform stream
    (the 64-bit LCG from |RNGSeed|), Maxwell speeds from the same table
    (CBuilder::MaxwellDist), the same bond fix-up and velocity normalisation
    (CRandomBuilder::Assemble).  The reference's one step with the forces toggled off moves
    every bead by dt v and keeps v, so its time-1 snapshot fixes the built x and v to its 6
    digits."""
    import re
    text = open(os.path.join(GOLDEN, case, "case.dmpci")).read()
    text = re.sub(r"State\s+restart RunId golden StateId 0", "State\trandom", text)
    inp = H.Input(text=text)
    assert inp.counts["state"] == 0
    x, u, v = inp.build_state(seed=abs(inp.counts["seed"]))
    ref = np.loadtxt(os.path.join(IO, f"random_{case}.dat.gz"))
    L = np.array(inp.cells, float)
    assert len(ref) == len(x)
    rv, ru = ref[:, 10:13], ref[:, 7:10]
    dt = inp.counts["dt"]
    tol = lambda r: 5.01e-6 * np.abs(r) + 1e-12
    assert np.all(np.abs(v - rv) <= tol(rv))
    assert np.all(np.abs(u + dt * v - ru) <= tol(ru) + 1e-9)
    assert np.allclose(x, np.mod(u, L))
    # kT exactly, no drift, as the reference normalises
    assert abs((v ** 2).sum() / (3 * len(v)) - inp.counts["kT"]) < 1e-12
    assert np.abs(v.sum(axis=0)).max() < 1e-9


@pytest.mark.parametrize("tag", ["z", "x", "y"])
def test_lamella_state_is_the_references(tag):
    """State lamella builds the reference's own starting state (CLamellaBuilder::Assemble,
    src/LamellaBuilder.cpp:126-1003): lipids on the hexagonal lattices of the two monolayers,
    heads and tails on the lamella's planes and the beads between on the line (Linearise 1) or
    drawn about it (Linearise 0), the solvent in the free slabs (CBuilder::isRandom), 2-d
    Maxwell velocities in the plane; normals along z, x and y, monolayer fractions 0.5, 0.4,
    0.6.  The reference's time-1 snapshot with the forces toggled off fixes x and v to its 6
    digits (as for the random builder)."""
    text = open(os.path.join(IO, f"lamella_{tag}.dmpci")).read()
    inp = H.Input(text=text)
    x, u, v = inp.build_state(seed=abs(inp.counts["seed"]))
    ref = np.loadtxt(os.path.join(IO, f"lamella_{tag}.dat.gz"))
    L = np.array(inp.cells, float)
    assert len(ref) == len(x)
    rv, ru = ref[:, 10:13], ref[:, 7:10]
    dt = inp.counts["dt"]
    tol = lambda r: 5.01e-6 * np.abs(r) + 1e-12
    assert np.all(np.abs(v - rv) <= tol(rv))
    assert np.all(np.abs(u + dt * v - ru) <= tol(ru) + 1e-9)
    assert np.array_equal(x, u)
    # the velocities lie in the plane, with kT per in-plane component and no drift
    normal = {"z": 2, "x": 0, "y": 1}[tag]
    assert np.all(v[:, normal] == 0.0)
    assert abs((v ** 2).sum() / (2 * len(v)) - inp.counts["kT"]) < 1e-12
    assert np.abs(v.sum(axis=0)).max() < 1e-9
    assert (x >= 0).all() and (x < L).all()
