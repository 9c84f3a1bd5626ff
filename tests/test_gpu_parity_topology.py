"""GPU parity on the reference's own topologies and at the benchmark sizes.

* The chain, branched-lipid and lipid-with-phi0 snapshots written by the reference: one
  noise-free step against the reference's own result, and ten thermostatted steps against
  the oracle started from the same snapshot (the oracle's Philox mode replaces the
  reference's serial LCG stream, which no parallel code reproduces).
* The command schedule of the reference's commands fixture (time-step changes, thermostat,
  conservative-force and force toggles, interaction changes) executed through the engine's
  own hooks, against the oracle running the same schedule.
* Forces and one full step of the water/oil 40^3, bilayer 48x48x32 and condensate 64^3
  boxes (bonds and angles included) against the oracle, and ten steps of the bilayer.

Every test prints the normwise and per-bead maximum force errors it measured."""
import json
import os
import sys

import numpy as np
import pytest

import dpd_amd as D
import oracle as O
from dpd_amd import host as H
from dpd_amd import workloads as W
from helpers import GOLDEN, SEEDS, clamped_and_bonded, consistent, force_errors, load_beads, rel_norm, restart_clamp

pytestmark = pytest.mark.gpu

KERNELS = [pytest.param(D.PAIR_HITLIST, id="hitlist"), pytest.param(D.PAIR_FULLSHELL, id="fullshell")]
TOPO_CASES = ["chains7", "lipid8", "lipidphi8"]
FORCE_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden_input(case):
    inp = H.Input(os.path.join(GOLDEN, case, "case.dmpci"))
    (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
    return inp, (h, t, k, l0, fi, se, mo, ph)


def engine_for(inp, topo, d, kernel, gamma=None, step0=0, seed=None):
    g = inp.gamma if gamma is None else gamma
    e = D.Engine(inp.cells, inp.a, g, dt=0.02, seed=seed if seed is not None else 1, pair_kernel=kernel)
    e.upload(d["id"], d["type"], d["x"], d["v"], f=d["f"], unpbc=d["u"], step0=step0)
    e.upload_topology(*topo)
    return e


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", TOPO_CASES)
def test_golden_noise_free_step_with_topology(kernel, case):
    """One noise-free step (gamma = 0) from the reference's snapshot with bonds and bond
    pairs (linear chains, branched H3(T4)2 lipids, bond pairs with phi0 = 0, 30 and 60
    degrees): against the FP64 oracle on the same reading of the snapshot within 1e-5, and
    against the reference's own result within what the snapshot's rounding moves.

    The snapshot's wrapped and unwrapped coordinates are each rounded to six digits, so they
    disagree by up to 5e-5 for a quarter of the beads.  The reference takes pair distances
    from the wrapped ones and bond vectors from the unwrapped ones (src/Bond.cpp:111-135,
    src/BondPair.cpp:138-160); the engine keeps one position per bead (wrapped offset plus
    image counts), which reads the unwrapped coordinate as x + image * L.  The oracle fed
    that same reading is the like-for-like check; the oracle fed the snapshot as the
    reference reads it meets the reference's result to six digits (test_oracle)."""
    inp, topo = golden_input(case)
    d = load_beads(case, "in")
    ref = load_beads(case, "nf1")
    say("golden with topology", kernel, case, len(d["id"]), "beads", len(topo[0]), "bonds", len(topo[4]), "angles")
    e = engine_for(inp, topo, d, kernel, gamma=np.zeros_like(inp.gamma), seed=SEEDS[case])
    e.step(1)
    X, U, V, F = e.download()
    L = np.array(inp.cells, dtype=float)
    # the oracle on the engine's reading of the snapshot
    img = np.round((d["u"] - d["x"]) / L)
    _, s = oracle_system(inp, topo, dict(d, u=d["x"] + img * L), SEEDS[case], gamma=np.zeros_like(inp.gamma))
    s.run(1)
    xo, uo, vo, fo = s.state()
    norm_o, per_o = force_errors(F, fo)
    # beads the restart writer moved off a face, and their bond partners, are compared
    # loosely: the reference's pair and bond terms see two positions for them
    aff = clamped_and_bonded(d, L, topo[0], topo[1], topo[4], topo[5])
    ok = ~aff
    norm, per = force_errors(F[ok], ref["f"][ok])
    norm_ro = force_errors(fo[ok], ref["f"][ok])[0]
    dv_ro = np.abs(vo[ok] - ref["v"][ok]).max()
    say(f"  against the oracle: force normwise {norm_o:.2e} per-bead {per_o:.2e}, v {rel_norm(V, vo):.2e}")
    say(f"  against the reference: force normwise {norm:.2e} per-bead {per:.2e} (oracle on the same "
        f"reading: {norm_ro:.2e}), v max {np.abs(V[ok] - ref['v'][ok]).max():.2e} (oracle {dv_ro:.2e}); "
        f"{int(aff.sum())} beads off a face or bonded to one")
    assert norm_o < FORCE_TOL
    assert np.abs(V - vo).max() < 1e-5 and np.abs(U - uo).max() < 1e-5
    assert norm <= norm_ro + FORCE_TOL
    assert np.abs(V[ok] - ref["v"][ok]).max() <= dv_ro + 1e-5
    assert aff.sum() < 0.02 * len(aff)
    # unwrapped coordinates: within the snapshot's own disagreement between the two readings
    # (beads the restart writer moved off a face carry up to 0.001 of it)
    gap = np.abs(d["u"] - (d["x"] + img * L))
    assert np.all(np.abs(U - ref["u"]) <= gap + 1e-5 * np.abs(ref["u"]) + 1e-5)
    assert np.all(np.abs(restart_clamp(X, L) - ref["x"]) <= 1e-5 * np.abs(ref["x"]) + 1e-5)


def oracle_system(inp, topo, d, seed, step0=0, gamma=None):
    box = O.Box(inp.cells, inp.a, inp.gamma if gamma is None else gamma, seed=seed, rng_mode=O.RNG_PHILOX)
    return box, O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"], topo=O.Topo(*topo))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", TOPO_CASES)
def test_golden_thermostatted_steps_track_oracle(kernel, case):
    """Ten thermostatted steps from the reference's snapshot, one box with bonds and bond
    pairs (beads crossing the periodic faces in an 8^3 box), against the oracle from the
    same snapshot with the same pair stream: positions, velocities, forces and the FP64
    observables after every step."""
    inp, topo = golden_input(case)
    # both sides start from the same self-consistent state (wrapped = unwrapped mod L)
    d = consistent(load_beads(case, "in"), inp.cells)
    seed = SEEDS[case]
    say("thermostatted steps", kernel, case)
    e = engine_for(inp, topo, d, kernel, seed=seed)
    _, s = oracle_system(inp, topo, d, seed)
    L = np.array(inp.cells, dtype=float)
    for t in range(10):
        e.step(1, observe=True)
        s.run(1, step0=t)
        X, U, V, F = e.download()
        xr, ur, vr, fr = s.state()
        norm, per = force_errors(F, fr)
        ob, obr = e.observables(), s.observables()
        if t in (0, 9):
            say(f"  step {t + 1}: force normwise {norm:.2e} per-bead {per:.2e}, v {rel_norm(V, vr):.2e}, "
                f"u {rel_norm(U, ur):.2e}, kT {ob['kT']:.6f} / {obr['kT']:.6f}")
        assert rel_norm(U, ur) < 1e-6
        assert rel_norm(V, vr) < 1e-4
        assert norm < 1e-4
        dx = (X - xr + 0.5 * L) % L - 0.5 * L
        assert np.abs(dx).max() < 1e-4
        assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < 1e-4
        assert abs(ob["pressure"] - obr["pressure"]) / abs(obr["pressure"]) < 1e-4
        assert np.allclose(ob["pe"], s.potential_energies(t + 1), rtol=1e-4, atol=1e-6)
        assert ob["step"] == t + 1


class EngineCommands:
    """The command handlers of the reference as the engine's hooks carry them out."""

    def __init__(self, e):
        self.e = e

    def run(self, names, cmd, state):
        tok = cmd.split()
        name = tok[0]
        e = self.e
        if name == "SetTimeStepSize":
            state["dt"] = float(tok[2])
            e.set_timestep(state["dt"])
        elif name == "ToggleDPDBeadThermostat":
            state["thermo"] = not state["thermo"]
            e.toggle(D.TOGGLE_THERMOSTAT, state["thermo"])
        elif name == "ToggleDPDBeadConservativeForces":
            state["cons"] = not state["cons"]
            e.toggle(D.TOGGLE_CONSERVATIVE, state["cons"])
        elif name == "ToggleDPDBeadForces":
            state["forces"] = not state["forces"]
            e.toggle(D.TOGGLE_FORCES, state["forces"])
        elif name in ("SetDPDBeadConsInt", "SetDPDBeadDissInt"):
            i, j, val = names.index(tok[2]), names.index(tok[3]), float(tok[4])
            key = "a" if name == "SetDPDBeadConsInt" else "g"
            state[key][i, j] = state[key][j, i] = val
            # the engine keeps the matrices and applies its toggles on top of them
            e.set_interactions(a=state["a"], gamma=state["g"])
        else:
            raise ValueError(cmd)


@pytest.mark.parametrize("kernel", KERNELS)
def test_commands_track_oracle(kernel):
    """The reference's 14-step command schedule (SetTimeStepSize, ToggleDPDBeadThermostat,
    ToggleDPDBeadConservativeForces, ToggleDPDBeadForces, SetDPDBeadConsInt/DissInt,
    src/CNTCell.cpp:223-377, :434-456), executed before the step of their time, on the
    engine through its hooks and on the oracle (which follows the reference's run of the same
    schedule to 6 digits, tests/test_oracle.py): state and observables after every step."""
    from test_oracle import apply_command
    with open(os.path.join(GOLDEN, "commands", "commands.json")) as fh:
        spec = json.load(fh)
    case = spec["case"]
    inp, topo = golden_input(case)
    d = consistent(load_beads(case, "in"), inp.cells)
    seed = SEEDS[case]
    e = engine_for(inp, topo, d, kernel, seed=seed)
    _, s = oracle_system(inp, topo, d, seed)
    names = inp.bead_names
    st_e = {"dt": 0.02, "thermo": True, "cons": True, "forces": True, "a": inp.a.copy(), "g": inp.gamma.copy()}
    st_o = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st_e.items()}
    ec = EngineCommands(e)
    sched = {}
    for t, c in spec["commands"]:
        sched.setdefault(t, []).append(c)
    say("commands", kernel, spec["steps"], "steps")
    for t in range(1, spec["steps"] + 1):
        for c in sched.get(t, []):
            ec.run(names, c, st_e)
            apply_command(c, s, names, st_o)
        e.step(1, observe=True)
        s.run(1, step0=t - 1)
        X, U, V, F = e.download()
        xr, ur, vr, fr = s.state()
        ob, obr = e.observables(), s.observables()
        norm, per = force_errors(F, fr)
        say(f"  step {t:2d} {' / '.join(sched.get(t, [])) or '-':40s} force {norm:.2e} (per-bead {per:.2e}) "
            f"v {rel_norm(V, vr):.2e} kT {ob['kT']:.6f} / {obr['kT']:.6f}")
        assert rel_norm(U, ur) < 1e-6
        assert rel_norm(V, vr) < 1e-4
        assert norm < 1e-4
        assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < 1e-4
        assert abs(ob["pressure"] - obr["pressure"]) / abs(obr["pressure"]) < 1e-4


# ------------------------------------------------ the benchmark configurations ------

def workload(name):
    return {"binary40": W.binary, "bilayer": W.bilayer, "condensate": W.condensate}[name]()


def wl_topo(wl):
    h, t, k, l0 = wl.bonds if wl.bonds is not None else ((), (), (), ())
    fi, se, mo, ph = wl.pairs if wl.pairs is not None else ((), (), (), ())
    return (h, t, k, l0, fi, se, mo, ph)


@pytest.mark.parametrize("kernel", [pytest.param(D.PAIR_HITLIST, id="hitlist"),
                                    pytest.param(D.PAIR_FULLSHELL, id="fullshell")])
@pytest.mark.parametrize("name", ["binary40", "bilayer", "condensate"])
def test_full_size_forces_and_one_step(kernel, name):
    """Forces of the full benchmark boxes (192,000 / 221,175 beads with 37k bonds and 22k
    angles / 786,432 beads with 147,450 bonds) against the oracle, then one full step."""
    wl = workload(name)
    topo = wl_topo(wl)
    ot = O.Topo(*topo)
    box = O.Box(wl.cells, wl.a, wl.gamma, seed=wl.seed, rng_mode=O.RNG_PHILOX)
    say("full size", name, kernel, wl.n, "beads", wl.nbonds, "bonds", wl.nangles, "angles")
    e = D.Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed, pair_kernel=kernel)
    e.upload(wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u)
    e.upload_topology(*topo)
    f, vir, pe = e.compute_forces(step=3)
    fr, virr, per = O.forces(box, wl.ids, wl.types, wl.x, wl.v, step=3, unpbc=wl.u, topo=ot)
    norm, pb = force_errors(f, fr)
    say(f"  forces: normwise {norm:.2e} per-bead {pb:.2e}, virial {rel_norm(vir, virr):.2e}")
    assert norm < FORCE_TOL
    assert pb < 20 * FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL
    assert np.allclose(pe, per, rtol=1e-5, atol=1e-6)
    # one step from a force-consistent state
    e.upload(wl.ids, wl.types, wl.x, wl.v, f=fr, unpbc=wl.u, step0=0)
    e.upload_topology(*topo)
    e.step(1, observe=True)
    X, U, V, F = e.download()
    s = O.System(box, wl.ids, wl.types, wl.x, wl.v, fr, unpbc=wl.u, topo=ot)
    s.run(1, step0=0)
    xr, ur, vr, frr = s.state()
    norm, pb = force_errors(F, frr)
    say(f"  one step: force normwise {norm:.2e} per-bead {pb:.2e}, v {rel_norm(V, vr):.2e}, u {rel_norm(U, ur):.2e}")
    assert rel_norm(U, ur) < 1e-7
    assert rel_norm(V, vr) < FORCE_TOL
    assert norm < FORCE_TOL
    ob, obr = e.observables(), s.observables()
    assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < FORCE_TOL
    assert abs(ob["pressure"] - obr["pressure"]) / abs(obr["pressure"]) < FORCE_TOL


def test_bilayer_ten_steps_track_oracle():
    """Ten steps of the 48x48x32 bilayer (221,175 beads, branched lipids, bonds and bond
    pairs) with the hit-list kernel, after 20 oracle steps of relaxation."""
    wl = workload("bilayer")
    topo = wl_topo(wl)
    ot = O.Topo(*topo)
    box = O.Box(wl.cells, wl.a, wl.gamma, seed=wl.seed, rng_mode=O.RNG_PHILOX)
    s = O.System(box, wl.ids, wl.types, wl.x, wl.v, unpbc=wl.u, topo=ot)
    s.run(20)
    x, u, v, f = s.state()
    e = D.Engine(wl.cells, wl.a, wl.gamma, dt=wl.dt, seed=wl.seed, pair_kernel=D.PAIR_HITLIST)
    e.upload(wl.ids, wl.types, x, v, f=f, unpbc=u, step0=20)
    e.upload_topology(*topo)
    e.step(10, observe=True)
    s.run(10, step0=20)
    X, U, V, F = e.download()
    xr, ur, vr, fr = s.state()
    norm, pb = force_errors(F, fr)
    say(f"  bilayer 10 steps: force normwise {norm:.2e} per-bead {pb:.2e}, v {rel_norm(V, vr):.2e}, "
        f"u {rel_norm(U, ur):.2e}")
    assert rel_norm(U, ur) < 1e-6
    assert rel_norm(V, vr) < 1e-4
    assert norm < 1e-4
    ob, obr = e.observables(), s.observables()
    assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < 1e-4
    assert abs(ob["pressure"] - obr["pressure"]) / abs(obr["pressure"]) < 1e-4
