"""The whole-run driver (dpd_run, the equivalent of the reference's `dpd <runId>`,
src/dmpc.cpp:64-226 and CSimBox::Run, src/SimBox.cpp:1272-1355) on the GPU: history rows
every SamplePeriod, restart snapshots every RestartPeriod that read back and restart a run,
commands applied at their times, and the command-line front end."""
import os
import subprocess

import numpy as np
import pytest

from dpd_amd import host as H
from helpers import ROOT

pytestmark = pytest.mark.gpu

WATER = """dpd

Title	" water 8 "
Date    02/12/26
Comment	" driver test "

State	{state}

Bead  W
      0.5
      25
      4.5

Polymer	Water    1.0   " (W) "

Box         8 8 8       1  1  1
Density		3
Temp        1
RNGSeed		-4242
Lambda		0.5
Step		0.02
Time		{time}
SamplePeriod     20
AnalysisPeriod	 {time}
DensityPeriod    {time}
DisplayPeriod    {time}
RestartPeriod    200
Grid		1  1  1
{commands}"""


def write(tmp, run_id, state="random", time=400, commands=""):
    with open(os.path.join(tmp, "dmpci." + run_id), "w") as fh:
        fh.write(WATER.format(state=state, time=time, commands=commands))


def history(tmp, run_id):
    return np.loadtxt(os.path.join(tmp, "dmpchs." + run_id), ndmin=2)


def test_run_writes_history_and_restarts(tmp_path):
    write(tmp_path, "w")
    st = H.run(tmp_path, "w")
    print(f"  driver: {st['steps']} steps, {st['nbeads']} beads, kT {st['last_kT']:.4f}, P {st['last_pressure']:.3f}",
          flush=True)
    assert st["steps"] == 400 and st["nbeads"] == 1536 and st["samples"] == 20
    h = history(tmp_path, "w")
    assert h.shape == (20, 8)   # time kT P KE PE E, MSD of W, end-to-end of Water
    assert np.array_equal(h[:, 0], np.arange(20, 401, 20))
    late = h[h[:, 0] > 200]
    # water at a = 25, rho = 3, dt = 0.02: kT ~1.01 and P ~23.7 (reference dmpchs rows)
    assert abs(late[:, 1].mean() - 1.0) < 0.05
    assert abs(late[:, 2].mean() - 23.7) < 1.0
    for t in (200, 400):
        r = H.read_restart(os.path.join(tmp_path, f"dmpcrs.w.con.{t}.dat"), 1536)
        assert np.all((r["x"] >= 0.0) & (r["x"] <= 8.0))
        assert np.isfinite(r["v"]).all()
        assert abs((r["v"] ** 2).sum() / (3 * 1536) - 1.0) < 0.15


def test_restart_continues_the_state(tmp_path):
    write(tmp_path, "w", time=200)
    H.run(tmp_path, "w")
    first = history(tmp_path, "w")
    write(tmp_path, "r", state="restart RunId w StateId 200", time=200)
    H.run(tmp_path, "r")
    second = history(tmp_path, "r")
    # the restarted run starts from the equilibrated snapshot: no start-up transient
    assert abs(second[0, 1] - 1.0) < 0.1
    assert abs(second[:, 2].mean() - first[first[:, 0] > 100][:, 2].mean()) < 1.0


def test_commands_are_applied(tmp_path):
    cmds = ("Command SetTimeStepSize 100 0.01\n"
            "Command ToggleDPDBeadThermostat 200\n"
            "Command SetDPDBeadConsInt 300 W W 30.0\n")
    write(tmp_path, "c", commands=cmds)
    st = H.run(tmp_path, "c")
    assert st["commands_applied"] == 3
    h = history(tmp_path, "c")
    assert np.isfinite(h).all()
    # stiffer repulsion after t = 300 raises the pressure
    assert h[h[:, 0] > 320][:, 2].mean() > h[(h[:, 0] > 120) & (h[:, 0] <= 300)][:, 2].mean() + 1.0


def test_command_line_front_end(tmp_path):
    exe = os.path.join(ROOT, "osprey-dpd_amd", "bin", "dpd_amd")
    if not os.path.exists(exe):
        pytest.skip("command-line front end not built (make -C osprey-dpd_amd/csrc)")
    write(tmp_path, "cli", time=200)
    r = subprocess.run([exe, "cli"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print("  " + r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    assert "run cli: 200 steps, 1536 beads" in r.stdout
    assert os.path.exists(os.path.join(tmp_path, "dmpcrs.cli.con.200.dat"))


def test_forces_toggle_and_momentum_commands(tmp_path):
    """ZeroTotalMom, RenormaliseMomenta (every 40 steps, before the step) and
    ToggleDPDBeadForces: once the forces are off the velocities only change by the
    renormalisation, so every sample from the first renormalisation after the toggle on
    has kT = 1 exactly."""
    cmds = ("Command ZeroTotalMom 50\n"
            "Command RenormaliseMomenta 60 40 1\n"
            "Command ToggleDPDBeadForces 300\n")
    write(tmp_path, "m", commands=cmds)
    st = H.run(tmp_path, "m")
    assert st["commands_applied"] == 3
    h = history(tmp_path, "m")
    late = h[h[:, 0] >= 320]
    print(f"  kT after the toggle: {late[:, 1]}", flush=True)
    assert len(late) == 5 and np.all(np.abs(late[:, 1] - 1.0) < 2e-6)
    # with forces on, the sample after a renormalised step is near but not at kT
    assert np.any(np.abs(h[(h[:, 0] > 60) & (h[:, 0] < 300), 1] - 1.0) > 1e-4)


CHAINS = """dpd

Title	" chains "
Date    02/12/26
Comment	" bond commands "

State	random

Bead  W
      0.5
      25
      4.5

Bead  A
      0.5
      25  25
      4.5 4.5

Bond  A A  128  0.5
BondPair  A A A  20.0  0.0

Polymer	Water    0.8   " (W) "
Polymer	Chain    0.2   " (A A A A A A A A) "

Box         8 8 8       1  1  1
Density		3
Temp        1
RNGSeed		-77
Lambda		0.5
Step		0.02
Time		300
SamplePeriod     20
AnalysisPeriod	 300
DensityPeriod    300
DisplayPeriod    300
RestartPeriod    150
Grid		1  1  1
Command SetBondStrength 151 AA 128 0.9
Command SetBondStiffness 151 AAA 40 60
"""


def test_bond_commands(tmp_path):
    """SetBondStrength (by the bond's name, head and tail names joined) and
    SetBondStiffness (bond pair name, angle in degrees): the bonds relax to the new
    unstretched length."""
    with open(os.path.join(tmp_path, "dmpci.b"), "w") as fh:
        fh.write(CHAINS)
    st = H.run(tmp_path, "b")
    assert st["commands_applied"] == 2
    inp = H.Input(text=CHAINS)
    (h, t, k, l0), _ = inp.topology()
    n = inp.counts["nbeads"]
    lengths = []
    for step in (150, 300):
        r = H.read_restart(os.path.join(tmp_path, f"dmpcrs.b.con.{step}.dat"), n)
        lengths.append(np.linalg.norm(r["u"][h] - r["u"][t], axis=1).mean())
    print(f"  mean bond length {lengths[0]:.3f} -> {lengths[1]:.3f}", flush=True)
    assert lengths[0] < 0.75 and lengths[1] > 0.8


def test_history_and_summary_match_the_reference(tmp_path):
    """dpd_run from the restart file the reference wrote, thermostat off from step 1 (a
    deterministic trajectory) and ToggleEnergyOutput on: the dmpcis summary is the
    reference's byte for byte, and every dmpchs row -- time, kT, P, kinetic, potential
    (the reference's pair sum) and total energy, MSD per bead type, end-to-end distance per
    polymer type -- matches the reference's within 1e-4 relative (FP32 trajectory against
    the reference's FP64 over ten steps), in the reference's column layout."""
    import gzip
    import shutil
    from helpers import GOLDEN
    io = os.path.join(GOLDEN, "io")
    with gzip.open(os.path.join(io, "dmpcrs.lipidphi8.con.0.dat.gz"), "rb") as fi, \
            open(os.path.join(tmp_path, "dmpcrs.lipidphi8.con.0.dat"), "wb") as fo:
        shutil.copyfileobj(fi, fo)
    shutil.copy(os.path.join(io, "dmpci.lipidphi8_nf10"), os.path.join(tmp_path, "dmpci.nf"))
    st = H.run(tmp_path, "nf")
    assert st["steps"] == 10 and st["samples"] == 10
    assert open(os.path.join(tmp_path, "dmpcis.nf")).read() == open(os.path.join(io, "dmpcis.lipidphi8")).read()
    got_txt = open(os.path.join(tmp_path, "dmpchs.nf")).read().splitlines()
    ref_txt = open(os.path.join(io, "dmpchs.lipidphi8_nf10")).read().splitlines()
    got = np.array([[float(t) for t in l.split()] for l in got_txt])
    ref = np.array([[float(t) for t in l.split()] for l in ref_txt])
    assert got.shape == ref.shape == (10, 11)
    # fields of 12 columns and a space, as the reference writes them
    assert all(len(l) == 13 * 11 for l in got_txt)
    assert np.array_equal(got[:, 0], ref[:, 0])
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-12)
    err[ref == 0] = np.abs(got[ref == 0])
    for c, name in enumerate(["t", "kT", "P", "KE", "PE", "E", "MSD H", "MSD T", "MSD W", "EE water", "EE lipid"]):
        print(f"  {name:9s} max rel diff {err[:, c].max():.2e}", flush=True)
    assert err.max() < 1e-4
    # the final snapshot: against the oracle run from the same restart file as the engine
    # reads it, and against the reference's within what the file's rounding moves.  The file
    # rounds wrapped and unwrapped coordinates to six digits each, up to 5e-5 apart; the
    # reference takes bond vectors from the unwrapped ones (src/Bond.cpp:111-135), the engine
    # from x + image * L, and ten steps carry that difference into the velocities.
    import oracle as O
    from helpers import clamped_and_bonded, load_beads
    ref_state = np.loadtxt(os.path.join(io, "lipidphi8_nf10.dat.gz"))
    r = H.read_restart(os.path.join(tmp_path, "dmpcrs.nf.con.10.dat"), len(ref_state))
    r0 = H.read_restart(os.path.join(tmp_path, "dmpcrs.lipidphi8.con.0.dat"), len(ref_state))
    inp = H.Input(os.path.join(tmp_path, "dmpci.nf"))
    (bh, bt, bk, bl), (fi, se, mo, ph) = inp.topology()
    L = np.array(inp.cells, dtype=float)
    img = np.round((r0["u"] - r0["x"]) / L)
    box = O.Box(inp.cells, inp.a, np.zeros_like(inp.gamma), seed=5150, rng_mode=O.RNG_PHILOX)
    s = O.System(box, r0["id"], r0["type"], r0["x"], r0["v"], r0["f"], unpbc=r0["x"] + img * L,
                 topo=O.Topo(bh, bt, bk, bl, fi, se, mo, ph))
    s.run(10)
    _, uo, vo, _ = s.state()
    ok = ~clamped_and_bonded(load_beads("lipidphi8", "in"), inp.cells, bh, bt, fi, se)
    dv = np.abs(r["v"] - ref_state[:, 10:13])
    dvo = np.abs(vo - ref_state[:, 10:13])
    print(f"  final velocities: against the oracle max diff {np.abs(r['v'] - vo).max():.2e}; against the "
          f"reference {dv[ok].max():.2e} (oracle {dvo[ok].max():.2e}; {int((~ok).sum())} beads set aside)",
          flush=True)
    # the restart file the engine wrote carries six digits
    assert np.allclose(r["v"], vo, rtol=1e-5, atol=2e-5)
    assert np.allclose(r["u"], uo, rtol=1e-5, atol=2e-5)
    assert np.all(dv[ok] <= dvo[ok] + 2e-5)
    assert np.allclose(r["u"][ok], ref_state[ok, 7:10], rtol=1e-5, atol=1e-4)
