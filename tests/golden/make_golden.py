#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ by running the compiled Osprey-DPD
reference (oracle/_ref/dpd, built from /root/reference/src by oracle/ref.mk).

Run in the development container only (the reference binary and sources do not
exist on the GPU machine):

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py

Every case is a small box.  The reference first builds a random state and relaxes it
for WARM steps (``State random``), then restarts from the 6-digit restart snapshot it
wrote (``State restart``, src/InitialStateRestart.cpp:96-139, src/RestartState.cpp:250-330)
and advances it:

* ``nf1``  one step with ``Command ToggleDPDBeadThermostat 1`` (gamma = 0, so the
           dissipative and random forces vanish: src/CNTCell.cpp:318-346);
* ``th1``  one step with the thermostat on (serial LCG stream, src/CNTCell.cpp:2454-2465);
* ``th10`` ten steps with the thermostat on, plus the ten ``dmpchs`` history rows
           (time, kT, P, ... : src/Monitor.cpp:1750-1823).

Both sides parse the same text snapshot, so their inputs are bit-identical.  The
fixtures keep the reference's own 6-significant-digit values
(src/RestartState.cpp:134-156).  Files are gzip'd text: one bead per line,
``polyId polyType beadId beadType x y z ux uy uz vx vy vz fx fy fz``.

``th10_energy.txt`` (every case) repeats the ten thermostatted steps with
``Command ToggleEnergyOutput 1 0`` (src/mcToggleEnergyOutput.cpp, src/Monitor.cpp:3171):
its ``dmpchs`` rows carry the kinetic and the potential energy (pair, bond and bond-pair
terms, src/Monitor.cpp:2755-2778, src/BondPair.cpp:320-378) after every step.

``commands/`` starts from the lipidphi8 snapshot and runs 14 steps under the command
schedule COMMANDS (time step changes, thermostat, conservative-force and force toggles,
interaction changes: src/CNTCell.cpp:223-377, :434-456): ``cmd14.dat.gz`` is the state
after the last step and ``cmd_history.txt`` the ``dmpchs`` rows of every step.

``io/`` (``make_golden.py io``): the dmpcis summary of every case, the lipidphi8 restart
file as the reference wrote it, and a deterministic 10-step run from it (thermostat off,
energy output on) with its dmpchs rows and final snapshot.

``rng.txt`` holds the first 1000 draws of the reference LCG for seed |RNGSeed|, logged
by ``Command GenerateRNGUniform`` (src/mcGenerateRNGUniformImpl.cpp:44-62).
"""
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "dpd")

WARM = 200

HEADER = """dpd

Title	" {title} "
Date    02/12/26
Comment	" golden fixture "

"""

TAIL = """Box         {nx} {ny} {nz}       1  1  1
Density		3
Temp        1
RNGSeed		{seed}
Lambda		0.5
Step		{dt}
Time		{time}
SamplePeriod     1
AnalysisPeriod	 {time}
DensityPeriod    {time}
DisplayPeriod    {time}
RestartPeriod    {time}
Grid		1  1  1
"""

# name -> (box, seed, dt, body text between State and Box)
CASES = {
    "water6": ((6, 6, 6), -4321, 0.02, """Bead  W
      0.5
      25
      4.5

Polymer	Water    1.0   " (W) "

"""),
    "binary6": ((6, 6, 6), -977, 0.02, """Bead  W
      0.5
      25
      4.5

Bead  O
      0.5
      80  25
      4.5 4.5

Polymer	Water    0.5   " (W) "
Polymer	Oil      0.5   " (O) "

"""),
    "chains7": ((7, 7, 7), -31337, 0.02, """Bead  W
      0.5
      25
      4.5

Bead  S
      0.5
      25 10
      4.5 4.5

Bead  P
      0.5
      25 25 25
      4.5 4.5 4.5

Bond    S P  128  0.5
Bond    P P  128  0.5

Polymer	Water    0.984615   " (W) "
Polymer	Chain    0.015385   " (S (3 (P P P S)) P P P) "

"""),
    "lipid8": ((8, 8, 8), -8086, 0.02, """Bead  H
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

Polymer	Water    0.90   " (W) "
Polymer	Lipid    0.10   " (H H (* (T T T T)) H T T T T) "

"""),
    "lipidphi8": ((8, 8, 8), -5150, 0.02, """Bead  H
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

BondPair  H H T  10.0  60.0
BondPair  H T T  15.0  30.0
BondPair  T T T  20.0  0.0

Polymer	Water    0.90   " (W) "
Polymer	Lipid    0.10   " (H H (* (T T T T)) H T T T T) "

"""),
}


def dmpci(name, state, time, extra=""):
    box, seed, dt, body = CASES[name]
    txt = HEADER.format(title=name) + "State\t" + state + "\n\n" + body
    txt += TAIL.format(nx=box[0], ny=box[1], nz=box[2], seed=seed, dt=dt, time=time)
    return txt + extra


def run(work, run_id, text):
    with open(os.path.join(work, "dmpci." + run_id), "w") as fh:
        fh.write(text)
    r = subprocess.run([REF, run_id], cwd=work, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference run {run_id} failed: {r.stdout.decode()[-2000:]}")


def bead_lines(path):
    """Bead lines of a reference restart file (drops the flag line, the radius column and
    the 'inclusive' tail)."""
    out = []
    with open(path) as fh:
        fh.readline()
        for line in fh:
            tok = line.split()
            if not tok or tok[0] == "inclusive":
                break
            out.append(" ".join(tok[:4] + tok[5:]))
    return out


def write_gz(path, lines):
    with gzip.open(path, "wt") as fh:
        fh.write("\n".join(lines) + "\n")


def history(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if tok:
                rows.append(tok)
    return rows


def count_summary(path):
    keys = {"polymers": r"Number of polymers created\s+(\d+)",
            "beads": r"Number of beads created\s+(\d+)",
            "bonds": r"Number of bonds created\s+(\d+)",
            "bondpairs": r"Number of bond pairs created\s+(\d+)"}
    txt = open(path).read()
    return {k: int(re.search(p, txt).group(1)) for k, p in keys.items()}


def make_case(name, work):
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    w = name + "w"
    run(work, w, dmpci(name, "random", WARM))
    restart = f"restart RunId {w} StateId {WARM}"
    write_gz(os.path.join(out, "in.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.{w}.con.{WARM}.dat")))
    run(work, name + "n", dmpci(name, restart, 1, "Command ToggleDPDBeadThermostat 1\n"))
    write_gz(os.path.join(out, "nf1.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.{name}n.con.1.dat")))
    run(work, name + "t", dmpci(name, restart, 1))
    write_gz(os.path.join(out, "th1.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.{name}t.con.1.dat")))
    run(work, name + "u", dmpci(name, restart, 10))
    write_gz(os.path.join(out, "th10.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.{name}u.con.10.dat")))
    hist = history(os.path.join(work, f"dmpchs.{name}u"))
    with open(os.path.join(out, "th10_history.txt"), "w") as fh:
        fh.write("# time kT P ... (reference dmpchs rows, thermostat on)\n")
        for row in hist:
            fh.write(" ".join(row) + "\n")
    with open(os.path.join(out, "case.dmpci"), "w") as fh:
        fh.write(dmpci(name, "restart RunId golden StateId 0", 10))
    with open(os.path.join(out, "counts.json"), "w") as fh:
        json.dump(count_summary(os.path.join(work, f"dmpcis.{name}u")), fh)


# (time, command text) of the commands fixture; a command of time t runs before step t
# (CSimBox::Run, src/SimBox.cpp:1272-1320)
COMMANDS = [
    (3, "SetTimeStepSize 3 0.01"),
    (4, "ToggleDPDBeadThermostat 4"),
    (6, "ToggleDPDBeadConservativeForces 6"),
    (7, "ToggleDPDBeadThermostat 7"),
    (8, "ToggleDPDBeadConservativeForces 8"),
    (9, "SetDPDBeadConsInt 9 T W 60"),
    (10, "SetDPDBeadDissInt 10 H W 3.0"),
    (11, "ToggleDPDBeadForces 11"),
    (12, "ToggleDPDBeadForces 12"),
    (13, "SetTimeStepSize 13 0.02"),
]
COMMAND_STEPS = 14


def warm_restart(name, work):
    """The relaxed snapshot every fixture of `name` starts from (the run is deterministic:
    it reproduces the committed in.dat.gz)."""
    w = name + "w"
    run(work, w, dmpci(name, "random", WARM))
    return w, f"restart RunId {w} StateId {WARM}"


def make_energy(name, work):
    w, restart = warm_restart(name, work)
    run(work, name + "e", dmpci(name, restart, 10, "Command ToggleEnergyOutput 1 0\n"))
    hist = history(os.path.join(work, f"dmpchs.{name}e"))
    with open(os.path.join(HERE, name, "th10_energy.txt"), "w") as fh:
        fh.write("# time kT P KE PE E ... (reference dmpchs rows, thermostat on, ToggleEnergyOutput)\n")
        for row in hist:
            fh.write(" ".join(row) + "\n")


def make_commands(work):
    name = "lipidphi8"
    out = os.path.join(HERE, "commands")
    os.makedirs(out, exist_ok=True)
    w, restart = warm_restart(name, work)
    inp = bead_lines(os.path.join(work, f"dmpcrs.{w}.con.{WARM}.dat"))
    with gzip.open(os.path.join(HERE, name, "in.dat.gz"), "rt") as fh:
        assert fh.read().split("\n")[:-1] == inp, "warm-up no longer reproduces the lipidphi8 snapshot"
    cmds = "".join(f"Command {c}\n" for _, c in COMMANDS)
    run(work, "cmd", dmpci(name, restart, COMMAND_STEPS, cmds))
    write_gz(os.path.join(out, f"cmd{COMMAND_STEPS}.dat.gz"),
             bead_lines(os.path.join(work, f"dmpcrs.cmd.con.{COMMAND_STEPS}.dat")))
    hist = history(os.path.join(work, "dmpchs.cmd"))
    with open(os.path.join(out, "cmd_history.txt"), "w") as fh:
        fh.write("# time kT P ... (reference dmpchs rows under the command schedule)\n")
        for row in hist:
            fh.write(" ".join(row) + "\n")
    with open(os.path.join(out, "commands.json"), "w") as fh:
        json.dump({"case": name, "steps": COMMAND_STEPS, "commands": COMMANDS}, fh, indent=1)


def make_rng(work):
    # a restart-based run draws nothing before step 1, so the logged draws are the
    # first draws of the stream seeded with |RNGSeed|
    run(work, "rngw", dmpci("water6", "random", WARM))
    run(work, "rng", dmpci("water6", f"restart RunId rngw StateId {WARM}", 1,
                           "Command GenerateRNGUniform 1 1000\n"))
    txt = open(os.path.join(work, "dmpcls.rng")).read()
    i = txt.index("GenerateRNGSequence\n1000 Uniform\n")
    vals = txt[i:].split("\n")[2:1002]
    with open(os.path.join(HERE, "rng.txt"), "w") as fh:
        fh.write(f"# seed {abs(CASES['water6'][1])}: first 1000 reference LCG uniforms (6 digits)\n")
        fh.write("\n".join(vals) + "\n")


def make_io(work):
    """Osprey file fixtures: the dmpcis summary the reference writes for every case
    (CInitialState::SerializeAsciiFile, src/InitialState.cpp:1153-1347), a restart file as
    the reference writes it (flag line, 17 columns, src/RestartState.cpp:100-156), and a
    deterministic run of the lipidphi8 snapshot (thermostat off from step 1, energy output
    on) with its full dmpchs rows (src/Monitor.cpp:1750-1823) and final snapshot."""
    out = os.path.join(HERE, "io")
    os.makedirs(out, exist_ok=True)
    for name in CASES:
        w, restart = warm_restart(name, work)
        run(work, name + "s", dmpci(name, restart, 1))
        shutil.copy(os.path.join(work, f"dmpcis.{name}s"), os.path.join(out, f"dmpcis.{name}"))
    # the reference's random starting state (CRandomBuilder::Assemble, Maxwell speeds of
    # CBuilder::MaxwellDist): one step with the forces toggled off moves every bead by
    # dt v and leaves v alone, so the time-1 snapshot carries the built state
    for name in ("water6", "chains7", "lipid8"):
        run(work, name + "r", dmpci(name, "random", 1, "Command ToggleDPDBeadForces 1\n"))
        write_gz(os.path.join(out, f"random_{name}.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.{name}r.con.1.dat")))
    w, restart = warm_restart("lipidphi8", work)
    src = os.path.join(work, f"dmpcrs.{w}.con.{WARM}.dat")
    with open(src, "rb") as fi, gzip.open(os.path.join(out, "dmpcrs.lipidphi8.con.0.dat.gz"), "wb") as fo:
        fo.write(fi.read())
    cmds = "Command ToggleDPDBeadThermostat 1\nCommand ToggleEnergyOutput 1 0\n"
    text = dmpci("lipidphi8", restart, 10, cmds)
    run(work, "nfe", text)
    shutil.copy(os.path.join(work, "dmpchs.nfe"), os.path.join(out, "dmpchs.lipidphi8_nf10"))
    write_gz(os.path.join(out, "lipidphi8_nf10.dat.gz"), bead_lines(os.path.join(work, "dmpcrs.nfe.con.10.dat")))
    with open(os.path.join(out, "dmpci.lipidphi8_nf10"), "w") as fh:
        fh.write(dmpci("lipidphi8", "restart RunId lipidphi8 StateId 0", 10, cmds))


LAMELLA_BODY = """Bead  H
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

Polymer	Water    {water}   " (W) "
Polymer	Lipid    {lipid}   " (H H (* (T T T T)) H T T T T) "

"""

# tag -> (box, normal, centre, thickness, linearise, upper fraction, lipid number fraction, seed)
LAMELLA = {
    "z": ((10, 10, 10), (0, 0, 1), 0.5, 5.0, 1, 0.5, 0.05, -2468),
    "x": ((8, 10, 10), (1, 0, 0), 0.45, 4.0, 0, 0.4, 0.05, -1357),
    "y": ((10, 8, 10), (0, 1, 0), 0.55, 5.0, 1, 0.6, 0.05, -9753),
}


def lamella_dmpci(tag, time, extra=""):
    box, normal, centre, thk, lin, upper, frac, seed = LAMELLA[tag]
    state = ("lamella\n        Polymer\t\t\tLipid\n        Normal\t\t\t{} {} {}\n        Centre\t\t\t{}\n"
             "        Thickness\t\t{}\n        Linearise\t\t{}\n        UpperFraction\t{}\n"
             "        Polymerise\t\t0").format(*normal, centre, thk, lin, upper)
    txt = HEADER.format(title="lamella " + tag) + "State\t" + state + "\n\n"
    txt += LAMELLA_BODY.format(water=f"{1 - frac:.5f}", lipid=f"{frac:.5f}")
    txt += TAIL.format(nx=box[0], ny=box[1], nz=box[2], seed=seed, dt=0.02, time=time)
    return txt + extra


def make_lamella(work):
    """The reference's lamella starting states (CLamellaBuilder::Assemble,
    src/LamellaBuilder.cpp:126-1003: the hexagonal lattice of lipids in two monolayers,
    isRandom for the solvent, 2-d Maxwell velocities): one step with the forces toggled off
    moves every bead by dt v and leaves v alone, so the time-1 snapshot carries the built
    state.  Normals along z, x and y; straight and randomly placed lipid beads; symmetric
    and asymmetric monolayers."""
    out = os.path.join(HERE, "io")
    os.makedirs(out, exist_ok=True)
    for tag in LAMELLA:
        text = lamella_dmpci(tag, 1)
        with open(os.path.join(out, f"lamella_{tag}.dmpci"), "w") as fh:
            fh.write(text)
        run(work, "lam" + tag, text + "Command ToggleDPDBeadForces 1\n")
        write_gz(os.path.join(out, f"lamella_{tag}.dat.gz"), bead_lines(os.path.join(work, f"dmpcrs.lam{tag}.con.1.dat")))


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    work = tempfile.mkdtemp(prefix="dpd_golden_")
    try:
        args = sys.argv[1:]
        if args[:1] == ["io"]:
            make_io(work)
            print("wrote io fixtures")
            return
        if args[:1] == ["lamella"]:
            make_lamella(work)
            print("wrote lamella fixtures")
            return
        if args[:1] == ["extras"]:
            # only the energy and command fixtures (the snapshots stay as committed)
            for name in CASES:
                make_energy(name, work)
                print("wrote energy rows of", name)
            make_commands(work)
            print("wrote commands")
            return
        make_rng(work)
        names = args or list(CASES)
        for name in names:
            make_case(name, work)
            make_energy(name, work)
            print("wrote", name)
        make_commands(work)
        print("wrote commands")
    finally:
        shutil.rmtree(work)


if __name__ == "__main__":
    main()
