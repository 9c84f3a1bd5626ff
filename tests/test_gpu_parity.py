"""GPU parity: the HIP engine against the FP64 oracle (Philox pair stream) and the
reference's golden snapshots.  Tolerances: forces within 1e-5 normwise relative
(FP32 pair math), positions/velocities within 1e-5 relative after one step."""
import sys

import numpy as np
import pytest

import dpd_amd as D
import oracle as O
from helpers import (SEEDS, chain_state, force_errors, load_beads, random_state, rel_norm,
                     restart_clamp)

pytestmark = pytest.mark.gpu

KERNELS = [pytest.param(D.PAIR_FULLSHELL, id="fullshell"), pytest.param(D.PAIR_HITLIST, id="hitlist")]
BINARY_A = np.array([[25.0, 80.0], [80.0, 25.0]])
BINARY_G = np.array([[4.5, 4.5], [4.5, 4.5]])
C4_A = np.array([[25.0, 25.0, 25.0], [25.0, 10.0, 25.0], [25.0, 25.0, 25.0]])
C4_G = np.full((3, 3), 4.5)
FORCE_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine(cells, a, g, kernel, seed=1234, dt=0.02, widths=(1.0, 1.0, 1.0)):
    return D.Engine(cells, a, g, dt=dt, seed=seed, widths=widths, pair_kernel=kernel)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("cells,ntypes,thermo", [((10, 10, 10), 2, True), ((10, 10, 10), 2, False),
                                                 ((20, 20, 20), 2, True), ((9, 11, 7), 2, True)])
def test_forces_match_oracle(kernel, cells, ntypes, thermo):
    say("forces", kernel, cells, thermo)
    n = 3 * int(np.prod(cells))
    ids, t, x, v = random_state(cells, n, ntypes, seed=7)
    e = engine(cells, BINARY_A, BINARY_G, kernel)
    e.upload(ids, t, x, v)
    f, vir, pe = e.compute_forces(step=5, thermostat=thermo)
    box = O.Box(cells, BINARY_A, BINARY_G, seed=1234)
    fr, virr, per = O.forces(box, ids, t, x, v, step=5, thermostat=thermo)
    norm, per_bead = force_errors(f, fr)
    say(f"  normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < FORCE_TOL
    assert per_bead < 20 * FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL
    assert abs(pe[0] - per[0]) / per[0] < FORCE_TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_forces_condensate_box_full_size(kernel):
    """64^3 cells at rho = 3 (786,432 beads, three types with a sticker a = 10)."""
    cells = (64, 64, 64)
    n = 3 * 64 ** 3
    say("condensate full size", kernel)
    ids, t, x, v = random_state(cells, n, 3, seed=11)
    e = engine(cells, C4_A, C4_G, kernel)
    e.upload(ids, t, x, v)
    f, vir, _ = e.compute_forces(step=17)
    fr, virr, _ = O.forces(O.Box(cells, C4_A, C4_G, seed=1234), ids, t, x, v, step=17)
    norm, per_bead = force_errors(f, fr)
    say(f"  normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL
    # Newton's third law on the GPU sum
    assert np.abs(f.sum(axis=0)).max() / np.abs(f).sum() < 1e-6


@pytest.mark.parametrize("cells", [(26, 26, 26), (84, 84, 84)], ids=["groups275", "groups9261"])
def test_forces_after_steps_sort_paths(cells):
    """The counting sort's two cell-start paths above 16,384 cell keys: 26^3 cells (52,728
    beads, 275 scan groups of 64 keys, so the last workgroup of the one-launch path holds
    three groups) and 84^3 cells (1,778,112 beads, 9,261 groups: above 8,192 groups the group
    counts are scanned in a launch of their own).  Three steps, then the conservative forces
    of the state reached against the oracle's on the same positions (after steps the engine's
    dissipative term reads the last predicted velocities, which the download does not
    return)."""
    n = 3 * int(np.prod(cells))
    say("sort paths", cells)
    ids, t, x, v = random_state(cells, n, 2, seed=13)
    e = engine(cells, BINARY_A, BINARY_G, D.PAIR_HITLIST)
    e.upload(ids, t, x, v)
    e.step(3)
    X, U, V, _ = e.download()
    L = np.array(cells, dtype=float)
    assert (X >= 0).all() and (X < L).all()
    f, vir, _ = e.compute_forces(step=9, thermostat=False)
    fr, virr, _ = O.forces(O.Box(cells, BINARY_A, BINARY_G, seed=1234), ids, t, X, V, step=9, thermostat=False)
    norm, per_bead = force_errors(f, fr)
    say(f"  normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_bonded_forces_match_oracle(kernel):
    cells = (12, 12, 12)
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, 60, 16, 4224, seed=3)
    n = len(ids)
    k = np.full(len(h), 128.0)
    l0 = np.full(len(h), 0.5)
    mod = np.where(np.arange(len(fi)) % 3 == 0, 15.0, 20.0)
    phi0 = np.where(np.arange(len(fi)) % 3 == 0, np.radians(30.0), 0.0)
    say("bonded", kernel, n)
    e = engine(cells, C4_A, C4_G, kernel)
    e.upload(ids, t, x, v, unpbc=u)
    e.upload_topology(h, tl, k, l0, fi, se, mod, phi0)
    f, vir, pe = e.compute_forces(step=2)
    topo = O.Topo(h, tl, k, l0, fi, se, mod, phi0)
    fr, virr, per = O.forces(O.Box(cells, C4_A, C4_G, seed=1234), ids, t, x, v, step=2, unpbc=u, topo=topo)
    norm, per_bead = force_errors(f, fr)
    say(f"  normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL
    assert np.allclose(pe[1:], per[1:], rtol=1e-6, atol=1e-6 * max(1.0, float(np.abs(per).max())))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("thermo", [False, True])
def test_one_step_matches_oracle(kernel, thermo):
    """One full integration step (UpdatePos, forces, bonds, angles, UpdateMom)."""
    cells = (12, 12, 12)
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, 40, 16, 4544, seed=5)
    k = np.full(len(h), 128.0)
    l0 = np.full(len(h), 0.5)
    mod = np.full(len(fi), 20.0)
    phi0 = np.zeros(len(fi))
    topo = O.Topo(h, tl, k, l0, fi, se, mod, phi0)
    g = C4_G if thermo else np.zeros_like(C4_G)
    box = O.Box(cells, C4_A, g, seed=99, rng_mode=O.RNG_PHILOX)
    # start from a force-consistent state: f = F(x, v)
    f0, _, _ = O.forces(box, ids, t, x, v, step=0, unpbc=u, topo=topo)
    say("one step", kernel, thermo)
    e = D.Engine(cells, C4_A, g, seed=99, pair_kernel=kernel)
    e.upload(ids, t, x, v, f=f0, unpbc=u, step0=0)
    e.upload_topology(h, tl, k, l0, fi, se, mod, phi0)
    e.step(1, observe=True)
    X, U, V, F = e.download()
    s = O.System(box, ids, t, x, v, f0, unpbc=u, topo=topo)
    s.run(1, step0=0)
    xr, ur, vr, fr = s.state()
    assert rel_norm(U, ur) < 1e-7
    assert rel_norm(V, vr) < FORCE_TOL
    norm, _ = force_errors(F, fr)
    say(f"  force normwise {norm:.2e}")
    assert norm < FORCE_TOL
    dx = (X - xr + 0.5 * np.array(cells)) % np.array(cells) - 0.5 * np.array(cells)
    assert np.abs(dx).max() < 1e-5
    ob = e.observables()
    obr = s.observables()
    assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < FORCE_TOL
    assert abs(ob["pressure"] - obr["pressure"]) / obr["pressure"] < FORCE_TOL
    assert rel_norm(ob["stress"], obr["stress"]) < FORCE_TOL


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", ["water6", "binary6"])
def test_golden_noise_free_step(kernel, case):
    """One noise-free step from the reference's snapshot vs the reference's own result."""
    a = np.array([[25.0]]) if case == "water6" else BINARY_A
    d = load_beads(case, "in")
    ref = load_beads(case, "nf1")
    say("golden", kernel, case)
    e = D.Engine((6, 6, 6), a, np.zeros_like(a), seed=SEEDS[case], pair_kernel=kernel)
    e.upload(d["id"], d["type"], d["x"], d["v"], f=d["f"], unpbc=d["u"])
    e.step(1)
    X, U, V, F = e.download()
    tol = 2e-5
    assert np.all(np.abs(V - ref["v"]) <= tol * np.abs(ref["v"]) + 1e-5)
    assert rel_norm(F, ref["f"]) < tol
    # The reference keeps the unwrapped coordinate independently of the wrapped one; its
    # restart writer moves a bead within 0.001 of a face to 0.001 without moving the
    # unwrapped coordinate (src/RestartState.cpp:492-540), so such input beads are
    # inconsistent by up to 0.001.  The engine derives unwrapped = wrapped + image * L.
    img = (d["u"] - d["x"]) / 6.0
    consistent = (np.abs(img - np.round(img)) < 1e-5).all(axis=1)
    ok = np.abs(U - ref["u"]) <= 1e-5 * np.abs(ref["u"]) + 1e-5
    assert ok[consistent].all()
    assert (~consistent).sum() <= 2
    Xc = restart_clamp(X, np.array([6.0, 6.0, 6.0]))
    assert np.all(np.abs(Xc - ref["x"]) <= 1e-5 * np.abs(ref["x"]) + 1e-5)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("lam", [0.5, 0.65])
def test_ten_steps_track_oracle(kernel, lam):
    """Ten steps with lambda = 1/2 (the steps read the predicted velocities from the stored
    half steps) and with another lambda (the predicted velocities stored apart)."""
    cells = (10, 10, 10)
    n = 3000
    ids, t, x, v = random_state(cells, n, 2, seed=21)
    box = O.Box(cells, BINARY_A, BINARY_G, seed=5, lam=lam)
    # relax the random start so forces are moderate
    s = O.System(box, ids, t, x, v)
    s.run(30)
    x, u, v, f = s.state()
    say("ten steps", kernel, "lambda", lam)
    e = D.Engine(cells, BINARY_A, BINARY_G, seed=5, lam=lam, pair_kernel=kernel)
    e.upload(ids, t, x, v, f=f, unpbc=u, step0=30)
    e.step(10, observe=True)
    X, U, V, F = e.download()
    s.run(10, step0=30)
    xr, ur, vr, fr = s.state()
    say(f"  u {rel_norm(U, ur):.2e} v {rel_norm(V, vr):.2e} f {rel_norm(F, fr):.2e}")
    assert rel_norm(U, ur) < 1e-6
    assert rel_norm(V, vr) < 1e-4
    assert rel_norm(F, fr) < 1e-4
    ob, obr = e.observables(), s.observables()
    assert abs(ob["kT"] - obr["kT"]) / obr["kT"] < 1e-4
    assert abs(ob["pressure"] - obr["pressure"]) / obr["pressure"] < 1e-4
    assert ob["step"] == 40


@pytest.mark.parametrize("kernel", KERNELS)
def test_momentum_conserved(kernel):
    cells = (16, 16, 16)
    ids, t, x, v = random_state(cells, 3 * 16 ** 3, 2, seed=4)
    e = D.Engine(cells, BINARY_A, BINARY_G, seed=8, pair_kernel=kernel)
    e.upload(ids, t, x, v)
    e.step(200, observe=True)
    ob = e.observables()
    say("momentum", ob["momentum"], "kT", ob["kT"])
    assert np.abs(ob["momentum"]).max() < 1e-2 * np.sqrt(len(ids))
    assert ob["jumps"] == 0
    X, U, V, F = e.download()
    L = np.array(cells, dtype=float)
    img = (U - X) / L
    assert np.abs(img - np.round(img)).max() < 1e-6
    assert (X >= 0).all() and (X < L).all()


def test_hitlist_overflow_matches_fullshell():
    """A dense cluster (~90 partners per bead inside it, several times the 24-entry
    per-thread partner list, so lists are evaluated and cleared mid-walk): the hit-list
    kernel matches the oracle and the full-shell kernel, which evaluates the same partners in
    the same order (to rounding: the two kernels' multiply-adds are contracted differently)."""
    cells = (8, 8, 8)
    rng = np.random.default_rng(21)
    n0 = 3 * 8 ** 3
    xs = [rng.uniform(0, 8, (n0, 3)), 3.0 + 2.0 * rng.uniform(0, 1, (300, 3))]
    x = np.concatenate(xs)
    n = len(x)
    v = rng.normal(size=(n, 3))
    t = rng.integers(0, 2, n).astype(np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    out = {}
    for k in (D.PAIR_FULLSHELL, D.PAIR_HITLIST):
        e = engine(cells, BINARY_A, BINARY_G, k)
        e.upload(ids, t, x, v)
        assert e.pair_kernel == k
        out[k] = e.compute_forces(step=3)
    fr, virr, _ = O.forces(O.Box(cells, BINARY_A, BINARY_G, seed=1234), ids, t, x, v, step=3)
    norm, _ = force_errors(out[D.PAIR_HITLIST][0], fr)
    say(f"  dense cluster: {n} beads, normwise vs oracle {norm:.2e}")
    assert norm < FORCE_TOL
    fh, fs = out[D.PAIR_HITLIST][0], out[D.PAIR_FULLSHELL][0]
    say(f"  hit list vs full shell: normwise {rel_norm(fh, fs):.2e}")
    assert rel_norm(fh, fs) < 1e-6
    assert rel_norm(out[D.PAIR_HITLIST][1], out[D.PAIR_FULLSHELL][1]) < 1e-6


@pytest.mark.parametrize("widths", [(1.0, 1.0, 1.0), (1.25, 1.0, 1.5)])
def test_hitlist_pairs_at_the_cutoff(widths):
    """Isolated pairs at distances just inside and just outside the cut-off, and coincident
    beads: the hit-list kernel tests candidates on 8-bit cell offsets with a widened bound and
    drops the extra partners at evaluation, so every pair inside must be kept (with the force
    of the oracle and of the full-shell kernel) and every pair outside must contribute
    nothing.  Pair centres sit two cells apart, on cell faces as well as inside cells."""
    cells = (12, 12, 12)
    w = np.array(widths)
    L = np.array(cells) * w
    rng = np.random.default_rng(5)
    dists = [1.0 - 1e-6, 1.0 - 1e-4, 1.0 - 3e-3, 0.999999, 1.0 + 1e-6, 1.0 + 1e-4, 1.0 + 3e-3, 0.5, 0.0]
    xs, k = [], 0
    for cz in range(1, 12, 3):
        for cy in range(1, 12, 3):
            for cx in range(1, 12, 3):
                d = dists[k % len(dists)]
                k += 1
                c = (np.array([cx, cy, cz]) + (0.0 if k % 3 == 0 else rng.uniform(0.1, 0.9, 3))) * w
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                xs += [c - 0.5 * d * u, c + 0.5 * d * u]
    x = np.mod(np.array(xs), L)
    n = len(x)
    v = rng.normal(size=(n, 3))
    t = rng.integers(0, 2, n).astype(np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    out = {}
    for kern in (D.PAIR_FULLSHELL, D.PAIR_HITLIST):
        e = engine(cells, BINARY_A, BINARY_G, kern, widths=widths)
        e.upload(ids, t, x, v)
        out[kern] = e.compute_forces(step=2)
    fr, _, _ = O.forces(O.Box(cells, BINARY_A, BINARY_G, seed=1234, widths=widths), ids, t, x, v, step=2)
    fh = out[D.PAIR_HITLIST][0]
    dr = np.linalg.norm(x[1::2] - x[0::2], axis=1)
    # (pairs within 1e-5 of the cut-off may fall either side in FP32: only the norms below)
    outside = (dr > 1.0 + 1e-5) | (dr < 1e-9)
    inside = (dr < 1.0 - 1e-5) & (dr > 1e-9)
    say(f"  widths {widths}: {n // 2} pairs, {int(inside.sum())} inside, {int(outside.sum())} outside or coincident")
    assert np.all(fh[0::2][outside] == 0) and np.all(fh[1::2][outside] == 0)
    assert np.all(np.linalg.norm(fh[0::2][inside], axis=1) > 0)
    assert rel_norm(fh, fr) < FORCE_TOL
    assert rel_norm(fh, out[D.PAIR_FULLSHELL][0]) < 1e-6


@pytest.mark.parametrize("kernel", KERNELS)
def test_forces_with_wider_cells(kernel):
    """CNT cells wider than the cutoff (Box ... 1.25 1 1.5 in a dmpci file): the
    neighbour shifts and cell keys use each axis' width."""
    cells, widths = (8, 9, 7), (1.25, 1.0, 1.5)
    L = np.array(cells) * np.array(widths)
    n = int(3 * np.prod(L))
    ids, t, x, v = random_state(cells, n, 2, seed=44, widths=widths)
    e = engine(cells, BINARY_A, BINARY_G, kernel, widths=widths)
    e.upload(ids, t, x, v)
    f, vir, _ = e.compute_forces(step=6)
    fr, virr, _ = O.forces(O.Box(cells, BINARY_A, BINARY_G, seed=1234, widths=widths), ids, t, x, v, step=6)
    norm, per_bead = force_errors(f, fr)
    say(f"  widths {widths}: normwise {norm:.2e} per-bead {per_bead:.2e}")
    assert norm < FORCE_TOL
    assert rel_norm(vir, virr) < FORCE_TOL
