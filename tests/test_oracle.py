"""CPU tests of the FP64 oracle: pinned against the reference's own outputs (golden
snapshots written by the compiled reference), rocRAND, closed forms and brute force."""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, SEEDS, load_beads, load_history, restart_clamp, six_digit_close

WATER_A = [[25.0]]
WATER_G = [[4.5]]
BINARY_A = [[25.0, 80.0], [80.0, 25.0]]
BINARY_G = [[4.5, 4.5], [4.5, 4.5]]
SIMPLE = {"water6": ((6, 6, 6), WATER_A, WATER_G), "binary6": ((6, 6, 6), BINARY_A, BINARY_G)}


def test_lcg_matches_reference_draws():
    """CCNTCell::Randf (src/CNTCell.cpp:2454-2465) vs the reference's GenerateRNGUniform log."""
    ref = np.loadtxt(os.path.join(GOLDEN, "rng.txt"))
    got = O.lcg_sequence(SEEDS["water6"], len(ref))
    assert six_digit_close(got, ref).all()
    # plausible mistakes fail: another multiplier, or the low word instead of the high word
    def seq(mult, shift):
        s, out = SEEDS["water6"], []
        for _ in range(len(ref)):
            s = (mult * s + 1) % 2**64
            out.append(((s >> shift) & 0xFFFFFFFF) / 2**32)
        return np.array(out)
    assert six_digit_close(seq(6364136223846793005, 32), ref).all()
    assert not six_digit_close(seq(6364136223846793007, 32), ref).all()
    assert not six_digit_close(seq(6364136223846793005, 0), ref).all()


def test_philox_matches_rocrand():
    rows = [l.split() for l in open(os.path.join(GOLDEN, "philox_kat.txt")) if not l.startswith("#")]
    assert len(rows) >= 6
    for r in rows:
        seed = int(r[0], 16)
        ctr = [int(t, 16) for t in r[1:5]]
        exp = tuple(int(t, 16) for t in r[5:9])
        assert O.philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32)) == exp


def test_pair_uniform_symmetric_and_step_keyed():
    u1 = O.pair_uniform(7, 11, 3, 9)
    assert u1 == O.pair_uniform(7, 11, 9, 3)
    assert u1 != O.pair_uniform(7, 12, 3, 9)
    assert 0.0 <= u1 < 1.0


def test_pair_uniform_moments():
    rng = np.random.default_rng(0)
    ids = rng.integers(1, 10**6, size=(20000, 2))
    u = np.array([O.pair_uniform(99, 5, int(a), int(b)) for a, b in ids])
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.003
    # consecutive steps are uncorrelated
    u2 = np.array([O.pair_uniform(99, 6, int(a), int(b)) for a, b in ids])
    assert abs(np.corrcoef(u, u2)[0, 1]) < 0.03


def _run_golden(case, name, nsteps, lam=0.5, thermostat=None, seed=None):
    cells, a, g = SIMPLE[case]
    d = load_beads(case, "in")
    thermo = name != "nf1" if thermostat is None else thermostat
    gam = np.array(g) if thermo else np.zeros_like(np.array(g))
    box = O.Box(cells, np.array(a), gam, seed=SEEDS[case] if seed is None else seed, lam=lam,
                rng_mode=O.RNG_LCG)
    s = O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"])
    s.run(nsteps)
    return s, box


@pytest.mark.parametrize("case", ["water6", "binary6"])
@pytest.mark.parametrize("name,nsteps", [("nf1", 1), ("th1", 1), ("th10", 10)])
def test_oracle_matches_reference_snapshot(case, name, nsteps):
    s, box = _run_golden(case, name, nsteps)
    x, u, v, f = s.state()
    ref = load_beads(case, name)
    assert (ref["id"] == load_beads(case, "in")["id"]).all()
    assert six_digit_close(restart_clamp(x, box.lengths), ref["x"]).all()
    assert six_digit_close(u, ref["u"]).all()
    assert six_digit_close(v, ref["v"]).all()
    assert six_digit_close(f, ref["f"]).all()


@pytest.mark.parametrize("case", ["water6", "binary6"])
def test_oracle_matches_reference_history(case):
    """kT and P rows of dmpchs (src/Monitor.cpp:979-982) over 10 thermostatted steps."""
    cells, a, g = SIMPLE[case]
    d = load_beads(case, "in")
    box = O.Box(cells, np.array(a), np.array(g), seed=SEEDS[case], rng_mode=O.RNG_LCG)
    s = O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"])
    hist = load_history(case)
    for row in hist:
        s.run(1)
        ob = s.observables()
        assert six_digit_close(np.array([ob["kT"], ob["pressure"]]), row[1:3]).all(), row
        assert ob["wrong_cell"] == 0


def test_planted_mistakes_are_caught():
    """The snapshot checks fail for plausible mistakes planted in the oracle's inputs."""
    ref = load_beads("water6", "th1")
    # wrong lambda (velocity prediction) changes the dissipative force
    s, _ = _run_golden("water6", "th1", 1, lam=0.0)
    assert not six_digit_close(s.state()[3], ref["f"]).all()
    # wrong seed changes the random force
    s, _ = _run_golden("water6", "th1", 1, seed=SEEDS["water6"] + 1)
    assert not six_digit_close(s.state()[3], ref["f"]).all()
    # thermostat dropped
    s, _ = _run_golden("water6", "th1", 1, thermostat=False)
    assert not six_digit_close(s.state()[3], ref["f"]).all()


# ------------------------------------------------------------ closed forms --------

def _two_beads(r, dv=(0.0, 0.0, 0.0), a=25.0, g=4.5, thermostat=2):
    box = O.Box((4, 4, 4), np.array([[a]]), np.array([[g]]), seed=3)
    x = np.array([[1.5, 1.5, 1.5], [1.5 + r, 1.5, 1.5]])
    v = np.array([[0.0, 0.0, 0.0], list(dv)])
    f, vir, pe = O.forces(box, [1, 2], [0, 0], x, v, step=1, thermostat=thermostat)
    return f, vir, pe


def test_conservative_closed_form():
    for r in (0.1, 0.5, 0.9):
        f, vir, pe = _two_beads(r, g=0.0)
        # force on bead 0 points away from bead 1 (-x): a (1 - r)
        assert np.allclose(f[0], [-25.0 * (1 - r), 0, 0], rtol=1e-14)
        assert np.allclose(f[1], -f[0], rtol=0, atol=0)
        assert np.isclose(vir[0], r * 25.0 * (1 - r), rtol=1e-14)
        assert np.isclose(pe[0], 0.5 * 25.0 * (1 - r) ** 2, rtol=1e-14)
    f, _, _ = _two_beads(1.0, g=0.0)
    assert np.all(f == 0.0)


def test_dissipative_closed_form():
    r, vrel = 0.6, 2.0
    f, _, _ = _two_beads(r, dv=(vrel, 0.0, 0.0), a=0.0)
    # bead 1 recedes along +x: dissipative force pulls bead 0 towards it
    expect = 4.5 * (1 - r) ** 2 * vrel
    assert np.allclose(f[0], [expect, 0, 0], rtol=1e-14)


def test_random_force_moments():
    """sqrt(gamma) w sqrt(24 kT/dt) (0.5 - U): mean 0, variance 2 gamma kT w^2 / dt."""
    r = 0.5
    box = O.Box((4, 4, 4), np.array([[0.0]]), np.array([[4.5]]), seed=11)
    x = np.array([[1.5, 1.5, 1.5], [1.5 + r, 1.5, 1.5]])
    v = np.zeros((2, 3))
    fr = np.array([O.forces(box, [1, 2], [0, 0], x, v, step=t)[0][0, 0] for t in range(1, 4001)])
    var_expect = 2 * 4.5 * 1.0 * (1 - r) ** 2 / 0.02
    assert abs(fr.mean()) < 4 * np.sqrt(var_expect / len(fr))
    assert abs(fr.var() / var_expect - 1) < 0.08


def _random_state(cells, n, ntypes, seed):
    rng = np.random.default_rng(seed)
    L = np.array(cells, dtype=float)
    x = rng.uniform(0, 1, size=(n, 3)) * L
    v = rng.normal(0, 1, size=(n, 3))
    t = rng.integers(0, ntypes, size=n).astype(np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    return ids, t, x, v


def test_cell_list_matches_brute_force():
    cells = (5, 4, 6)
    ids, t, x, v = _random_state(cells, 360, 2, 1)
    box = O.Box(cells, np.array(BINARY_A), np.array(BINARY_G), seed=5)
    f1, vir1, _ = O.forces(box, ids, t, x, v, step=3)
    f2, vir2 = O.pair_forces_brute(box, ids, t, x, v, step=3)
    assert np.allclose(f1, f2, rtol=1e-12, atol=1e-10)
    assert np.allclose(vir1, vir2, rtol=1e-12, atol=1e-9)
    # Newton's third law: total force vanishes
    assert np.abs(f1.sum(axis=0)).max() < 1e-9


def test_bond_closed_form():
    box = O.Box((4, 4, 4), np.array([[0.0]]), np.array([[0.0]]))
    x = np.array([[1.0, 1.0, 1.0], [1.8, 1.0, 1.0]])
    topo = O.Topo(head=[0], tail=[1], k=[128.0], l0=[0.5])
    f, vir, pe = O.forces(box, [1, 2], [0, 0], x, np.zeros((2, 3)), step=1, topo=topo)
    # stretched by 0.3: head pulled towards tail (+x) with k (L - l0)
    assert np.allclose(f[0], [128.0 * 0.3, 0, 0], rtol=1e-12)
    assert np.allclose(f[1], -f[0])
    assert np.isclose(pe[1], 0.5 * 128 * 0.3 ** 2)
    # head bead carries the stress d_x F_x with d = head - tail
    assert np.isclose(vir[0], (-0.8) * (128.0 * 0.3))


def _angle(theta_deg, k=20.0, phi0=0.0):
    """Beads a-b-c with bonds (a->b), (b->c) and bond pair first=(b->c), second=(a->b).
    theta is the bond angle at b (180 = straight)."""
    th = np.radians(theta_deg)
    b = np.array([2.0, 2.0, 2.0])
    a = b + 0.5 * np.array([-1.0, 0.0, 0.0])
    c = b + 0.5 * np.array([-np.cos(th), np.sin(th), 0.0])
    x = np.array([a, b, c])
    topo = O.Topo(head=[0, 1], tail=[1, 2], k=[0.0, 0.0], l0=[0.5, 0.5], first=[1], second=[0],
                  modulus=[k], phi0=[phi0])
    box = O.Box((4, 4, 4), np.array([[0.0]]), np.array([[0.0]]))
    f, _, pe = O.forces(box, [1, 2, 3], [0, 0, 0], x, np.zeros((3, 3)), step=1, topo=topo)
    return x, f, pe[2]


def test_angle_closed_forms():
    # straight chain: no force, zero energy (phi measured from straight)
    _, f, e = _angle(180.0)
    assert np.abs(f).max() < 1e-12 and abs(e) < 1e-12
    # right angle: U = k, force on the end bead is k / |b| perpendicular to its bond
    x, f, e = _angle(90.0)
    assert np.isclose(e, 20.0)
    assert np.allclose(np.linalg.norm(f[2]), 20.0 / 0.5)
    assert np.abs(f.sum(axis=0)).max() < 1e-12


def test_angle_finite_difference_phi0_zero():
    """F = -dU/dx for U = k (1 - cos phi), phi from straight (src/BondPair.cpp:189-193)."""
    x, f, _ = _angle(130.0)
    topo = O.Topo(head=[0, 1], tail=[1, 2], k=[0.0, 0.0], l0=[0.5, 0.5], first=[1], second=[0],
                  modulus=[20.0], phi0=[0.0])
    box = O.Box((4, 4, 4), np.array([[0.0]]), np.array([[0.0]]))
    h = 1e-6
    for bead in range(3):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[bead, d] += h
            xm[bead, d] -= h
            ep = O.forces(box, [1, 2, 3], [0] * 3, xp, np.zeros((3, 3)), 1, topo=topo)[2][2]
            em = O.forces(box, [1, 2, 3], [0] * 3, xm, np.zeros((3, 3)), 1, topo=topo)[2][2]
            assert np.isclose(f[bead, d], -(ep - em) / (2 * h), rtol=1e-6, atol=1e-6)


# --------------------------------------------- bonded cases via the dmpci parser -----

@pytest.mark.parametrize("case", ["chains7", "lipid8", "lipidphi8"])
@pytest.mark.parametrize("name,nsteps", [("nf1", 1), ("th1", 1), ("th10", 10)])
def test_oracle_bonded_matches_reference(case, name, nsteps):
    """Bonds (src/Bond.cpp:111-162) and bond pairs incl. phi0 > 0 (src/BondPair.cpp:138-286),
    topology from the product's dmpci parser, against the reference's snapshots."""
    from dpd_amd import host as H
    inp = H.Input(os.path.join(GOLDEN, case, "case.dmpci"))
    (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
    ids, ty, pid, _ = inp.beads()
    d = load_beads(case, "in")
    assert (d["id"] == ids).all() and (d["type"] == ty).all() and (d["poly"] == pid).all()
    g = inp.gamma if name != "nf1" else np.zeros_like(inp.gamma)
    box = O.Box(inp.cells, inp.a, g, seed=SEEDS[case], rng_mode=O.RNG_LCG)
    s = O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"],
                 topo=O.Topo(h, t, k, l0, fi, se, mo, ph))
    s.run(nsteps)
    x, u, v, f = s.state()
    ref = load_beads(case, name)
    assert six_digit_close(restart_clamp(x, box.lengths), ref["x"]).all()
    assert six_digit_close(u, ref["u"]).all()
    assert six_digit_close(v, ref["v"]).all()
    assert six_digit_close(f, ref["f"]).all()


# ---- slice stress (CMonitor::AddBeadStress, src/Monitor.cpp:1900-2104) ----

def _shares(nslice, Lz, z1, z2, dz):
    """Per-slice share of one interaction, from the oracle's split (F = dx = unit z-ish)."""
    tab = np.zeros((nslice, 1, 9))
    F = np.array([0.0, 0.0, 1.0])
    dx = np.array([0.0, 0.0, dz])
    O.slice_add(tab, Lz, 0, z1, z2, F, dx)
    return tab[:, 0, 8] / dz if abs(dz) > 1e-12 else tab[:, 0, 8]


def test_slice_split_closed_forms():
    # same slice: all of it
    s = _shares(10, 10.0, 3.7, 3.4, 0.3)
    assert np.allclose(s, np.eye(10)[3])
    # adjacent slices, width 1: lower (1 - 0.6)/0.7, upper 0.3/0.7
    s = _shares(10, 10.0, 4.3, 3.6, 0.7)
    want = np.zeros(10); want[3] = 0.4 / 0.7; want[4] = 0.3 / 0.7
    assert np.allclose(s, want, rtol=1e-12)
    # four slices of width 0.25 spanned: inner 0.25/0.7 each, ends by their fractions
    s = _shares(40, 10.0, 4.3, 3.6, 0.7)
    want = np.zeros(40); want[14] = 0.6 * 0.25 / 0.7; want[15] = want[16] = 0.25 / 0.7; want[17] = 0.2 * 0.25 / 0.7
    assert np.allclose(s, want, rtol=1e-12)
    # bead 2 above bead 1 (dz < 0): the same split with the roles swapped
    s2 = _shares(40, 10.0, 3.6, 4.3, -0.7)
    assert np.allclose(s2, want, rtol=1e-12)
    # across the periodic boundary: bead 1 at 0.3 sits above bead 2 at 9.6
    s = _shares(10, 10.0, 0.3, 9.6, 0.7)
    want = np.zeros(10); want[9] = 0.4 / 0.7; want[0] = 0.3 / 0.7
    assert np.allclose(s, want, rtol=1e-12)
    s = _shares(40, 10.0, 0.3, 9.6, 0.7)
    want = np.zeros(40); want[38] = 0.6 * 0.25 / 0.7; want[39] = want[0] = 0.25 / 0.7; want[1] = 0.2 * 0.25 / 0.7
    assert np.allclose(s, want, rtol=1e-12)
    # equal heights on a slice border: half to each side of the border
    tab = np.zeros((10, 1, 9))
    O.slice_add(tab, 10.0, 0, 3.0, 3.0, [1.0, 0.0, 0.0], [0.5, 0.0, 0.0])
    assert np.isclose(tab[3, 0, 0], 0.25) and np.isclose(tab[2, 0, 0], 0.25) and np.isclose(tab.sum(), 0.5)
    # every split sums to one
    rng = np.random.default_rng(4)
    for _ in range(200):
        z1, z2 = rng.uniform(0, 10, 2)
        dz = z1 - z2
        dz -= 10.0 * np.round(dz / 10.0)
        if abs(dz) >= 1.0:
            continue
        for ns in (1, 7, 10, 40):
            assert np.isclose(_shares(ns, 10.0, z1, z2, dz).sum(), 1.0)


def test_slice_stress_sums_to_the_virial():
    """Summed over slices and type pairs, the slice tables are the pair (and bond)
    virial of the same force evaluation: every pair counted once with F (x) dx."""
    cells = (6, 5, 7)
    ids, t, x, v = _random_state(cells, 630, 2, 11)
    box = O.Box(cells, np.array(BINARY_A), np.array(BINARY_G), seed=7)
    pair, bond = O.slice_stress(box, ids, t, x, v, step=4, nslice=14)
    _, vir, _ = O.forces(box, ids, t, x, v, step=4)
    assert np.allclose(pair.sum(axis=(0, 1)), vir, rtol=1e-10, atol=1e-9)
    assert np.abs(bond).max() == 0.0
    # the same system with chains: the bond table carries the bond virial
    n = len(ids)
    head = np.arange(0, 60, 2, dtype=np.int32)
    tail = head + 1
    topo = O.Topo(head=head, tail=tail, k=np.full(30, 64.0), l0=np.full(30, 0.5))
    xb = x.copy()
    xb[tail] = np.mod(x[head] + 0.6, np.array(cells, float))
    pair_b, bond_b = O.slice_stress(box, ids, t, xb, v, step=4, nslice=14, topo=topo)
    _, vir_b, _ = O.forces(box, ids, t, xb, v, step=4, topo=topo)
    _, vir_nb, _ = O.forces(box, ids, t, xb, v, step=4)
    assert np.allclose(bond_b.sum(axis=(0, 1)), vir_b - vir_nb, rtol=1e-10, atol=1e-9)
    assert np.allclose(pair_b.sum(axis=(0, 1)), vir_nb, rtol=1e-10, atol=1e-9)
    # slices see where the pairs are: a profile along z of a uniform fluid is flat
    zz = pair.sum(axis=1)[:, [0, 4, 8]].sum(axis=1)
    assert zz.std() < 0.25 * zz.mean()


# ---- energies (CMonitor::ZeroTotalEnergy, CCNTCell::UpdateTotalEnergy) ----

def _golden_system(case, rng_mode=O.RNG_LCG, thermo=True):
    from dpd_amd import host as H
    inp = H.Input(os.path.join(GOLDEN, case, "case.dmpci"))
    (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
    d = load_beads(case, "in")
    g = inp.gamma if thermo else np.zeros_like(inp.gamma)
    box = O.Box(inp.cells, inp.a, g, seed=SEEDS[case], rng_mode=rng_mode)
    s = O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"],
                 topo=O.Topo(h, t, k, l0, fi, se, mo, ph))
    return inp, box, s


def _history(path):
    return np.array([[float(t) for t in l.split()] for l in open(path) if l.strip() and not l.startswith("#")])


@pytest.mark.parametrize("case", ["water6", "binary6", "chains7", "lipid8", "lipidphi8"])
def test_oracle_energies_match_reference(case):
    """Kinetic and potential energy after each of ten thermostatted steps against the
    reference's ToggleEnergyOutput rows (KE = 1/2 sum v^2; PE = pair 1/2 a (1-r)^2 + bond
    1/2 k (L-l0)^2 + bond-pair terms incl. phi0 > 0, src/BondPair.cpp:320-378)."""
    if case in SIMPLE:
        cells, a, g = SIMPLE[case]
        d = load_beads(case, "in")
        box = O.Box(cells, np.array(a), np.array(g), seed=SEEDS[case], rng_mode=O.RNG_LCG)
        s = O.System(box, d["id"], d["type"], d["x"], d["v"], d["f"], unpbc=d["u"])
    else:
        _, box, s = _golden_system(case)
    rows = _history(os.path.join(GOLDEN, case, "th10_energy.txt"))
    assert len(rows) == 10
    for row in rows:
        s.run(1)
        ke, pe, parts = s.energies()
        assert six_digit_close(np.array([ke, pe, ke + pe]), row[3:6]).all(), (row[:6], ke, pe)


def _phi_state(phi_deg, l1=0.5, l2=0.6):
    """Beads a-b-c whose bond vectors (head - tail) b1 = c-b ... make angle phi: the
    reference's phi is the angle between the two bond vectors (0 = straight chain)."""
    ph = np.radians(phi_deg)
    b = np.array([2.0, 2.0, 2.0])
    a = b - l2 * np.array([1.0, 0.0, 0.0])          # second bond a -> b along +x
    c = b + l1 * np.array([np.cos(ph), np.sin(ph), 0.0])   # first bond b -> c at phi from it
    return np.array([a, b, c])


def _angle_eval(x, k, phi0):
    topo = O.Topo(head=[0, 1], tail=[1, 2], k=[0.0, 0.0], l0=[0.5, 0.5], first=[1], second=[0],
                  modulus=[k], phi0=[phi0])
    box = O.Box((4, 4, 4), np.array([[0.0]]), np.array([[0.0]]))
    f, _, pe = O.forces(box, [1, 2, 3], [0, 0, 0], x, np.zeros((3, 3)), step=1, topo=topo)
    return f, pe[2]


@pytest.mark.parametrize("phi0_deg", [30.0, 60.0])
def test_angle_energy_closed_form_phi0(phi0_deg):
    """For phi0 > 0 the bond-pair energy is k (1 - cos(phi - phi0)) (the reference writes
    it as k (1 -/+ |cos phi| cos phi0 - sin phi sin phi0) on either side of 90 degrees):
    zero at phi = phi0, k (1 - sin phi0) at 90 degrees, k (1 - cos phi0) straight."""
    k, phi0 = 15.0, np.radians(phi0_deg)
    for phi_deg in list(np.linspace(1.0, 179.0, 37)) + [phi0_deg, 90.0]:
        _, e = _angle_eval(_phi_state(phi_deg), k, phi0)
        want = k * (1.0 - np.cos(np.radians(phi_deg) - phi0))
        assert np.isclose(e, want, rtol=1e-12, atol=1e-12), (phi_deg, e, want)
    assert abs(_angle_eval(_phi_state(phi0_deg), k, phi0)[1]) < 1e-12
    assert np.isclose(_angle_eval(_phi_state(90.0), k, phi0)[1], k * (1 - np.sin(phi0)))


@pytest.mark.parametrize("phi_deg", [20.0, 45.0, 75.0])
def test_angle_finite_difference_phi0(phi_deg):
    """Below 90 degrees the phi0 > 0 force is -grad of k (1 - cos(phi - phi0))
    (src/BondPair.cpp:184-188 with prefactor |tan phi| = tan phi there)."""
    k, phi0 = 15.0, np.radians(30.0)
    x = _phi_state(phi_deg)
    f, _ = _angle_eval(x, k, phi0)
    h = 1e-6
    for bead in range(3):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[bead, d] += h
            xm[bead, d] -= h
            ep = _angle_eval(xp, k, phi0)[1]
            em = _angle_eval(xm, k, phi0)[1]
            assert np.isclose(f[bead, d], -(ep - em) / (2 * h), rtol=1e-6, atol=1e-6)


def _commands():
    with open(os.path.join(GOLDEN, "commands", "commands.json")) as fh:
        return json.load(fh)


def apply_command(cmd, sys_, names, state):
    """Apply one command of the fixture schedule to an oracle System or an engine wrapper
    with the same methods (set_timestep, set_halfstep, set_gamma, set_a)."""
    tok = cmd.split()
    name = tok[0]
    nt = len(names)
    if name == "SetTimeStepSize":
        state["dt"] = float(tok[2])
        sys_.set_timestep(state["dt"])
    elif name == "ToggleDPDBeadThermostat":
        state["thermo"] = not state["thermo"]
        sys_.set_gamma(state["g"] if state["thermo"] else np.zeros((nt, nt)))
    elif name == "ToggleDPDBeadConservativeForces":
        state["cons"] = not state["cons"]
        sys_.set_a(state["a"] if state["cons"] else np.zeros((nt, nt)))
    elif name == "ToggleDPDBeadForces":
        state["forces"] = not state["forces"]
        sys_.set_halfstep(state["dt"] if state["forces"] else 0.0)
    elif name in ("SetDPDBeadConsInt", "SetDPDBeadDissInt"):
        i, j, val = names.index(tok[2]), names.index(tok[3]), float(tok[4])
        key = "a" if name == "SetDPDBeadConsInt" else "g"
        state[key][i, j] = state[key][j, i] = val
        on = state["cons"] if key == "a" else state["thermo"]
        if on:
            (sys_.set_a if key == "a" else sys_.set_gamma)(state[key])
    else:
        raise ValueError(cmd)


def test_oracle_commands_match_reference():
    """Time step changes, thermostat / conservative / force toggles and interaction changes
    (src/CNTCell.cpp:223-377, :434-456) executed before the step of their time, against the
    reference's run of the same schedule: every step's kT and P, and the final snapshot."""
    spec = _commands()
    inp, box, s = _golden_system(spec["case"])
    state = {"dt": box.dt, "thermo": True, "cons": True, "forces": True, "a": inp.a.copy(), "g": inp.gamma.copy()}
    sched = {}
    for t, c in spec["commands"]:
        sched.setdefault(t, []).append(c)
    rows = _history(os.path.join(GOLDEN, "commands", "cmd_history.txt"))
    for t in range(1, spec["steps"] + 1):
        for c in sched.get(t, []):
            apply_command(c, s, inp.bead_names, state)
        s.run(1)
        ob = s.observables()
        assert six_digit_close(np.array([ob["kT"], ob["pressure"]]), rows[t - 1][1:3]).all(), (t, ob["kT"], rows[t - 1])
    x, u, v, f = s.state()
    ref = load_beads("commands", "cmd%d" % spec["steps"])
    assert six_digit_close(restart_clamp(x, box.lengths), ref["x"]).all()
    assert six_digit_close(u, ref["u"]).all()
    assert six_digit_close(v, ref["v"]).all()
    assert six_digit_close(f, ref["f"]).all()


# ---- bond-pair slice stress (CMonitor::AddBondPairStress) ----

def _angle_only(cells, ntypes, ids, types, x, topo):
    """Forces of the bond pairs alone (a = gamma = 0, bond springs of zero strength)."""
    z = np.zeros((ntypes, ntypes))
    box = O.Box(cells, z, z, seed=1)
    t0 = O.Topo(topo.head, topo.tail, np.zeros(len(topo.head)), topo.l0, topo.first, topo.second,
                topo.modulus, topo.phi0)
    f, _, _ = O.forces(box, ids, types, x, np.zeros_like(x), step=1, topo=t0)
    return box, t0, f


def test_bondpair_slices_closed_form():
    """One bond pair a - b - c (types 0, 1, 2) spanning three of four slices (width 2): the
    pair c -> b lies in slice 0; b -> a spans 1.5 .. 5.0 (slice 0 gets 0.5/3.5 of it,
    slice 1 2/3.5, slice 2 1/3.5); a -> c spans 1.0 .. 5.0 (1/4, 2/4, 1/4).  Each pair adds
    (F[j+1] - F[j]) / 3 (x) (u[j+1] - u[j]), F the bond pair's force on each bead."""
    cells = (8, 8, 8)
    x = np.array([[3.0, 4.0, 5.0],      # a: second bond's head
                  [3.4, 3.8, 1.5],      # b: shared (first bond's head)
                  [3.9, 3.3, 1.0]])     # c: first bond's tail
    ids = np.array([1, 2, 3], np.int32)
    types = np.array([0, 1, 2], np.int32)
    topo = O.Topo(head=[0, 1], tail=[1, 2], k=[0.0, 0.0], l0=[0.5, 0.5], first=[1], second=[0],
                  modulus=[17.0], phi0=[np.radians(40.0)])
    box, t0, f = _angle_only(cells, 3, ids, types, x, topo)
    assert np.abs(f).max() > 1.0
    tab = O.slice_stress_bondpairs(box, ids, types, x, 4, topo=t0)
    idx = (0 * 3 + 1) * 3 + 2
    assert np.abs(np.delete(tab, idx, axis=1)).max() == 0.0

    def T(p, q):   # bead p -> bead q
        return np.outer(0.33333333 * (f[q] - f[p]), x[q] - x[p]).ravel()
    a, b, c = 0, 1, 2
    want = np.zeros((4, 9))
    want[0] = T(c, b) + (0.5 / 3.5) * T(b, a) + 0.25 * T(a, c)
    want[1] = (2.0 / 3.5) * T(b, a) + 0.5 * T(a, c)
    want[2] = (1.0 / 3.5) * T(b, a) + 0.25 * T(a, c)
    assert np.allclose(tab[:, idx, :], want, rtol=1e-12, atol=1e-12)
    # a plausible mistake fails: the factor of a third dropped, or the pair order reversed
    # in the force but not in the separation
    assert not np.allclose(tab[:, idx, :], 3 * want)
    bad = np.zeros((4, 9))
    bad[0] = -T(c, b)
    assert not np.allclose(tab[:, idx, :], want + 2 * bad)


def test_bondpair_slices_sum_to_the_bondpair_virial():
    """Summed over slices and types, the table is the bond pairs' virial sum_i F_i (x) u_i
    (times 0.99999999, the reference's 0.33333333 times three) -- on the reference's branched
    lipid snapshot with its bond pairs (phi0 = 0, 30 and 60 degrees)."""
    from dpd_amd import host as H
    inp = H.Input(os.path.join(GOLDEN, "lipidphi8", "case.dmpci"))
    (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
    d = load_beads("lipidphi8", "in")
    topo = O.Topo(h, t, k, l0, fi, se, mo, ph)
    # wrapped positions consistent with the unwrapped ones (the snapshot's writer moved beads
    # within 0.001 of a face without moving their unwrapped coordinate; the slice split
    # takes heights from the one and separations from the other)
    x = np.mod(d["u"], np.array(inp.cells, float))
    box, t0, f = _angle_only(inp.cells, len(inp.a), d["id"], d["type"], x, topo)
    tab = O.slice_stress_bondpairs(box, d["id"], d["type"], x, 16, unpbc=d["u"], topo=t0)
    f, _, _ = O.forces(box, d["id"], d["type"], x, np.zeros_like(x), step=1, unpbc=d["u"], topo=t0)
    vir = np.einsum("ia,ib->ab", f, d["u"]).ravel()
    assert np.allclose(tab.sum(axis=(0, 1)), 0.99999999 * vir, rtol=1e-9, atol=1e-9)
    assert np.abs(vir).max() > 1.0
    # the profile sees where the lipids are
    assert (np.abs(tab).sum(axis=(1, 2)) > 0).sum() >= 8
