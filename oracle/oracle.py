"""ctypes binding of the FP64 CPU oracle (oracle/dpd_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing measured.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

RNG_LCG = 0
RNG_PHILOX = 1


class Params(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("wx", C.c_double), ("wy", C.c_double), ("wz", C.c_double),
                ("dt", C.c_double), ("lam", C.c_double), ("kT", C.c_double),
                ("seed", C.c_uint64), ("ntypes", C.c_int32),
                ("a", C.POINTER(C.c_double)), ("g", C.POINTER(C.c_double)),
                ("rng_mode", C.c_int32)]


class Topology(C.Structure):
    _fields_ = [("nbonds", C.c_int32), ("head", C.POINTER(C.c_int32)),
                ("tail", C.POINTER(C.c_int32)), ("k", C.POINTER(C.c_double)),
                ("l0", C.POINTER(C.c_double)), ("npairs", C.c_int32),
                ("first", C.POINTER(C.c_int32)), ("second", C.POINTER(C.c_int32)),
                ("modulus", C.POINTER(C.c_double)), ("phi0", C.POINTER(C.c_double))]


class Obs(C.Structure):
    _fields_ = [("kT", C.c_double), ("pressure", C.c_double), ("stress", C.c_double * 9),
                ("sum_v2", C.c_double * 3), ("momentum", C.c_double * 3),
                ("pe_pair", C.c_double), ("pe_bond", C.c_double), ("pe_angle", C.c_double),
                ("wrong_cell", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: run `make -C oracle` ({LIB_PATH})")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Topology)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
        L.orc_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_observables.argtypes = [C.c_void_p, C.POINTER(Obs)]
        L.orc_set_gamma.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_a.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_timestep.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_halfstep.argtypes = [C.c_void_p, C.c_double]
        L.orc_forces.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.POINTER(Topology), C.c_int64, C.c_int32,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_pair_forces_brute.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                            C.c_void_p, C.c_void_p]
        L.orc_slice_add.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_double,
                                    C.c_double, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_slice_stress.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.POINTER(Topology), C.c_int64, C.c_int32,
                                       C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_slice_stress_bondpairs.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.POINTER(Topology), C.c_int32,
                                                 C.c_void_p]
        L.orc_energy_output.argtypes = [C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.POINTER(Topology), C.c_void_p]
        L.orc_lcg_uniform.restype = C.c_double
        L.orc_lcg_uniform.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_pair_uniform.restype = C.c_double
        L.orc_pair_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_int32, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Box:
    """Box/thermostat parameters in reference units (dmpci keywords)."""

    def __init__(self, cells, a, g, dt=0.02, lam=0.5, kT=1.0, seed=1, widths=(1.0, 1.0, 1.0),
                 rng_mode=RNG_PHILOX):
        self.cells = tuple(int(c) for c in cells)
        self.widths = tuple(float(w) for w in widths)
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.g = np.ascontiguousarray(g, dtype=np.float64)
        self.ntypes = self.a.shape[0]
        self.dt, self.lam, self.kT, self.seed = dt, lam, kT, seed
        self.rng_mode = rng_mode

    def params(self):
        p = Params()
        p.nx, p.ny, p.nz = self.cells
        p.wx, p.wy, p.wz = self.widths
        p.dt, p.lam, p.kT = self.dt, self.lam, self.kT
        p.seed = self.seed
        p.ntypes = self.ntypes
        p.a = self.a.ctypes.data_as(C.POINTER(C.c_double))
        p.g = self.g.ctypes.data_as(C.POINTER(C.c_double))
        p.rng_mode = self.rng_mode
        return p

    @property
    def lengths(self):
        return np.array(self.cells, dtype=np.float64) * np.array(self.widths)


class Topo:
    """Bonds by bead index, bond pairs by bond index (phi0 in radians)."""

    def __init__(self, head=(), tail=(), k=(), l0=(), first=(), second=(), modulus=(), phi0=()):
        self.head = np.ascontiguousarray(head, dtype=np.int32)
        self.tail = np.ascontiguousarray(tail, dtype=np.int32)
        self.k = np.ascontiguousarray(k, dtype=np.float64)
        self.l0 = np.ascontiguousarray(l0, dtype=np.float64)
        self.first = np.ascontiguousarray(first, dtype=np.int32)
        self.second = np.ascontiguousarray(second, dtype=np.int32)
        self.modulus = np.ascontiguousarray(modulus, dtype=np.float64)
        self.phi0 = np.ascontiguousarray(phi0, dtype=np.float64)

    def struct(self):
        t = Topology()
        t.nbonds = len(self.head)
        t.head = self.head.ctypes.data_as(C.POINTER(C.c_int32))
        t.tail = self.tail.ctypes.data_as(C.POINTER(C.c_int32))
        t.k = self.k.ctypes.data_as(C.POINTER(C.c_double))
        t.l0 = self.l0.ctypes.data_as(C.POINTER(C.c_double))
        t.npairs = len(self.first)
        t.first = self.first.ctypes.data_as(C.POINTER(C.c_int32))
        t.second = self.second.ctypes.data_as(C.POINTER(C.c_int32))
        t.modulus = self.modulus.ctypes.data_as(C.POINTER(C.c_double))
        t.phi0 = self.phi0.ctypes.data_as(C.POINTER(C.c_double))
        return t


class System:
    """Stateful oracle run (reference step order)."""

    def __init__(self, box, ids, types, x, v, f=None, unpbc=None, topo=None):
        L = lib()
        self.n = len(ids)
        self._keep = [np.ascontiguousarray(a, dtype=dt) if a is not None else None
                      for a, dt in ((ids, np.int32), (types, np.int32), (x, np.float64),
                                    (unpbc, np.float64), (v, np.float64), (f, np.float64))]
        self.box = box
        self.dt = box.dt
        self._p = box.params()
        self._t = (topo or Topo()).struct()
        self._topo = topo or Topo()
        ids_, types_, x_, u_, v_, f_ = self._keep
        self.h = L.orc_create(C.byref(self._p), self.n, _p(ids_), _p(types_), _p(x_), _p(u_), _p(v_),
                              _p(f_), C.byref(self._t))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def run(self, nsteps, step0=0):
        lib().orc_run(self.h, int(nsteps), int(step0))

    def set_gamma(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        lib().orc_set_gamma(self.h, _p(g))

    def set_a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_set_a(self.h, _p(a))
        self._a_now = a.copy()

    def set_timestep(self, dt):
        """SetTimeStepSize: CCNTCell::ChangeTimeStepConstants."""
        lib().orc_set_timestep(self.h, float(dt))
        self.dt = float(dt)

    def set_halfstep(self, dt):
        """ToggleDPDBeadForces: the half-step constants (dt = 0 switches the forces off)."""
        lib().orc_set_halfstep(self.h, float(dt))

    def _box_now(self):
        b = self.box
        if getattr(self, "_a_now", None) is not None:
            b = Box(b.cells, self._a_now, b.g, dt=b.dt, lam=b.lam, kT=b.kT, seed=b.seed, widths=b.widths,
                    rng_mode=b.rng_mode)
        return b

    def energies(self):
        """(KE, PE, parts) as the reference's energy output writes them after a step
        (CMonitor::ZeroTotalEnergy / CCNTCell::UpdateTotalEnergy, src/Monitor.cpp:2755-2778):
        KE = 1/2 sum v^2; PE = pair + bond + bond-pair energy at the current positions, the
        pair part summed the reference's way (energy_output)."""
        x, u, v, _ = self.state()
        pe = energy_output(self._box_now(), self._keep[0], self._keep[1], x, unpbc=u, topo=self._topo)
        return 0.5 * float((v * v).sum()), float(pe.sum()), pe

    def potential_energies(self, step=0):
        """The true pair, bond and bond-pair energies at the current positions (every pair
        once)."""
        x, u, v, _ = self.state()
        _, _, pe = forces(self._box_now(), self._keep[0], self._keep[1], x, v, step, thermostat=False, unpbc=u,
                          topo=self._topo)
        return pe

    def state(self):
        x = np.empty((self.n, 3)); u = np.empty((self.n, 3))
        v = np.empty((self.n, 3)); f = np.empty((self.n, 3))
        lib().orc_get(self.h, _p(x), _p(u), _p(v), _p(f))
        return x, u, v, f

    def observables(self):
        o = Obs()
        lib().orc_observables(self.h, C.byref(o))
        return {"kT": o.kT, "pressure": o.pressure, "stress": np.array(o.stress[:]),
                "sum_v2": np.array(o.sum_v2[:]), "momentum": np.array(o.momentum[:]),
                "wrong_cell": o.wrong_cell}


def forces(box, ids, types, x, v, step, thermostat=True, unpbc=None, topo=None):
    """One FP64 force evaluation (Philox pair stream).  Returns f[n,3], virial[9], pe[3].
    thermostat: True/1 = dissipative + random, False/0 = neither, 2 = dissipative only."""
    L = lib()
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    u = np.ascontiguousarray(unpbc, dtype=np.float64) if unpbc is not None else None
    p = box.params()
    t = (topo or Topo())
    ts = t.struct()
    f = np.zeros((n, 3)); vir = np.zeros(9); pe = np.zeros(3)
    L.orc_forces(C.byref(p), n, _p(ids), _p(types), _p(x), _p(u), _p(v), C.byref(ts), int(step),
                 int(thermostat), _p(f), _p(vir), _p(pe))
    return f, vir, pe


def slice_stress(box, ids, types, x, v, step, nslice, thermostat=True, unpbc=None, topo=None):
    """The slice stress tables of one force evaluation (CMonitor::AddBeadStress /
    AddBondStress): pair[nslice, ntypes*ntypes, 9] in the reference's half-shell order
    (type of the bead the force acts on first), bond[nslice, ntypes*ntypes, 9] keyed by
    (head type, tail type)."""
    L = lib()
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    u = np.ascontiguousarray(unpbc, dtype=np.float64) if unpbc is not None else None
    p = box.params()
    t = (topo or Topo())
    ts = t.struct()
    nt = box.ntypes
    pair = np.zeros((nslice, nt * nt, 9))
    bond = np.zeros((nslice, nt * nt, 9))
    L.orc_slice_stress(C.byref(p), n, _p(ids), _p(types), _p(x), _p(u), _p(v), C.byref(ts), int(step),
                       int(thermostat), int(nslice), _p(pair), _p(bond))
    return pair, bond


def slice_stress_bondpairs(box, ids, types, x, nslice, unpbc=None, topo=None):
    """The bond-pair slice stress table (CMonitor::AddBondPairStress): [nslice,
    ntypes**3, 9] keyed by the bead types (second head, first head, first tail)."""
    L = lib()
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.ascontiguousarray(unpbc, dtype=np.float64) if unpbc is not None else None
    p = box.params()
    ts = (topo or Topo()).struct()
    nt = box.ntypes
    tab = np.zeros((nslice, nt ** 3, 9))
    L.orc_slice_stress_bondpairs(C.byref(p), n, _p(ids), _p(types), _p(x), _p(u), C.byref(ts), int(nslice), _p(tab))
    return tab


def slice_add(tab, Lz, idx, z1, z2, F, dx, wgt=1.0):
    """One interaction's z-slice split into tab[nslice, nidx, 9] (in place)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    lib().orc_slice_add(_p(tab), tab.shape[0], float(Lz), int(idx), tab.shape[1], float(z1), float(z2), _p(F),
                        _p(dx), float(wgt))
    return tab


def energy_output(box, ids, types, x, unpbc=None, topo=None):
    """[pair, bond, bond-pair] potential energy as the reference's ToggleEnergyOutput rows
    sum it (pairs across cells only when the neighbour's id is larger)."""
    L = lib()
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.ascontiguousarray(unpbc, dtype=np.float64) if unpbc is not None else None
    p = box.params()
    ts = (topo or Topo()).struct()
    out = np.zeros(3)
    L.orc_energy_output(C.byref(p), n, _p(ids), _p(types), _p(x), _p(u), C.byref(ts), _p(out))
    return out


def pair_forces_brute(box, ids, types, x, v, step, thermostat=True):
    L = lib()
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = box.params()
    f = np.zeros((n, 3)); vir = np.zeros(9)
    L.orc_pair_forces_brute(C.byref(p), n, _p(ids), _p(types), _p(x), _p(v), int(step),
                            int(thermostat), _p(f), _p(vir))
    return f, vir


def lcg_sequence(seed, n):
    L = lib()
    s = C.c_uint64(seed)
    return np.array([L.orc_lcg_uniform(C.byref(s)) for _ in range(n)])


def philox4x32_10(ctr, key):
    """ctr: 4 words, key: 2 words -> 4 words"""
    L = lib()
    out = (C.c_uint32 * 4)()
    L.orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return tuple(out)


def pair_uniform(seed, step, i, j):
    return lib().orc_pair_uniform(seed, step, i, j)
