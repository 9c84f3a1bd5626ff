"""Python binding of the MI355X DPD engine (ctypes over include/dpd_engine.h).

The HIP library lib/libospreydpd_amd.so is required: importing works without it, but
every call that needs it raises ``EngineMissing``; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libospreydpd_amd.so")

OK = 0
ERR_NAMES = {1: "PARAM", 2: "OOM", 3: "HIP", 4: "NAN", 5: "OUT_OF_BOX", 6: "STATE", 7: "IO",
             8: "UNSUPPORTED"}
(PAIR_AUTO, PAIR_BRICK, PAIR_FULLSHELL, PAIR_BRICK512, PAIR_STREAM, PAIR_STREAM1024, PAIR_COMPACT, PAIR_CELLS,
 PAIR_HITLIST) = range(9)
PAIR_KINDS = 8
# the pair kernels the library carries (ids 1, 3-7 name retired kernels, refused with
# ERR_UNSUPPORTED) and their device kernel names
PAIR_AVAILABLE = (PAIR_FULLSHELL, PAIR_HITLIST)
PAIR_KERNEL_NAMES = {PAIR_FULLSHELL: "k_pair_fullshell", PAIR_HITLIST: "k_pair_hitlist"}
STEP_OBSERVE, NO_THERMOSTAT, STEP_TIMING, STEP_GRAPH, STEP_TIMING_SAMPLED = 1, 2, 4, 8, 16
TOGGLE_THERMOSTAT, TOGGLE_CONSERVATIVE, TOGGLE_FORCES = 1, 2, 3


class EngineMissing(RuntimeError):
    pass


class DpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dpd error {ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("wx", C.c_double), ("wy", C.c_double), ("wz", C.c_double),
                ("dt", C.c_double), ("lam", C.c_double), ("kT", C.c_double),
                ("seed", C.c_uint64), ("ntypes", C.c_int32),
                ("a", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double)),
                ("device", C.c_int32), ("pair_kernel", C.c_int32)]


class Obs(C.Structure):
    _fields_ = [("step", C.c_int64), ("kT", C.c_double), ("pressure", C.c_double),
                ("stress", C.c_double * 9), ("sum_v2", C.c_double * 3),
                ("momentum", C.c_double * 3), ("pe_pair", C.c_double), ("pe_bond", C.c_double),
                ("pe_angle", C.c_double), ("jumps", C.c_int64), ("nbeads", C.c_int64)]


# exported symbols of include/dpd_engine.h: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "dpd_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "dpd_destroy": (None, [_P]),
    "dpd_last_error": (C.c_char_p, [_P]),
    "dpd_upload_beads": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, _P, _P, C.c_int64]),
    "dpd_upload_topology": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, C.c_int64, _P, _P, _P, _P]),
    "dpd_step": (C.c_int, [_P, C.c_int64, C.c_uint32]),
    "dpd_compute_forces": (C.c_int, [_P, C.c_int64, C.c_uint32, _P, _P, _P]),
    "dpd_observables": (C.c_int, [_P, C.POINTER(Obs)]),
    "dpd_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "dpd_set_interactions": (C.c_int, [_P, _P, _P]),
    "dpd_set_timestep": (C.c_int, [_P, C.c_double]),
    "dpd_toggle": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "dpd_last_timing": (C.c_int64, [_P, _P]),
    "dpd_get_step": (C.c_int64, [_P]),
    "dpd_time_pair": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "dpd_set_pair_kernel": (C.c_int, [_P, C.c_int32]),
    "dpd_get_pair_kernel": (C.c_int32, [_P]),
    "dpd_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _P]),
    "dpd_schedule_count": (C.c_int32, []),
    "dpd_schedule_name": (C.c_char_p, [C.c_int32]),
    "dpd_set_schedule": (C.c_int, [_P, C.c_int32]),
    "dpd_get_schedule": (C.c_int32, [_P]),
    "dpd_autotune_schedules": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _P]),
    "dpd_apply_tuning": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int32)]),
    "dpd_create_batch": (C.c_int, [C.c_int32, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "dpd_upload_box": (C.c_int, [_P, C.c_int32, C.c_int64, _P, _P, _P, _P, _P, _P]),
    "dpd_upload_box_topology": (C.c_int, [_P, C.c_int32, C.c_int64, _P, _P, _P, _P, C.c_int64, _P, _P,
                                          _P, _P]),
    "dpd_commit": (C.c_int, [_P, C.c_int64]),
    "dpd_box_count": (C.c_int32, [_P]),
    "dpd_observables_box": (C.c_int, [_P, C.c_int32, C.POINTER(Obs)]),
    "dpd_download_box": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    "dpd_set_box_interactions": (C.c_int, [_P, C.c_int32, _P, _P]),
    "dpd_slice_stress": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P]),
    "dpd_slice_stress_bondpairs": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "dpd_slice_stress_noalloc": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P]),
    "dpd_energy_output": (C.c_int, [_P, C.c_int32, _P]),
    "dpd_zero_total_momentum": (C.c_int, [_P, C.c_int32]),
    "dpd_renormalise_momenta": (C.c_int, [_P, C.c_int32]),
}

_lib = None


def lib():
    """Load the HIP library (raises EngineMissing if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineMissing(f"{LIB_PATH} not built: run __graft_entry__.build() or "
                                f"make -C osprey-dpd_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Engine:
    """One DPD box on one GPU."""

    def __init__(self, cells, a, gamma, dt=0.02, lam=0.5, kT=1.0, seed=1, widths=(1.0, 1.0, 1.0),
                 device=0, pair_kernel=PAIR_AUTO):
        L = lib()
        self._a = np.ascontiguousarray(a, dtype=np.float64)
        self._g = np.ascontiguousarray(gamma, dtype=np.float64)
        p = Params()
        p.nx, p.ny, p.nz = (int(c) for c in cells)
        p.wx, p.wy, p.wz = (float(w) for w in widths)
        p.dt, p.lam, p.kT = dt, lam, kT
        p.seed = int(seed)
        p.ntypes = self._a.shape[0]
        p.a = self._a.ctypes.data_as(C.POINTER(C.c_double))
        p.gamma = self._g.ctypes.data_as(C.POINTER(C.c_double))
        p.device = device
        p.pair_kernel = pair_kernel
        h = C.c_void_p()
        rc = L.dpd_create(C.byref(p), C.byref(h))
        self.h = h
        if rc != OK:
            msg = L.dpd_last_error(h).decode() if h else ""
            self.close()
            raise DpdError(rc, msg)
        self.n = 0
        self.cells = tuple(int(c) for c in cells)
        self.widths = tuple(float(w) for w in widths)

    def _check(self, rc):
        if rc != OK:
            raise DpdError(rc, lib().dpd_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().dpd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, ids, types, x, v, f=None, unpbc=None, step0=0):
        n = len(ids)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        types = np.ascontiguousarray(types, dtype=np.int32)
        x, v, f, u = _f64(x, (n, 3)), _f64(v, (n, 3)), _f64(f, (n, 3)), _f64(unpbc, (n, 3))
        self._check(lib().dpd_upload_beads(self.h, n, _ptr(ids), _ptr(types), _ptr(x), _ptr(u),
                                           _ptr(v), _ptr(f), int(step0)))
        self.n = n

    def upload_topology(self, head=(), tail=(), k=(), l0=(), first=(), second=(), modulus=(), phi0=()):
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        head, tail, first, second = i32(head), i32(tail), i32(first), i32(second)
        k, l0, modulus, phi0 = _f64(k), _f64(l0), _f64(modulus), _f64(phi0)
        self._check(lib().dpd_upload_topology(self.h, len(head), _ptr(head), _ptr(tail), _ptr(k),
                                              _ptr(l0), len(first), _ptr(first), _ptr(second),
                                              _ptr(modulus), _ptr(phi0)))

    def step(self, nsteps=1, observe=False, timing=False, graph=False):
        """timing: False, True (every pair launch timed) or "sampled" (every 16th)."""
        flags = (STEP_OBSERVE if observe else 0) | (STEP_TIMING if timing else 0) | \
            (STEP_TIMING_SAMPLED if timing == "sampled" else 0) | (STEP_GRAPH if graph else 0)
        self._check(lib().dpd_step(self.h, int(nsteps), flags))

    def compute_forces(self, step, thermostat=True):
        f = np.zeros((self.n, 3))
        vir = np.zeros(9)
        pe = np.zeros(3)
        flags = 0 if thermostat else NO_THERMOSTAT
        self._check(lib().dpd_compute_forces(self.h, int(step), flags, _ptr(f), _ptr(vir), _ptr(pe)))
        return f, vir, pe

    def observables(self, box=0):
        o = Obs()
        self._check(lib().dpd_observables_box(self.h, int(box), C.byref(o)))
        return {"step": o.step, "kT": o.kT, "pressure": o.pressure, "stress": np.array(o.stress[:]),
                "sum_v2": np.array(o.sum_v2[:]), "momentum": np.array(o.momentum[:]),
                "pe": np.array([o.pe_pair, o.pe_bond, o.pe_angle]), "jumps": o.jumps,
                "nbeads": o.nbeads}

    def download(self):
        n = self.n
        x, u, v, f = (np.empty((n, 3)) for _ in range(4))
        self._check(lib().dpd_download(self.h, _ptr(x), _ptr(u), _ptr(v), _ptr(f)))
        return x, u, v, f

    def set_interactions(self, a=None, gamma=None):
        a, g = _f64(a), _f64(gamma)
        self._check(lib().dpd_set_interactions(self.h, _ptr(a), _ptr(g)))

    def set_timestep(self, dt):
        self._check(lib().dpd_set_timestep(self.h, float(dt)))

    def toggle(self, what, on):
        self._check(lib().dpd_toggle(self.h, int(what), int(bool(on))))

    def last_timing(self):
        ms = np.zeros(3)
        n = lib().dpd_last_timing(self.h, _ptr(ms))
        return int(n), ms

    def time_pair(self, nlaunch=20):
        """Average duration (ms) of the step's pair-force launch on the current state, from
        `nlaunch` back-to-back launches between two HIP events (dpd_time_pair)."""
        ms = C.c_double()
        self._check(lib().dpd_time_pair(self.h, int(nlaunch), C.byref(ms)))
        return ms.value

    @property
    def step_count(self):
        return lib().dpd_get_step(self.h)

    @property
    def pair_kernel(self):
        return int(lib().dpd_get_pair_kernel(self.h))

    def set_pair_kernel(self, kernel):
        self._check(lib().dpd_set_pair_kernel(self.h, int(kernel)))

    def autotune(self, nsteps=10):
        """Time every pair kernel on the current state; keep the fastest.  Returns the
        choice and {kernel: ms per step}."""
        chosen = C.c_int32()
        ms = np.zeros(PAIR_KINDS)
        self._check(lib().dpd_autotune(self.h, int(nsteps), C.byref(chosen), _ptr(ms)))
        return chosen.value, {k + 1: float(ms[k]) for k in range(PAIR_KINDS)}

    @property
    def schedule(self):
        return int(lib().dpd_get_schedule(self.h))

    def set_schedule(self, sched):
        self._check(lib().dpd_set_schedule(self.h, int(sched)))

    def autotune_schedules(self, nsteps=10):
        """Time every pair-launch schedule on the current state; keep the fastest.  Returns
        the choice and {schedule name: ms per step}."""
        chosen = C.c_int32()
        n = schedule_count()
        ms = np.zeros(n)
        self._check(lib().dpd_autotune_schedules(self.h, int(nsteps), C.byref(chosen), _ptr(ms)))
        return chosen.value, {schedule_name(k): float(ms[k]) for k in range(n)}

    def apply_tuning(self, path):
        """Set the schedule the tuning file records for this engine's shape (dpd_apply_tuning);
        returns it, or None when the file has none."""
        chosen = C.c_int32()
        self._check(lib().dpd_apply_tuning(self.h, str(path).encode(), C.byref(chosen)))
        return chosen.value if chosen.value >= 0 else None

    def zero_total_momentum(self, box=-1):
        self._check(lib().dpd_zero_total_momentum(self.h, int(box)))

    def renormalise_momenta(self, box=-1):
        self._check(lib().dpd_renormalise_momenta(self.h, int(box)))

    def _ntypes(self, box):
        return self._a.shape[0]

    def slice_stress(self, nslices, box=0, bonds=True):
        """Stress profile over `nslices` z-slices of the current state (CMonitor::AddBeadStress
        / AddBondStress, src/Monitor.cpp:1900-2104, :2283-2518): (pair, bond) tables of shape
        [nslices, ntypes*ntypes, 9], element 3a + b = F_a dx_b; bond is None when not asked."""
        nt = self._ntypes(box)
        pair = np.zeros((int(nslices), nt * nt, 9))
        bond = np.zeros((int(nslices), nt * nt, 9)) if bonds else None
        self._check(lib().dpd_slice_stress(self.h, int(box), int(nslices), _ptr(pair), _ptr(bond)))
        return pair, bond

    def energy_output(self, box=0):
        """[pair, bond, bond-pair] potential energy as the reference's energy output sums it
        (cross-cell pairs only when the neighbour's id is larger)."""
        out = np.zeros(3)
        self._check(lib().dpd_energy_output(self.h, int(box), _ptr(out)))
        return out

    def slice_stress_bondpairs(self, nslices, box=0):
        """Bond-pair stress profile (CMonitor::AddBondPairStress, src/Monitor.cpp:2520-2739):
        [nslices, ntypes**3, 9] keyed by the bead types (second head, first head, first tail)."""
        return _slice_bondpairs(self, nslices, box)


def schedule_count():
    return int(lib().dpd_schedule_count())


def schedule_name(k):
    name = lib().dpd_schedule_name(int(k))
    return name.decode() if name else None


def _slice_bondpairs(eng, nslices, box):
    nt = eng._ntypes(box)
    tab = np.zeros((int(nslices), nt ** 3, 9))
    eng._check(lib().dpd_slice_stress_bondpairs(eng.h, int(box), int(nslices), _ptr(tab)))
    return tab


def _params(cells, a, gamma, dt, lam, kT, seed, widths, device, pair_kernel, keep):
    a = np.ascontiguousarray(a, dtype=np.float64)
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    keep += [a, g]
    p = Params()
    p.nx, p.ny, p.nz = (int(c) for c in cells)
    p.wx, p.wy, p.wz = (float(w) for w in widths)
    p.dt, p.lam, p.kT = dt, lam, kT
    p.seed = int(seed)
    p.ntypes = a.shape[0]
    p.a = a.ctypes.data_as(C.POINTER(C.c_double))
    p.gamma = g.ctypes.data_as(C.POINTER(C.c_double))
    p.device = device
    p.pair_kernel = pair_kernel
    return p


class Batch(Engine):
    """Many independent boxes advanced together (each kernel spans all boxes).

    boxes: list of dicts with keys cells, a, gamma and optional dt, lam, kT, seed, widths.
    """

    def __init__(self, boxes, device=0, pair_kernel=PAIR_AUTO):
        L = lib()
        keep = []
        arr = (Params * len(boxes))()
        for i, b in enumerate(boxes):
            arr[i] = _params(b["cells"], b["a"], b["gamma"], b.get("dt", 0.02), b.get("lam", 0.5),
                             b.get("kT", 1.0), b.get("seed", 1), b.get("widths", (1.0, 1.0, 1.0)),
                             device, pair_kernel, keep)
        h = C.c_void_p()
        rc = L.dpd_create_batch(len(boxes), arr, C.byref(h))
        self.h = h
        if rc != OK:
            msg = L.dpd_last_error(h).decode() if h else ""
            self.close()
            raise DpdError(rc, msg)
        self.nbox = len(boxes)
        self.box_ntypes = [np.asarray(b["a"]).shape[0] for b in boxes]
        self.box_n = [0] * self.nbox
        self.n = 0

    def _ntypes(self, box):
        return self.box_ntypes[box]

    def upload_box(self, box, ids, types, x, v, f=None, unpbc=None):
        n = len(ids)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        types = np.ascontiguousarray(types, dtype=np.int32)
        x, v, f, u = _f64(x, (n, 3)), _f64(v, (n, 3)), _f64(f, (n, 3)), _f64(unpbc, (n, 3))
        self._check(lib().dpd_upload_box(self.h, int(box), n, _ptr(ids), _ptr(types), _ptr(x), _ptr(u),
                                         _ptr(v), _ptr(f)))
        self.box_n[box] = n

    def upload_box_topology(self, box, head=(), tail=(), k=(), l0=(), first=(), second=(), modulus=(),
                            phi0=()):
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        head, tail, first, second = i32(head), i32(tail), i32(first), i32(second)
        k, l0, modulus, phi0 = _f64(k), _f64(l0), _f64(modulus), _f64(phi0)
        self._check(lib().dpd_upload_box_topology(self.h, int(box), len(head), _ptr(head), _ptr(tail),
                                                  _ptr(k), _ptr(l0), len(first), _ptr(first),
                                                  _ptr(second), _ptr(modulus), _ptr(phi0)))

    def commit(self, step0=0):
        self._check(lib().dpd_commit(self.h, int(step0)))
        self.n = sum(self.box_n)

    def download_box(self, box):
        n = self.box_n[box]
        x, u, v, f = (np.empty((n, 3)) for _ in range(4))
        self._check(lib().dpd_download_box(self.h, int(box), _ptr(x), _ptr(u), _ptr(v), _ptr(f)))
        return x, u, v, f

    def set_box_interactions(self, box, a=None, gamma=None):
        a, g = _f64(a), _f64(gamma)
        self._check(lib().dpd_set_box_interactions(self.h, int(box), _ptr(a), _ptr(g)))
