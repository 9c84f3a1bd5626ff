"""Binding of the host layer (include/dpd_host.h): dmpci control files, polymer topology,
starting states, restart snapshots and the whole-run driver."""
import ctypes as C

import numpy as np

from . import Batch, Engine, PAIR_AUTO, _ptr, lib

STATE_NAMES = {0: "random", 1: "restart", 2: "lamella", 3: "other"}


class Counts(C.Structure):
    _fields_ = [("ntypes", C.c_int32), ("nbeads", C.c_int64), ("nbonds", C.c_int64),
                ("npairs", C.c_int64), ("npolymers", C.c_int64), ("npolymer_types", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("wx", C.c_double), ("wy", C.c_double), ("wz", C.c_double),
                ("density", C.c_double), ("kT", C.c_double), ("lam", C.c_double), ("dt", C.c_double),
                ("seed", C.c_int64), ("total_time", C.c_int64), ("sample_period", C.c_int64),
                ("restart_period", C.c_int64), ("ncommands", C.c_int32), ("state", C.c_int32),
                ("restart_state", C.c_int64)]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("nbeads", C.c_int64), ("seconds", C.c_double),
                ("bead_steps_per_s", C.c_double), ("last_kT", C.c_double),
                ("last_pressure", C.c_double), ("samples", C.c_int64),
                ("commands_applied", C.c_int64)]


_P = C.c_void_p
HOST_SIGNATURES = {
    "dpd_input_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int32]),
    "dpd_input_parse": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int32]),
    "dpd_input_free": (None, [_P]),
    "dpd_input_counts": (C.c_int, [_P, C.POINTER(Counts)]),
    "dpd_input_types": (C.c_int, [_P, C.c_char_p, C.c_int32, _P, _P]),
    "dpd_input_beads": (C.c_int, [_P, _P, _P, _P, _P]),
    "dpd_input_topology": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dpd_input_bond_types": (C.c_int, [_P, C.POINTER(C.c_int32), _P, _P]),
    "dpd_input_params": (C.c_int, [_P, C.c_int32, _P]),
    "dpd_input_write_summary": (C.c_int, [_P, C.c_char_p]),
    "dpd_input_build_state": (C.c_int, [_P, C.c_uint64, _P, _P, _P]),
    "dpd_input_command": (C.c_int, [_P, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int64),
                                    C.c_char_p, C.c_int32]),
    "dpd_restart_read": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dpd_restart_write": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dpd_run": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(RunStats), C.c_char_p, C.c_int32]),
}

_bound = False


def hlib():
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in HOST_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class InputError(RuntimeError):
    pass


class Input:
    """A parsed dmpci control file with its polymers expanded."""

    def __init__(self, path=None, text=None):
        L = hlib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        if text is not None:
            rc = L.dpd_input_parse(text.encode(), C.byref(h), err, 512)
        else:
            rc = L.dpd_input_read(str(path).encode(), C.byref(h), err, 512)
        if rc != 0:
            raise InputError(err.value.decode())
        self.h = h
        c = Counts()
        L.dpd_input_counts(self.h, C.byref(c))
        self.counts = {k: getattr(c, k) for k, _ in Counts._fields_}
        self.state = STATE_NAMES[c.state]
        names = C.create_string_buffer(4096)
        nt = c.ntypes
        self.a = np.zeros((nt, nt))
        self.gamma = np.zeros((nt, nt))
        L.dpd_input_types(self.h, names, 4096, _ptr(self.a), _ptr(self.gamma))
        self.bead_names = names.value.decode().split(",")

    def __del__(self):
        if getattr(self, "h", None):
            hlib().dpd_input_free(self.h)
            self.h = None

    @property
    def cells(self):
        c = self.counts
        return (c["nx"], c["ny"], c["nz"])

    @property
    def widths(self):
        c = self.counts
        return (c["wx"], c["wy"], c["wz"])

    def beads(self):
        n = self.counts["nbeads"]
        ids, t, pid, pt = (np.zeros(n, np.int32) for _ in range(4))
        hlib().dpd_input_beads(self.h, _ptr(ids), _ptr(t), _ptr(pid), _ptr(pt))
        return ids, t, pid, pt

    def topology(self):
        nb, npair = self.counts["nbonds"], self.counts["npairs"]
        h, t = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        k, l0 = np.zeros(nb), np.zeros(nb)
        fi, se = np.zeros(npair, np.int32), np.zeros(npair, np.int32)
        mo, ph = np.zeros(npair), np.zeros(npair)
        hlib().dpd_input_topology(self.h, _ptr(h), _ptr(t), _ptr(k), _ptr(l0), _ptr(fi), _ptr(se),
                                  _ptr(mo), _ptr(ph))
        return (h, t, k, l0), (fi, se, mo, ph)

    def bond_types(self):
        """(head type, tail type) of each Bond declaration, in declaration order."""
        n = C.c_int32()
        hlib().dpd_input_bond_types(self.h, C.byref(n), None, None)
        h, t = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        hlib().dpd_input_bond_types(self.h, C.byref(n), _ptr(h), _ptr(t))
        return list(zip(h.tolist(), t.tolist()))

    def build_state(self, seed=None):
        n = self.counts["nbeads"]
        x, u, v = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        s = abs(self.counts["seed"]) if seed is None else seed
        rc = hlib().dpd_input_build_state(self.h, int(s), _ptr(x), _ptr(u), _ptr(v))
        if rc != 0:
            raise InputError(f"cannot build state ({rc})")
        return x, u, v

    def commands(self):
        out = []
        for i in range(self.counts["ncommands"]):
            name, args = C.create_string_buffer(256), C.create_string_buffer(1024)
            t = C.c_int64()
            hlib().dpd_input_command(self.h, i, name, 256, C.byref(t), args, 1024)
            out.append((name.value.decode(), t.value, args.value.decode().split()))
        return out

    def write_summary(self, path):
        """Write the dmpcis initial-state summary (CInitialState::SerializeAsciiFile)."""
        rc = hlib().dpd_input_write_summary(self.h, str(path).encode())
        if rc != 0:
            raise InputError(f"cannot write {path}")

    def engine(self, device=0, pair_kernel=PAIR_AUTO):
        c = self.counts
        return Engine(self.cells, self.a, self.gamma, dt=c["dt"], lam=c["lam"], kT=c["kT"],
                      seed=abs(c["seed"]), widths=self.widths, device=device, pair_kernel=pair_kernel)


def replica_batch(inputs, seeds, device=0, pair_kernel=PAIR_AUTO):
    """One batched engine with a box per (input, seed): box i runs inputs[i]'s control
    file from the starting state built with seeds[i], and draws its pair random stream
    from the same seed.  The same Input may appear several times.  Returns the engine
    (committed at step 0) and per-box (bead types, polymer types, box lengths)."""
    boxes = []
    for inp, s in zip(inputs, seeds):
        c = inp.counts
        boxes.append({"cells": inp.cells, "a": inp.a, "gamma": inp.gamma, "dt": c["dt"], "lam": c["lam"],
                      "kT": c["kT"], "seed": int(s), "widths": inp.widths})
    e = Batch(boxes, device=device, pair_kernel=pair_kernel)
    meta = []
    for b, (inp, s) in enumerate(zip(inputs, seeds)):
        ids, types, pid, pt = inp.beads()
        x, u, v = inp.build_state(seed=int(s))
        e.upload_box(b, ids, types, x, v, unpbc=u)
        (h, t, k, l0), (fi, se, mo, ph) = inp.topology()
        if len(h) or len(fi):
            e.upload_box_topology(b, h, t, k, l0, fi, se, mo, ph)
        meta.append((types, pt, np.array(inp.cells, float) * np.array(inp.widths)))
    e.commit(0)
    return e, meta


def read_restart(path, n):
    pid, pt, ids, t = (np.zeros(n, np.int32) for _ in range(4))
    x, u, v, f = (np.zeros((n, 3)) for _ in range(4))
    rc = hlib().dpd_restart_read(str(path).encode(), n, _ptr(pid), _ptr(pt), _ptr(ids), _ptr(t),
                                 _ptr(x), _ptr(u), _ptr(v), _ptr(f))
    if rc != 0:
        raise InputError(f"cannot read restart {path}")
    return {"poly": pid, "ptype": pt, "id": ids, "type": t, "x": x, "u": u, "v": v, "f": f}


def write_restart(path, poly, ptype, ids, types, L, x, u, v, f, radius=None):
    n = len(ids)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    poly, ptype, ids, types = i32(poly), i32(ptype), i32(ids), i32(types)
    L, x, u, v, f = f64(L), f64(x), f64(u), f64(v), f64(f)
    rad = f64(radius) if radius is not None else None
    rc = hlib().dpd_restart_write(str(path).encode(), n, _ptr(poly), _ptr(ptype), _ptr(ids), _ptr(types),
                                  _ptr(rad), _ptr(L), _ptr(x), _ptr(u), _ptr(v), _ptr(f))
    if rc != 0:
        raise InputError(f"cannot write restart {path} ({rc})")


def run(directory, run_id, device=0):
    st = RunStats()
    err = C.create_string_buffer(512)
    rc = hlib().dpd_run(str(directory).encode(), str(run_id).encode(), device, C.byref(st), err, 512)
    if rc != 0:
        raise InputError(err.value.decode())
    return {k: getattr(st, k) for k, _ in RunStats._fields_}
