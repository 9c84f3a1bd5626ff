"""Binding of the surrogate network (include/dpd_mlp.h): BF16 matrix-core MLP."""
import ctypes as C

import numpy as np

from . import _ptr, lib

_P = C.c_void_p
MLP_SIGNATURES = {
    "dpd_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_int32,
                                 C.POINTER(C.c_void_p)]),
    "dpd_mlp_destroy": (None, [_P]),
    "dpd_mlp_last_error": (C.c_char_p, [_P]),
    "dpd_mlp_set_scaling": (C.c_int, [_P, _P, _P, _P, _P]),
    "dpd_mlp_param_count": (C.c_int64, [_P]),
    "dpd_mlp_get_params": (C.c_int, [_P, _P]),
    "dpd_mlp_set_params": (C.c_int, [_P, _P]),
    "dpd_mlp_gradients": (C.c_int, [_P, C.c_int64, _P, _P, C.POINTER(C.c_double), _P]),
    "dpd_mlp_train": (C.c_int, [_P, C.c_int64, _P, _P, C.c_int32, C.c_double, _P]),
    "dpd_mlp_predict": (C.c_int, [_P, C.c_int64, _P, _P]),
}
_bound = False


def mlib():
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in MLP_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class MLPError(RuntimeError):
    pass


class Surrogate:
    """nin -> hidden -> hidden -> nout tanh network on the GPU."""

    def __init__(self, nin=6, hidden=256, nout=2, seed=1, device=0):
        h = C.c_void_p()
        rc = mlib().dpd_mlp_create(nin, hidden, nout, seed, device, C.byref(h))
        self.h = h
        if rc != 0:
            msg = mlib().dpd_mlp_last_error(h).decode() if h else "bad sizes"
            self.close()
            raise MLPError(msg)
        self.nin, self.hidden, self.nout = nin, hidden, nout

    def close(self):
        if getattr(self, "h", None):
            mlib().dpd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MLPError(mlib().dpd_mlp_last_error(self.h).decode())

    def set_scaling(self, in_mean, in_scale, out_mean, out_scale):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (in_mean, in_scale, out_mean, out_scale)]
        self._check(mlib().dpd_mlp_set_scaling(self.h, *[_ptr(v) for v in a]))

    def fit_scaling(self, X, Y):
        X, Y = np.asarray(X, float), np.asarray(Y, float)
        self.set_scaling(X.mean(0), X.std(0) + 1e-12, Y.mean(0), Y.std(0) + 1e-12)

    @property
    def nparams(self):
        return int(mlib().dpd_mlp_param_count(self.h))

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self._check(mlib().dpd_mlp_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(mlib().dpd_mlp_set_params(self.h, _ptr(p)))

    def gradients(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        self._check(mlib().dpd_mlp_gradients(self.h, len(X), _ptr(X), _ptr(Y), C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, X, Y, iters=1000, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        losses = np.zeros(max(iters, 1))
        self._check(mlib().dpd_mlp_train(self.h, len(X), _ptr(X), _ptr(Y), int(iters), float(lr),
                                         _ptr(losses)))
        return losses[:iters]

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros((len(X), self.nout))
        self._check(mlib().dpd_mlp_predict(self.h, len(X), _ptr(X), _ptr(Y)))
        return Y
