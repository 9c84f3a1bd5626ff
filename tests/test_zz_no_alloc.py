"""dpd_slice_stress_noalloc (the allocation-free variant of dpd_slice_stress, tables in a
static device scratch) through the C ABI against dpd_slice_stress on the same state: two
calls in a row, the second with a smaller input, and a size beyond the scratch refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import dpd_amd as D
from helpers import chain_state

pytestmark = pytest.mark.gpu
A3 = np.array([[25.0, 40.0, 30.0], [40.0, 25.0, 35.0], [30.0, 35.0, 25.0]])
G3 = np.full((3, 3), 4.5)
SCRATCH_DOUBLES = 1 << 19


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine(cells, seed):
    n = 3 * int(np.prod(cells))
    nch = n // 60
    ids, t, x, u, v, (h, tl, fi, se) = chain_state(cells, nch, 10, n - 10 * nch, seed=seed)
    e = D.Engine(cells, A3, G3, seed=seed)
    e.upload(ids, t, x, v, unpbc=u)
    e.upload_topology(h, tl, np.full(len(h), 128.0), np.full(len(h), 0.5), fi, se, np.full(len(fi), 20.0),
                      np.zeros(len(fi)))
    e.step(4, observe=True)
    return e


def noalloc(e, nslices, box=0):
    nt = e._ntypes(box)
    pair = np.zeros((nslices, nt * nt, 9))
    bond = np.zeros((nslices, nt * nt, 9))
    rc = D.lib().dpd_slice_stress_noalloc(e.h, box, nslices, pair.ctypes.data_as(C.c_void_p),
                                          bond.ctypes.data_as(C.c_void_p))
    return rc, pair, bond


def close(a, b):
    # FP64 atomics add in any order: equal to rounding
    return np.allclose(a, b, rtol=1e-12, atol=1e-10 * max(1.0, float(np.abs(b).max())))


def test_noalloc_matches_the_allocating_call():
    big = engine((14, 12, 16), seed=3)
    small = engine((8, 9, 7), seed=4)
    for e, nslices in ((big, 40), (small, 9)):
        rc, pg, bg = noalloc(e, nslices)
        assert rc == D.OK, D.lib().dpd_last_error(e.h).decode()
        pr, br = e.slice_stress(nslices)
        say(f"  {e.n} beads, {nslices} slices: max |pair diff| {np.abs(pg - pr).max():.2e}, "
            f"max |bond diff| {np.abs(bg - br).max():.2e}")
        assert close(pg, pr) and close(bg, br)
        assert np.abs(pr).max() > 0 and np.abs(br).max() > 0
    # the second, smaller call left nothing of the first in its tables (checked above), and
    # the first engine still gets its own tables after the smaller call
    rc, pg, bg = noalloc(big, 40)
    pr, br = big.slice_stress(40)
    assert rc == D.OK and close(pg, pr) and close(bg, br)


def test_noalloc_refuses_tables_beyond_the_scratch():
    e = engine((8, 8, 8), seed=5)
    too_many = SCRATCH_DOUBLES // (2 * 9 * 9) + 1   # 3 types: 2 tables of 9 * 9 doubles per slice
    nt = 3
    pair = np.zeros((too_many, nt * nt, 9))
    bond = np.zeros((too_many, nt * nt, 9))
    rc = D.lib().dpd_slice_stress_noalloc(e.h, 0, too_many, pair.ctypes.data_as(C.c_void_p),
                                          bond.ctypes.data_as(C.c_void_p))
    assert rc == 1 and "scratch" in D.lib().dpd_last_error(e.h).decode()
    # the largest size that fits still works
    rc, pg, bg = noalloc(e, too_many - 1)
    assert rc == D.OK
    pr, br = e.slice_stress(too_many - 1)
    assert close(pg, pr) and close(bg, br)
