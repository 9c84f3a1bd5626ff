"""Failure reporting through the C ABI, with every condition forced from the host and every
kernel exiting normally: beads outside the box are refused at upload, non-finite
velocities are caught by the step's device checks (DPD_ERR_NAN; a position that is not
finite is parked in a valid cell before it is binned, so no index leaves its array),
moves of more than one cell per step are counted, and steps out of order are refused."""
import numpy as np
import pytest

import dpd_amd as D
from helpers import random_state

pytestmark = pytest.mark.gpu
A2 = np.array([[25.0, 40.0], [40.0, 25.0]])
G2 = np.full((2, 2), 4.5)
ERR_NAN, ERR_OUT_OF_BOX, ERR_STATE, ERR_PARAM = 4, 5, 6, 1


def test_bead_outside_the_box_is_refused():
    cells = (8, 8, 8)
    ids, t, x, v = random_state(cells, 1536, 2, seed=1)
    e = D.Engine(cells, A2, G2)
    for bad in (8.0, -1e-9, np.nan):
        xb = x.copy()
        xb[17, 1] = bad
        with pytest.raises(D.DpdError) as err:
            e.upload(ids, t, xb, v)
        assert err.value.code == ERR_OUT_OF_BOX and "bead 17" in str(err.value)
    with pytest.raises(D.DpdError) as err:
        e.step(1)                       # nothing uploaded yet
    assert err.value.code == ERR_STATE
    e.upload(ids, t, x, v)
    e.step(2)
    with pytest.raises(D.DpdError) as err:
        e.step(-1)
    assert err.value.code == ERR_PARAM


@pytest.mark.parametrize("bad", [np.nan, np.inf, 1e12])
def test_non_finite_motion_is_reported(bad):
    cells = (8, 8, 8)
    ids, t, x, v = random_state(cells, 1536, 2, seed=2)
    v = v.copy()
    v[100] = bad
    e = D.Engine(cells, A2, G2)
    e.upload(ids, t, x, v)
    with pytest.raises(D.DpdError) as err:
        e.step(3)
    assert err.value.code == ERR_NAN


def test_multi_cell_moves_are_counted_and_binned():
    """A bead moving 1.6 cells in one step: counted in `jumps`, and re-binned correctly
    (the reference moves a bead at most one cell and logs the rest as errors)."""
    cells = (10, 10, 10)
    ids, t, x, v = random_state(cells, 3000, 2, seed=3)
    v = v.copy()
    v[7] = (80.0, 0.0, 0.0)
    x = x.copy()
    x[7] = (2.5, 5.0, 5.0)
    e = D.Engine(cells, A2, G2, dt=0.02, pair_kernel=D.PAIR_FULLSHELL)
    e.upload(ids, t, x, v)
    e.step(1, observe=True)
    ob = e.observables()
    assert ob["jumps"] >= 1
    X, U, V, F = e.download()
    assert 3.9 < X[7, 0] < 4.3                          # 2.5 + 0.02 * 80 (+ the force term)
    assert np.all((X >= 0) & (X < 10))
    e.step(3)                                           # the state stays usable

