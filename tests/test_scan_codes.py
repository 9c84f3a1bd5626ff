"""The hit-list kernel's candidate test on 8-bit scan codes (code_dist2 and the row scan in
pair_kernels.hpp, scan_code in engine_common.hpp), restated on the CPU with the kernel's
FP32 and 16-bit integer arithmetic: each axis of the tested difference is within one code
unit of the true one, so with the kernel's widened bound no pair inside the cut-off can be
missed (cubic cells and, as a wider superset, cells of unequal sides); the packed x field of
a row segment and the list entry's cell field are right for any first cell index, across
multiples of 256 too."""
import numpy as np

f32 = np.float32


def encode(o, w):
    """q = (int)(o * (256 / w)) in FP32, clamped to 0..255, as scan_code computes it"""
    q = np.floor(o.astype(f32) * (f32(256.0) / w.astype(f32))).astype(np.int64)
    return np.clip(q, 0, 255)


def scan_code(q, cx):
    """the 32-bit code: x offset, cell x index mod 256, y offset, z offset"""
    q = np.asarray(q, np.uint64)
    return (q[..., 0] | ((np.asarray(cx, np.uint64) & 0xFF) << 8) | (q[..., 1] << 16) | (q[..., 2] << 24)).astype(np.uint32)


def bound(w):
    """the kernel's threshold on the sum of squared unit differences"""
    w = w.astype(f32)
    inv = f32(256.0) / w
    m = np.sqrt(np.sum(w * w, dtype=f32), dtype=f32) * f32(1.01 / 256.0) + f32(1e-5)
    tu = (f32(1.0) + m) * inv.max()
    return int(np.uint32(tu * tu)) + 1


def bead_units(oi, shift, w):
    """the bead's position seen from the candidate's cell, in code units, floored (FP32)"""
    return np.floor((oi.astype(f32) - shift.astype(f32)) * (f32(256.0) / w.astype(f32))).astype(np.int64)


def test_row_segment_x_field():
    """For a candidate in cell fx + r (r = 0, 1, 2) of a row segment starting at cell fx, the
    16-bit difference of the packed x halves, (Qx + fxbits) - (code bytes 0-1) mod 2^16 read
    as signed, is Qx - (r * 256 + q_x) for first cells on either side of multiples of 256 and
    rows of up to 2^12 cells; ((code >> 8) & 63) - fx mod 64 is r (the list entry's cell);
    bytes 2 and 3 stay the y and z offsets of the original code."""
    rng = np.random.default_rng(6)
    for fx in [0, 1, 253, 254, 255, 256, 257, 510, 511, 4093] + list(rng.integers(0, 4094, 50)):
        fxbits = (fx & 0xFF) << 8
        for r in range(3):
            q = rng.integers(0, 256, (64, 3))
            c = scan_code(q, fx + r).astype(np.int64)
            Qx = rng.integers(-256, 512, 64)
            dx = ((Qx + fxbits) - (c & 0xFFFF)) & 0xFFFF
            dx = np.where(dx >= 0x8000, dx - 0x10000, dx)
            assert np.array_equal(dx, Qx - (r * 256 + q[:, 0]))
            assert np.all((((c >> 8) & 63) - (fx & 0xFF)) % 64 == r)
            assert np.array_equal((c >> 16) & 0xFF, q[:, 1]) and np.array_equal(c >> 24, q[:, 2])
    # a plausible mistake fails: the cell index without the modulo breaks past 255
    q = np.array([[10, 20, 30]])
    c = int(scan_code(q, 256)[0])
    assert ((c - (255 << 8)) & 0xFFFF) == 256 + 10
    assert (((c >> 8) & 0xFF) - 255) != 1


def test_axis_error_within_one_unit():
    """Bead and candidate both floored to units of w/256: the unit difference of one axis is
    within one unit of the true difference (plus FP32 rounding, ~1e-4 of a unit)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        w = rng.uniform(1.0, 2.0, 3).astype(f32)
        n = 20000
        oi = (rng.uniform(0, 1, (n, 3)) * w).astype(f32)
        oj = (rng.uniform(0, 1, (n, 3)) * w).astype(f32)
        oj[:5] = 0.0
        oj[5:10] = w * f32(0.99999994)          # the largest offset the engine stores
        shift = (rng.integers(-1, 2, (n, 3)) * w).astype(f32)
        d = bead_units(oi, shift, w) - encode(oj, w)
        true = (oi.astype(np.float64) - shift - oj) / (w.astype(np.float64) / 256.0)
        err = np.abs(d - true)
        assert np.all(err < 1 + 1e-3)
        assert err.max() > 0.99


def test_no_pair_inside_the_cutoff_fails():
    """Pairs of beads in neighbouring cells, cubic and unequal cells, many just inside the
    cut-off: every pair with exact |dr| < 1 passes sum d^2 < bound."""
    rng = np.random.default_rng(4)
    for trial in range(24):
        w = (np.full(3, rng.uniform(1.0, 1.6)) if trial % 2 == 0 else rng.uniform(1.0, 1.6, 3)).astype(f32)
        n = 40000
        oi = (rng.uniform(0, 1, (n, 3)) * w).astype(f32)
        # partners at 0.98 .. 1.0 from the bead, in one of the neighbour cells
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        xj = oi + d * rng.uniform(0.98, 1.0, (n, 1))
        cell = np.floor(xj / w)
        oj = np.minimum((xj - cell * w).astype(f32), w * f32(0.99999994))
        exact = np.linalg.norm(oi.astype(np.float64) - (oj.astype(np.float64) + cell * w), axis=1)
        du = bead_units(oi, (cell * w).astype(f32), w) - encode(oj, w)
        s = np.sum(du * du, axis=1)
        inside = exact < 1.0
        assert inside.sum() > 1000
        assert np.all(s[inside] < bound(w))


def test_margin_is_needed():
    """Without the widening some pairs just inside the cut-off fail the unit test (a check
    that the test above can fail)."""
    rng = np.random.default_rng(5)
    w = np.ones(3, f32)
    n = 200000
    oi = rng.uniform(0, 1, (n, 3)).astype(f32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xj = oi + (d * rng.uniform(0.995, 1.0, (n, 1))).astype(f32)
    cell = np.floor(xj)
    oj = np.minimum((xj - cell).astype(f32), f32(0.99999994))
    exact = np.linalg.norm(oi.astype(np.float64) - (oj.astype(np.float64) + cell), axis=1)
    du = bead_units(oi, cell.astype(f32), w) - encode(oj, w)
    s = np.sum(du * du, axis=1)
    assert np.any((exact < 1.0) & (s >= 256 * 256))


def test_superset_is_tight_for_cubic_cells():
    """The listed superset stays close to the sphere: for unit cells a pair further than
    1 + 2 m apart never passes."""
    rng = np.random.default_rng(7)
    w = np.ones(3, f32)
    m = np.sqrt(3.0) * 1.01 / 256 + 1e-5
    n = 100000
    oi = rng.uniform(0, 1, (n, 3)).astype(f32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xj = oi + d * rng.uniform(1.0 + 2 * m, 1.1, (n, 1))
    cell = np.floor(xj)
    keep = np.all(np.abs(cell) <= 1, axis=1)
    oi, xj, cell = oi[keep], xj[keep], cell[keep]
    oj = np.minimum((xj - cell).astype(f32), f32(0.99999994))
    du = bead_units(oi, cell.astype(f32), w) - encode(oj, w)
    s = np.sum(du * du, axis=1)
    assert np.all(s >= bound(w))
