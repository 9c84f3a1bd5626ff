"""The hit-list kernel's candidate test on 8-bit scan codes (pair_kernels.hpp, scan_code in
engine_common.hpp), checked on the CPU in FP32 arithmetic: the decoded offset is within
w/512 of the true one on each axis, the tested distance is within the margin m the kernel
widens its bound by, so no pair inside the cut-off can be missed, and the x field of a row
segment (bytes 0-1 of the code less 256 times the segment's first cell index) gives the
candidate's x in the segment for any cell index, across multiples of 256 too."""
import numpy as np


def encode(o, w):
    """q = (int)(o * (256 / w)) in FP32, as scan_code computes it"""
    q = np.floor(o.astype(np.float32) * (np.float32(256.0) / w.astype(np.float32))).astype(np.int64)
    return np.clip(q, 0, 255)


def scan_code(q, cx):
    """the 32-bit code: x offset, cell x index mod 256, y offset, z offset"""
    q = np.asarray(q, np.uint64)
    return (q[..., 0] | ((np.asarray(cx, np.uint64) & 0xFF) << 8) | (q[..., 1] << 16) | (q[..., 2] << 24)).astype(np.uint32)


def test_row_segment_x_field():
    """For a candidate in cell fx + r (r = 0, 1, 2) of a row segment starting at cell fx, the
    kernel's (code - ((fx mod 256) << 8)) mod 2^16 is r * 256 + q_x, for first cells on
    either side of multiples of 256 and rows of up to 2^12 cells; bytes 2 and 3 stay the y and
    z offsets of the original code."""
    rng = np.random.default_rng(6)
    for fx in [0, 1, 253, 254, 255, 256, 257, 510, 511, 4093] + list(rng.integers(0, 4094, 50)):
        for r in range(3):
            q = rng.integers(0, 256, (64, 3))
            c = scan_code(q, fx + r)
            x16 = (c.astype(np.int64) - ((fx & 0xFF) << 8)) & 0xFFFF
            assert np.array_equal(x16, r * 256 + q[:, 0])
            assert np.array_equal((c >> 16) & 0xFF, q[:, 1]) and np.array_equal(c >> 24, q[:, 2])
    # a plausible mistake fails: subtracting the index without the modulo breaks past 255
    q = np.array([[10, 20, 30]])
    c = scan_code(q, 256)
    assert ((c.astype(np.int64) - (255 << 8)) & 0xFFFF)[0] == 256 + 10
    assert ((c.astype(np.int64) - (255 << 8)))[0] != 256 + 10 + (20 << 16) + (30 << 24)


def test_decode_error_within_half_a_unit():
    rng = np.random.default_rng(3)
    for _ in range(20):
        w = rng.uniform(1.0, 2.0, 3).astype(np.float32)
        o = (rng.uniform(0, 1, (20000, 3)) * w).astype(np.float32)
        o[:5] = 0.0
        o[5:10] = w * np.float32(0.99999994)          # the largest offset the engine stores
        q = encode(o, w)
        dec = (q + 0.5) * (w / 256.0)
        # FP32 rounding of o * 256 / w can land on the next integer: the error then exceeds
        # w/512 by ~1e-5 of it, which the kernel's 1% allowance in m covers
        err = np.abs(dec - o) / (w / 512.0)
        assert np.all(err <= 1 + 1e-3)
        assert err.max() > 0.99


def test_tested_distance_within_the_margin():
    """Two beads in neighbouring cells: the kernel's test distance (own position exact plus
    the cell shift, the candidate decoded from its code) differs from the exact distance by
    less than m = 1.01 |w| / 512 + 1e-5, so dr < 1 implies dr_test < 1 + m."""
    rng = np.random.default_rng(4)
    f32 = np.float32
    for _ in range(20):
        w = rng.uniform(1.0, 1.6, 3).astype(f32)
        u = w / f32(256.0)
        m = np.sqrt(np.sum(w.astype(np.float64) ** 2)) * 1.01 / 512.0 + 1e-5
        n = 20000
        oi = (rng.uniform(0, 1, (n, 3)) * w).astype(f32)
        oj = (rng.uniform(0, 1, (n, 3)) * w).astype(f32)
        shift = rng.integers(-1, 2, (n, 3)).astype(f32) * w      # cell of j relative to i, times w
        exact = np.linalg.norm(oi.astype(np.float64) - (oj.astype(np.float64) + shift), axis=1)
        qx = (oi - shift) - f32(0.5) * u                          # the kernel's per-cell position
        dd = qx - encode(oj, w).astype(f32) * u
        test = np.sqrt(np.sum(dd.astype(f32) ** 2, axis=1, dtype=f32))
        assert np.all(np.abs(test - exact) < m)
        inside = exact < 1.0
        assert np.all(test[inside] ** 2 < (1.0 + m) ** 2)


def test_margin_is_needed():
    """Without the widening some pairs just inside the cut-off fail the code test (a check
    that the test above can fail)."""
    rng = np.random.default_rng(5)
    f32 = np.float32
    w = np.ones(3, f32)
    u = w / f32(256.0)
    n = 200000
    oi = rng.uniform(0, 1, (n, 3)).astype(f32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    oj_abs = oi + (d * rng.uniform(0.995, 1.0, (n, 1))).astype(f32)
    c = np.floor(oj_abs)
    oj = (oj_abs - c).astype(f32)
    exact = np.linalg.norm(oi.astype(np.float64) - (oj.astype(np.float64) + c), axis=1)
    qx = (oi - c.astype(f32)) - f32(0.5) * u
    dd = qx - encode(oj, w).astype(f32) * u
    test2 = np.sum(dd ** 2, axis=1)
    assert np.any((exact < 1.0) & (test2 >= 1.0))
