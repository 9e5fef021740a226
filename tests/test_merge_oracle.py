"""Pin the merge-scan / runaway / bounce oracle (CPU) with closed forms."""
import numpy as np

from oracle import merge_np as M


def test_head_on_bounce_conserves_momentum_and_reverses_by_e():
    n = 2
    # two bodies approaching along x, touching (overlap 0.2 r)
    x = np.array([0.0, 1.8e-3], np.float32)
    y = np.zeros(2, np.float32)
    vx = np.array([1.0, -1.0], np.float32)
    vy = np.zeros(2, np.float32)
    sample = np.concatenate([x, y, vx, vy])
    r = np.array([1e-3, 1e-3])
    m = np.array([2.0, 1.0])
    hi, hj, aff, d = M.bounce_deltas(sample, [0], [1], r, m, 1e-3, 0.6)
    assert aff.tolist() == [0, 1]
    dv = d[:, 2]
    assert abs(m @ dv) < 1e-12                      # momentum conserved
    v_after = vx.astype(np.float64) + dv
    assert abs((v_after[1] - v_after[0]) - 0.6 * 2.0) < 1e-12
    # overlap 0.2e-3 pushed apart 1.1x, mass-weighted
    overlap = 2e-3 - float(x[1])            # f32 sample position
    assert abs((d[1, 0] - d[0, 0]) - 1.1 * overlap) < 1e-15
    assert abs(m @ d[:, 0]) < 1e-15


def test_greedy_filter_keeps_first_contact():
    n = 3
    x = np.array([0.0, 1.5e-3, 3.0e-3], np.float32)
    sample = np.concatenate([x, np.zeros(3, np.float32),
                             np.array([1.0, 0.0, -1.0], np.float32),
                             np.zeros(3, np.float32)])
    r = np.full(3, 1e-3)
    res = M.bounce_deltas(sample, [1, 0], [2, 1], r, np.ones(3), 1e-3)
    hi, hj, aff, _ = res
    assert (hi.tolist(), hj.tolist()) == ([1], [2])   # (0,1) uses body 1 again


def test_star_crossing_reports_fraction():
    """A body moving through a star between the previous and this sample
    hits on the first segment with frac = tt - 1 in (-1, 0]."""
    gx = np.array([0.0, 0.5], np.float32)
    gy = np.array([0.0, 0.0], np.float32)
    px = np.array([0.0, -0.5], np.float32)
    py = np.array([0.0, 0.0], np.float32)
    gv = np.zeros(2, np.float32)
    rr = np.array([0.01, 0.0], np.float32)
    rows, cols, frac = M.merge_scan(gx, gy, gv, gv, px, py, rr, np.ones(2), [0], 0.1)
    assert rows.tolist() == [0] and cols.tolist() == [1]
    assert abs(frac[0] - (0.5 - 1.0)) < 1e-6


def test_runaway_threshold():
    gx = np.array([1.0, 1.0], np.float32)
    gy = np.zeros(2, np.float32)
    v_esc = np.sqrt(2.0)             # G M = 1 at r = 1
    gvx = np.array([3.01 * v_esc, 2.99 * v_esc], np.float32)
    idx = M.runaway(gx, gy, gvx, np.zeros(2, np.float32), np.ones(2), np.ones(2),
                    0.0, 0.0, 2.0)
    assert idx.tolist() == [0]
