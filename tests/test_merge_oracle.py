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


def _merge_scene():
    n = 6
    f = np.float32
    sample = np.concatenate([np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0], f),      # x
                             np.array([0.0, 0.5, 1.0, 1.5, 2.0, 2.5], f),      # y
                             np.array([1.0, -1.0, 0.5, 0.0, 2.0, -2.0], f),    # vx
                             np.array([0.0, 0.25, -0.5, 1.0, 0.0, 0.0], f)])   # vy
    mass = np.array([3.0, 1.0, 0.5, 2.0, 1.0, 1.0])
    vis = np.ones(n, np.uint8)
    rr = np.array([0.3, 0.2, 0.1, 0.25, 0.1, 0.1])
    ast = np.zeros(n, np.uint8)
    star = np.zeros(n, np.uint8)
    return sample, mass, vis, rr, ast, star


def test_apply_merges_conserves_and_keeps_list_order():
    """One merge conserves mass, momentum, the mass-weighted centre and r^3;
    the chain (0, 1), (1, 2) merges once (1 is gone after the first pair);
    the record unpacks to the merge's time, bodies, mass and position."""
    import struct
    sample, mass, vis, rr, ast, star = _merge_scene()
    n = len(mass)
    m0, r0 = mass.copy(), rr.copy()
    body, dead, radius, ev, merges, sh = M.apply_merges(
        sample, [(0, 1, 0.25), (1, 2, -0.5)], [], 7, mass, vis, rr, ast, star, {},
        False, 1.0, 100.0, 0.5)
    assert sh is None and merges == 1 and dead == [1] and list(body) == [0]
    x, y, vx, vy, m = body[0]
    s = sample.astype(np.float64)
    assert m == m0[0] + m0[1] and mass[1] == 0.0 and vis[1] == 0 and mass[0] == m
    for k, v in enumerate((x, y, vx, vy)):
        w = (s[k * n] * m0[0] + s[k * n + 1] * m0[1]) / m
        assert abs(v - w) <= 1e-15 * max(1.0, abs(w))
    assert abs(radius[0] ** 3 - (r0[0] ** 3 + r0[1] ** 3)) < 1e-15
    t, a, b, mf, kind, ex, ey = struct.unpack("<dIIfIff", ev[0])
    assert (a, b, kind) == (0, 1, 0) and mf == np.float32(m)
    assert t == 100.0 + (7 + 1 + 0.25) * 0.5
    assert (ex, ey) == (float(np.float32(x)), float(np.float32(y)))


def test_apply_merges_contact_time_shatter_and_runaway():
    """A kernel contact time inside one raster replaces frac (and is
    consumed), one outside does not; a fast massive pair shatters and stops
    the call before later pairs and the runaway list; runaway removals skip
    invisible bodies and write kind-2 records with a = 0xFFFFFFFF."""
    import struct
    sample, mass, vis, rr, ast, star = _merge_scene()
    contact = {1: 8.6, 4: 20.0}
    _, _, _, ev, merges, sh = M.apply_merges(
        sample, [(0, 1, 0.25), (3, 4, 0.1)], [5, 1], 7, mass, vis, rr, ast, star,
        contact, False, 1.0, 0.0, 1.0)
    assert merges == 2 and contact == {}
    assert struct.unpack("<dIIfIff", ev[0])[0] == 7 + 1 + (8.6 - 8)
    assert struct.unpack("<dIIfIff", ev[1])[0] == 7 + 1 + 0.1
    assert len(ev) == 3                     # runaway 1 was merged away already
    t, a, b, mf, kind, ex, ey = struct.unpack("<dIIfIff", ev[2])
    assert (a, b, kind, mf) == (0xFFFFFFFF, 5, 2, 0.0) and t == 8.0
    # shatter: |dv| = |(-2, 0.25)| = 2.02 >= 1.5 v_esc = 1.5 sqrt(2 G 4 / 0.5) = 1.90
    # at G = 0.1
    sample, mass, vis, rr, ast, star = _merge_scene()
    _, dead, _, ev, merges, sh = M.apply_merges(
        sample, [(0, 1, 0.0), (3, 4, 0.0)], [5], 7, mass, vis, rr, ast, star, {},
        True, 0.1, 0.0, 1.0)
    assert sh == (0, 1, 8.0) and merges == 0 and dead == [] and ev == []
    # the same pair with the star flag set: no shatter, a merge
    sample, mass, vis, rr, ast, star = _merge_scene()
    star[0] = 1
    _, _, _, ev, merges, sh = M.apply_merges(
        sample, [(0, 1, 0.0)], [], 7, mass, vis, rr, ast, star, {}, True, 0.1, 0.0, 1.0)
    assert sh is None and merges == 1
