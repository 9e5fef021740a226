"""Pin the 3-D form of the raster-integrator oracle (CPU only).

The reference integrates the plane (nbody_kernel.py:97-560); its scheme in
space adds a z component to every vector.  What fixes the 3-D oracle:

* a planar scene (z = vz = 0) gives the plane oracle's results bit for bit;
* the same planar scene rotated into a tilted plane gives the rotated plane
  results (forces, encounter times, hot classification, contact distances
  and sub-samples all depend on |d| only, so any dropped or misplaced z
  term breaks it);
* a circular inclined orbit around the star follows its closed form;
* the 3-D form of the reference's NumPy `reference_advance`
  (test_kernel.py:16-81) agrees within the reference's 5e-6
  (test_kernel.py:116-131).
"""
import numpy as np

import oracle
from oracle import nbody_np as nb

DAY = 1 / 365.25


def _hot_scene(n=300, seed=5):
    """Asteroids crossing the Earth's position at up to 0.5 AU/yr relative
    speed: many are hot (fine loop, sub-samples), one makes contact."""
    rng = np.random.default_rng(seed)
    x = np.concatenate([[0.0, 1.0], 1.0 + rng.normal(0, 0.01, n)])
    y = np.concatenate([[0.0, 0.0], rng.normal(0, 0.01, n)])
    vx = np.concatenate([[0.0, 0.0], rng.normal(0, 0.5, n)])
    vy = np.concatenate([[0.0, 2 * np.pi], 2 * np.pi + rng.normal(0, 0.5, n)])
    m = np.concatenate([[1.0, 3e-6], np.full(n, 1e-12)])
    vis = np.ones(n + 2, np.uint8)
    ast = np.concatenate([[0, 0], np.ones(n)]).astype(np.uint8)
    rr = np.concatenate([[0.00465, 4.3e-5], np.full(n, 2e-5)])
    return x, y, vx, vy, m, vis, ast, rr


def _rotation(tilt_deg=30.0, turn_deg=40.0):
    th, ph = np.radians(tilt_deg), np.radians(turn_deg)
    rx = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]])
    rz = np.array([[np.cos(ph), -np.sin(ph), 0], [np.sin(ph), np.cos(ph), 0], [0, 0, 1]])
    return rz @ rx


def test_planar_scene_equals_plane_oracle():
    x, y, vx, vy, m, vis, ast, rr = _hot_scene()
    n, z = len(x), np.zeros(len(x))
    o2, s2, sub2, h2 = oracle.nbody_frames(x, y, vx, vy, m, vis, ast, rr, DAY, 4)
    o3, s3, sub3, h3 = oracle.nbody_frames3d(x, y, z, vx, vy, z, m, vis, ast, rr, DAY, 4)
    S2, S3 = s2.reshape(4, n), s3.reshape(6, n)
    assert np.array_equal(S3[[0, 1, 3, 4]], S2)
    assert not np.any(S3[[2, 5]])
    O3 = o3.reshape(4, 6, n)
    assert np.array_equal(O3[:, [0, 1, 3, 4]], o2.reshape(4, 4, n))
    assert sum(len(i) for i, _ in sub2) > 100          # the fine loop ran
    for (i2, p2), (i3, p3) in zip(sub2, sub3):
        assert np.array_equal(i2, i3)
        assert np.array_equal(p3[:, :2], p2) and not np.any(p3[:, 2])
    for u, v in zip(h2, h3):
        assert np.array_equal(u, v)


def test_rotated_plane_equals_plane():
    x, y, vx, vy, m, vis, ast, rr = _hot_scene()
    n, z = len(x), np.zeros(len(x))
    R = _rotation()
    P = R @ np.array([x, y, z])
    V = R @ np.array([vx, vy, z])
    _, s2, sub2, h2 = oracle.nbody_frames(x, y, vx, vy, m, vis, ast, rr, DAY, 4)
    _, s3, sub3, h3 = oracle.nbody_frames3d(*P, *V, m, vis, ast, rr, DAY, 4)
    S2, S3 = s2.reshape(4, n), s3.reshape(6, n)
    pb, vb = R.T @ S3[:3], R.T @ S3[3:]
    assert np.abs(pb[:2] - S2[:2]).max() < 1e-12
    assert np.abs(pb[2]).max() < 1e-12
    assert np.abs(vb[:2] - S2[2:]).max() < 1e-10 * np.abs(S2[2:]).max()
    assert np.abs(vb[2]).max() < 1e-10 * np.abs(S2[2:]).max()
    assert len(h2[0]) >= 1
    assert np.array_equal(h2[0], h3[0]) and np.array_equal(h2[2], h3[2])
    assert np.allclose(h2[1], h3[1], rtol=0, atol=1e-9)
    for (i2, p2), (i3, p3) in zip(sub2, sub3):
        assert len(i2) > 20 and np.array_equal(i2, i3)
        back = np.einsum("ij,mjn->min", R.T, p3.astype(np.float64))
        assert np.abs(back[:, :2] - p2).max() < 5e-7      # f32 rounding
        assert np.abs(back[:, 2]).max() < 5e-7


def test_inclined_circular_orbit():
    """A massless body on a circular orbit of radius 1 AU inclined by 25 deg
    (speed set for the softened force G r / (r^2 + soft)^1.5) stays on the
    closed-form circle over 16 one-day frames; the star does not move."""
    G, soft, r = nb.G_AU, nb.SOFTENING, 1.0
    w = np.sqrt(G / (r * r + soft) ** 1.5)            # angular velocity
    inc, node = np.radians(25.0), np.radians(70.0)
    e1 = np.array([np.cos(node), np.sin(node), 0.0])
    e2 = np.array([-np.sin(node) * np.cos(inc), np.cos(node) * np.cos(inc), np.sin(inc)])
    p0, v0 = r * e1, r * w * e2
    cols = [np.array([0.0, p0[c]]) for c in range(3)] + \
           [np.array([0.0, v0[c]]) for c in range(3)]
    out, state, _, _ = oracle.nbody_frames3d(
        *cols, np.array([1.0, 0.0]), np.ones(2, np.uint8), np.array([0, 1], np.uint8),
        None, DAY, 16)
    S = state.reshape(6, 2)
    t = 16 * DAY
    pe = r * (np.cos(w * t) * e1 + np.sin(w * t) * e2)
    ve = r * w * (-np.sin(w * t) * e1 + np.cos(w * t) * e2)
    assert np.abs(S[:3, 1] - pe).max() < 1e-9
    assert np.abs(S[3:, 1] - ve).max() < 1e-8
    assert not np.any(S[:, 0])
    # every frame on the circle (f32 snapshots)
    o = out.reshape(16, 6, 2)
    for k in range(16):
        tk = (k + 1) * DAY
        pk = r * (np.cos(w * tk) * e1 + np.sin(w * tk) * e2)
        assert np.abs(o[k, :3, 1] - pk).max() < 2e-7


def test_frame_matches_reference_advance_3d():
    """test_kernel.py:116-131 in space: the seed-7 scene of 200 asteroids
    given inclinations up to 10 deg (positions and velocities rotated about
    random in-plane axes), one 5-day frame: |dpos|, |dvel| < 5e-6 against
    the 3-D form of reference_advance; the frozen planet does not move."""
    x, y, vx, vy, m, vis, ast = nb.make_system(200)
    n = len(x)
    rng = np.random.default_rng(11)
    P = np.array([x, y, np.zeros(n)])
    V = np.array([vx, vy, np.zeros(n)])
    for i in range(1, n):                     # the star (index 0) stays put
        phi, inc = rng.uniform(0, 2 * np.pi), np.radians(rng.uniform(0, 10))
        u = np.array([np.cos(phi), np.sin(phi), 0.0])
        K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
        R = np.eye(3) + np.sin(inc) * K + (1 - np.cos(inc)) * K @ K
        P[:, i], V[:, i] = R @ P[:, i], R @ V[:, i]
    out, _, _, _ = oracle.nbody_frames3d(*P, *V, m, vis, ast, None, 5 * DAY, 1)
    p, v = nb.advance_nd(5 * DAY, list(P), list(V), m, vis, ast)
    o = out[0].reshape(6, n)
    assert np.abs(o[:3] - np.array(p)).max() < 5e-6
    assert np.abs(o[3:] - np.array(v)).max() < 5e-6
    assert np.abs(o[:3, 3] - P[:, 3]).max() < 1e-5      # frozen planet


def test_advance_nd_reduces_to_reference_advance():
    """The 2-component form of advance_nd is the reference's function; with
    z = vz = 0 the 3-component form agrees with it to rounding (it differs
    only in sqrt(dx^2 + dy^2 + 0) against hypot in the step control)."""
    x, y, vx, vy, m, vis, ast = nb.make_system(50)
    n = len(x)
    ref = nb.reference_advance(2 * DAY, x, y, vx, vy, m, vis, ast)
    p, v = nb.advance_nd(2 * DAY, [x, y, np.zeros(n)], [vx, vy, np.zeros(n)],
                         m, vis, ast)
    got = [p[0], p[1], v[0], v[1]]
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(b).max())
    assert not np.any(p[2]) and not np.any(v[2])
