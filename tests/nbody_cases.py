"""Scenes and reference integrations of the reference's nbody tests, shared
by the oracle tests (CPU) and the GPU tests.

* divers            — test_kernel.py:190-219 (20 bodies falling into the sun)
* sub_sample_scene  — test_subsamples.py:47-63 + its fine Verlet reference
  (referenz_bahn, test_subsamples.py:20-43)
* moving_star       — test_bewegter_stern.py:27-74 (close pass at a star in
  uniform motion) + its fine Verlet reference
"""
import numpy as np

G_AU = 4 * np.pi ** 2
SOFTENING = 1e-6


def divers():
    rng = np.random.default_rng(3)
    nd = 20
    x = np.zeros(1 + nd)
    y = np.zeros(1 + nd)
    vx = np.zeros(1 + nd)
    vy = np.zeros(1 + nd)
    mass = np.full(1 + nd, 1e-12)
    mass[0] = 1.0
    vis = np.ones(1 + nd, np.uint8)
    is_ast = np.ones(1 + nd, np.uint8)
    is_ast[0] = 0
    th = rng.random(nd) * 2 * np.pi
    r0 = 0.25 + 0.1 * rng.random(nd)
    x[1:] = r0 * np.cos(th)
    y[1:] = r0 * np.sin(th)
    return x, y, vx, vy, mass, vis, is_ast


def sub_sample_scene():
    diver = (0.15, 0.004, -160.0, 0.0)
    far = (3.0, 0.0, 0.0, np.sqrt(G_AU / 3.0))
    x = np.array([0.0, diver[0], far[0]])
    y = np.array([0.0, diver[1], far[1]])
    vx = np.array([0.0, diver[2], far[2]])
    vy = np.array([0.0, diver[3], far[3]])
    return (x, y, vx, vy, np.array([1.0, 0.0, 0.0]), np.ones(3, np.uint8),
            np.array([0, 1, 1], np.uint8)), diver


def fine_orbit(px, py, pvx, pvy, mass, dt, steps):
    """Velocity Verlet of a test particle around a fixed mass; the
    position after every step."""
    mu = G_AU * mass
    h = dt / steps

    def acc(a, b):
        r2 = a * a + b * b + SOFTENING
        f = -mu / (r2 * np.sqrt(r2))
        return f * a, f * b
    ax, ay = acc(px, py)
    track = np.empty((steps, 2))
    for s in range(steps):
        px += pvx * h + 0.5 * ax * h * h
        py += pvy * h + 0.5 * ay * h * h
        nax, nay = acc(px, py)
        pvx += 0.5 * (ax + nax) * h
        pvy += 0.5 * (ay + nay) * h
        ax, ay = nax, nay
        track[s] = (px, py)
    return track


def moving_star(v_star):
    x = np.array([0.0, -0.6])
    y = np.array([0.0, -0.25])
    vx = np.array([v_star, 11.0])
    vy = np.array([0.0, 3.0])
    return (x, y, vx, vy, np.array([1.0, 0.0]), np.ones(2, np.uint8),
            np.array([0, 1], np.uint8))


def moving_star_reference(v_star, dt, steps):
    """Test particle in the field of a star moving at (v_star, 0)."""
    px, py, pvx, pvy = -0.6, -0.25, 11.0, 3.0
    mu = G_AU
    h = dt / steps

    def acc(t):
        dx = v_star * t - px
        dy = -py
        r2 = dx * dx + dy * dy + SOFTENING
        f = mu / (r2 * np.sqrt(r2))
        return f * dx, f * dy
    t = 0.0
    ax, ay = acc(t)
    for _ in range(steps):
        px += pvx * h + 0.5 * ax * h * h
        py += pvy * h + 0.5 * ay * h * h
        t += h
        nax, nay = acc(t)
        pvx += 0.5 * (ax + nax) * h
        pvy += 0.5 * (ay + nay) * h
        ax, ay = nax, nay
    return px, py
