"""NumPy restatements of the reference's self-gravity test helpers.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* `wolke`       — test_selfgrav.py:43-51 (seeded uniform cloud)
* `energie`     — test_selfgrav.py:54-63 (Plummer energy, diagonal removed)
* `referenz`    — test_selfgrav.py:66-88 (f64 leapfrog KDK, all-pairs)
* `allpairs_accel` — test_pm.py:32-41
Constants BOX, M_KOERPER, EPS, E_SCHRANKE — test_selfgrav.py:33-40.
"""
from __future__ import annotations

import numpy as np

G_AU = 4 * np.pi * np.pi
BOX = 100_000.0
M_KOERPER = 1.12e7
EPS = 100.0
E_SCHRANKE = 5e-4


def wolke(n: int, v_faktor: float, seed: int = 7):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-BOX / 2, BOX / 2, n)
    y = rng.uniform(-BOX / 2, BOX / 2, n)
    mass = np.full(n, M_KOERPER)
    v_typ = np.sqrt(G_AU * mass.sum() / (BOX / 2))
    th = rng.uniform(0, 2 * np.pi, n)
    return (x, y, v_typ * v_faktor * np.cos(th),
            v_typ * v_faktor * np.sin(th), mass)


def energie(x, y, vx, vy, mass, eps2, G=G_AU):
    kin = 0.5 * (mass * (vx * vx + vy * vy)).sum()
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    r = np.sqrt(dx * dx + dy * dy + eps2)
    pot = -0.5 * G * (mass[None, :] * mass[:, None] / r).sum()
    pot += 0.5 * G * (mass * mass / np.sqrt(eps2)).sum()
    return kin + pot


def referenz(x, y, vx, vy, mass, eps2, dt, schritte, G=G_AU):
    rx, ry = x.copy(), y.copy()
    rvx, rvy = vx.copy(), vy.copy()

    def kraft(px, py):
        dx = px[None, :] - px[:, None]
        dy = py[None, :] - py[:, None]
        inv3 = (dx * dx + dy * dy + eps2) ** -1.5
        return (G * mass[None, :] * inv3 * dx).sum(1), \
               (G * mass[None, :] * inv3 * dy).sum(1)

    ax, ay = kraft(rx, ry)
    for _ in range(schritte):
        rvx += ax * (0.5 * dt)
        rvy += ay * (0.5 * dt)
        rx += rvx * dt
        ry += rvy * dt
        ax, ay = kraft(rx, ry)
        rvx += ax * (0.5 * dt)
        rvy += ay * (0.5 * dt)
    return rx, ry, rvx, rvy


def allpairs_accel(x, y, m, eps2, G=G_AU):
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    r2 = dx * dx + dy * dy + eps2
    inv = 1.0 / (r2 * np.sqrt(r2))
    ax = G * (m[None, :] * dx * inv).sum(axis=1)
    ay = G * (m[None, :] * dy * inv).sum(axis=1)
    return ax, ay
