"""NumPy restatement of the reference's nbody test helpers — TEST
INFRASTRUCTURE ONLY.

* `reference_advance` — test_kernel.py:16-81: Yoshida-4 with the global
  adaptive step of the JS worker (ASTAD: massive pairs AND asteroid x
  massive pairs set dt), every body integrated with that step.
* `make_system` — test_kernel.py:84-108 (seeded scene).
Constants follow nbody_kernel.py:58-65.
"""
from __future__ import annotations

import numpy as np

G_AU = 4 * np.pi * np.pi
SOFTENING = 1e-6
MAX_SUB_DT_YEARS = 0.5 / 365.25
YOSHIDA_W1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
YOSHIDA_W0 = -(2.0 ** (1.0 / 3.0)) / (2.0 - 2.0 ** (1.0 / 3.0))


def reference_advance(dt_years, x, y, vx, vy, mass, visible, is_ast):
    (x, y), (vx, vy) = advance_nd(dt_years, [x, y], [vx, vy], mass, visible,
                                  is_ast)
    return x, y, vx, vy


def advance_nd(dt_years, pos, vel, mass, visible, is_ast, G=G_AU,
               soft=SOFTENING, max_sub_dt=MAX_SUB_DT_YEARS):
    """`reference_advance` for 2 or 3 components (positions `pos`, velocities
    `vel`: lists of arrays); |d|^2 summed x, y, (z), and for 2 components
    exactly the reference's operations (hypot distances)."""
    dim = len(pos)
    p = [np.asarray(a, np.float64).copy() for a in pos]
    v = [np.asarray(a, np.float64).copy() for a in vel]
    vis = np.asarray(visible) != 0
    ast = np.asarray(is_ast) != 0
    mas_i = np.flatnonzero(~ast & vis)

    def norm(d):
        if dim == 2:
            return np.hypot(d[0], d[1])
        return np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])

    def accel():
        acc = [np.zeros_like(p[0]) for _ in range(dim)]
        for i in mas_i:
            d = [c - c[i] for c in p]
            r2 = d[0] * d[0]
            for c in range(1, dim):
                r2 = r2 + d[c] * d[c]
            r2 = r2 + soft
            f = G / (r2 * np.sqrt(r2))
            f[i] = 0.0
            f[~vis] = 0.0
            sel = np.ones_like(f, dtype=bool)
            sel[~vis] = False
            sel[i] = False
            sel[mas_i[mas_i < i]] = False
            for c in range(dim):
                acc[c][sel] += f[sel] * mass[i] * (-d[c][sel])
                acc[c][i] += np.sum(f[sel] * mass[sel] * d[c][sel])
        return acc

    def adapt_dt():
        dt = max_sub_dt
        for i in mas_i:
            sel = vis.copy()
            sel[i] = False
            sel[mas_i[mas_i < i]] = False
            dist = np.maximum(norm([c[sel] - c[i] for c in p]), 1e-12)
            vrel = norm([c[sel] - c[i] for c in v])
            ok = vrel > 1e-9
            if ok.any():
                dt = min(dt, float(np.min(dist[ok] / vrel[ok])) / 20.0)
        return max(dt, max_sub_dt / 1000.0)

    acc = accel()

    def verlet(dt):
        nonlocal acc
        for c in range(dim):
            p[c][vis] += v[c][vis] * dt + 0.5 * acc[c][vis] * dt * dt
        nacc = accel()
        for c in range(dim):
            v[c][vis] += 0.5 * (acc[c][vis] + nacc[c][vis]) * dt
        acc = nacc

    remaining, guard = dt_years, 0
    while remaining > 1e-12 and guard < 100_000:
        guard += 1
        sub = min(remaining, adapt_dt())
        for w in (YOSHIDA_W1, YOSHIDA_W0, YOSHIDA_W1):
            verlet(w * sub)
        remaining -= sub
    return p, v


def make_system(n_ast: int, seed: int = 7):
    rng = np.random.default_rng(seed)
    n = 10 + n_ast
    x = np.zeros(n)
    y = np.zeros(n)
    vx = np.zeros(n)
    vy = np.zeros(n)
    mass = np.full(n, 1e-6)
    mass[0] = 1.0
    is_ast = np.zeros(n, np.uint8)
    is_ast[10:] = 1
    visible = np.ones(n, np.uint8)
    visible[3] = 0
    rp = np.arange(1, 10, dtype=np.float64)
    x[1:10] = rp
    vy[1:10] = np.sqrt(G_AU / rp)
    r = 2 + 2 * rng.random(n_ast)
    th = rng.random(n_ast) * 2 * np.pi
    x[10:] = r * np.cos(th)
    y[10:] = r * np.sin(th)
    v = np.sqrt(G_AU / r)
    vx[10:] = -v * np.sin(th)
    vy[10:] = v * np.cos(th)
    mass[10:] = 1e-12
    return x, y, vx, vy, mass, visible, is_ast
