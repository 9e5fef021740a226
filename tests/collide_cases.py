"""Collision scenes: the four of test_erkennung_streifen.py:76-111 (seed
4711) and the asteroid cloud of the 50k-asteroid configuration."""
import numpy as np

DT_Y = 0.5 / 365.25
G = 1.0


def stripe_scenes():
    rng = np.random.default_rng(4711)
    out = []
    nb = 20_000
    out.append(("wolke", rng.uniform(-3, 3, nb), rng.uniform(-3, 3, nb),
                rng.normal(0, 4, nb), rng.normal(0, 4, nb), np.full(nb, 3e-3)))
    nb = 8_000
    out.append(("klumpen", rng.normal(2.0, 0.02, nb), rng.normal(0.0, 0.02, nb),
                rng.normal(0, 6, nb), rng.normal(0, 6, nb), np.full(nb, 2e-4)))
    nb = 5_000
    x = np.concatenate([rng.normal(-4.0, 0.1, nb), rng.normal(4.0, 0.1, nb)])
    y = rng.normal(0.0, 0.1, 2 * nb)
    out.append(("zwei-wolken", x, y, rng.normal(0, 3, 2 * nb),
                rng.normal(0, 3, 2 * nb), np.full(2 * nb, 1e-3)))
    out.append(("auf-der-grenze", np.array([5.0 - 5e-4, 5.0 + 5e-4]), np.zeros(2),
                np.array([0.5, -0.5]), np.zeros(2), np.full(2, 1e-4)))
    return out


def cell_size(vx, vy, rr):
    """h and halo as bounce_suche derives them (film_producer.py:357-362)."""
    v95 = float(np.percentile(np.hypot(np.float32(vx), np.float32(vy)), 95))
    h = max(1e-4, 2.0 * v95 * DT_Y)
    return h, 2.0 * h + 2.0 * float(np.max(rr))


def asteroid_cloud(n=50_000, seed=4):
    """50,000 asteroids in a Gaussian cloud (sigma 0.05 AU) at 2.5 AU with
    bulk circular velocity (G = 1 units, star mass 1), dispersion 0.02 of
    it, radii log-uniform in [1e-7, 1e-5] AU."""
    rng = np.random.default_rng(seed)
    x = 2.5 + rng.normal(0, 0.05, n)
    y = rng.normal(0, 0.05, n)
    vc = np.sqrt(G / 2.5)
    vx = rng.normal(0, 0.02 * vc, n)
    vy = vc + rng.normal(0, 0.02 * vc, n)
    rr = np.exp(rng.uniform(np.log(1e-7), np.log(1e-5), n))
    return x, y, vx, vy, rr
