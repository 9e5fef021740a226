"""Seeded synthetic scenes of the benchmark configurations (G = 1 units).

* `perturbed_cube` — equal-mass bodies in the unit cube: one uniformly
  jittered point per cell of an nx*ny*nz lattice, displaced by a Zel'dovich
  field of a Gaussian random density with P(k) ~ k^-2, scaled so the rms
  density contrast is `amp`.  Cold start.  `periodic=True` wraps positions
  into [0, 1).  Used for the 65,536-body all-pairs case (64x32x32, seed 2),
  the 262,144-body particle-mesh case (64^3, seed 3, periodic) and the
  250k-mass + 500k-tracer case (100x50x50 seed 5 / 100x100x50 seed 55).
* `solar_system` — star + the 8 planets on Kepler ellipses + test
  particles in the asteroid belt (2-D, AU, years, G = 4 pi^2).
* `asteroid_cloud` — star + 4 planets + a Gaussian asteroid cloud at
  2.5 AU for the collision search.
"""
from __future__ import annotations

import numpy as np


def _zeldovich(shape, rng, amp):
    """Displacement field psi (3, *shape) in lattice units with rms div = amp."""
    kx = np.fft.fftfreq(shape[0]) * 2 * np.pi
    ky = np.fft.fftfreq(shape[1]) * 2 * np.pi
    kz = np.fft.rfftfreq(shape[2]) * 2 * np.pi
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    white = np.fft.rfftn(rng.standard_normal(shape))
    delta_k = white / np.sqrt(k2)          # P(k) ~ k^-2
    delta_k[0, 0, 0] = 0.0
    delta = np.fft.irfftn(delta_k, s=shape, axes=(0, 1, 2))
    delta_k *= amp / delta.std()
    # div psi = -delta  =>  psi_k = i k delta_k / k^2
    psi = [np.fft.irfftn(1j * K * delta_k / k2, s=shape, axes=(0, 1, 2))
           for K in (KX, KY, KZ)]
    return np.stack(psi)


def perturbed_cube(nx: int, ny: int, nz: int, seed: int, amp: float = 0.05,
                   periodic: bool = False):
    """(x, y, z) float64 arrays of nx*ny*nz bodies in the unit cube."""
    rng = np.random.default_rng(seed)
    shape = (nx, ny, nz)
    idx = np.indices(shape).reshape(3, -1).astype(np.float64)
    jit = rng.uniform(0.0, 1.0, idx.shape)
    psi = _zeldovich(shape, rng, amp).reshape(3, -1)
    scale = np.array([nx, ny, nz], np.float64)[:, None]
    pos = (idx + jit + psi) / scale
    if periodic:
        pos = np.mod(pos, 1.0)
    return pos[0].copy(), pos[1].copy(), pos[2].copy()


def selfgrav_cube(n_shape=(64, 32, 32), seed=2, amp=0.05):
    """All-pairs case: equal masses (total 1) on a perturbed lattice."""
    x, y, z = perturbed_cube(*n_shape, seed=seed, amp=amp)
    m = np.full(x.size, 1.0 / x.size)
    return (x, y, z), m


def learned_scene(mass_shape=(100, 50, 50), tracer_shape=(100, 100, 50),
                  seed=5, tracer_seed=55, amp=0.05):
    """250k masses + 500k tracers (default sizes), cold start."""
    pos = perturbed_cube(*mass_shape, seed=seed, amp=amp)
    tpos = perturbed_cube(*tracer_shape, seed=tracer_seed, amp=amp)
    m = np.full(pos[0].size, 1.0 / pos[0].size)
    return pos, m, tpos


# planet masses (solar masses) and semi-major axes (AU), Mercury..Neptune
PLANETS = ((1.66e-7, 0.387), (2.45e-6, 0.723), (3.00e-6, 1.000),
           (3.23e-7, 1.524), (9.55e-4, 5.203), (2.86e-4, 9.537),
           (4.37e-5, 19.19), (5.15e-5, 30.07))
G_AU = 4.0 * np.pi * np.pi


def _kepler_2d(a, e, nu, omega, gm):
    """Position and velocity on the ellipse (a, e) at true anomaly nu, with
    periapsis direction omega (2-D)."""
    p = a * (1.0 - e * e)
    r = p / (1.0 + e * np.cos(nu))
    vr = np.sqrt(gm / p) * e * np.sin(nu)
    vt = np.sqrt(gm / p) * (1.0 + e * np.cos(nu))
    ang = nu + omega
    c, s = np.cos(ang), np.sin(ang)
    return r * c, r * s, vr * c - vt * s, vr * s + vt * c


def solar_system(n_test: int = 10_000, seed: int = 1):
    """(x, y, vx, vy, mass, visible, is_ast): star of mass 1 at rest, the
    8 planets with e ~ U[0, 0.1] and random phases, n_test massless test
    particles with a ~ U[2.1, 3.3] AU, e ~ U[0, 0.2], random angles."""
    rng = np.random.default_rng(seed)
    m = np.array([1.0] + [pm for pm, _ in PLANETS] + [0.0] * n_test)
    a = np.array([pa for _, pa in PLANETS] + list(rng.uniform(2.1, 3.3, n_test)))
    e = np.concatenate([rng.uniform(0.0, 0.1, 8), rng.uniform(0.0, 0.2, n_test)])
    nu = rng.uniform(0.0, 2 * np.pi, 8 + n_test)
    om = rng.uniform(0.0, 2 * np.pi, 8 + n_test)
    x, y, vx, vy = _kepler_2d(a, e, nu, om, G_AU)
    x, y = np.concatenate([[0.0], x]), np.concatenate([[0.0], y])
    vx, vy = np.concatenate([[0.0], vx]), np.concatenate([[0.0], vy])
    n = len(m)
    visible = np.ones(n, np.uint8)
    is_ast = np.zeros(n, np.uint8)
    is_ast[9:] = 1
    return x, y, vx, vy, m, visible, is_ast


def _kepler_3d(a, e, inc, node, peri, nu, gm):
    """Position and velocity on the ellipse (a, e) at true anomaly nu, in
    the plane inclined by inc with ascending node `node` and argument of
    periapsis `peri`: the perifocal state rotated by Rz(node) Rx(inc)
    Rz(peri).  Returns (x, y, z, vx, vy, vz)."""
    p = a * (1.0 - e * e)
    r = p / (1.0 + e * np.cos(nu))
    xp, yp = r * np.cos(nu), r * np.sin(nu)
    s = np.sqrt(gm / p)
    vxp, vyp = -s * np.sin(nu), s * (e + np.cos(nu))
    cO, sO, ci, si = np.cos(node), np.sin(node), np.cos(inc), np.sin(inc)
    cw, sw = np.cos(peri), np.sin(peri)
    r11, r12 = cO * cw - sO * sw * ci, -cO * sw - sO * cw * ci
    r21, r22 = sO * cw + cO * sw * ci, -sO * sw + cO * cw * ci
    r31, r32 = sw * si, cw * si
    return (r11 * xp + r12 * yp, r21 * xp + r22 * yp, r31 * xp + r32 * yp,
            r11 * vxp + r12 * vyp, r21 * vxp + r22 * vyp, r31 * vxp + r32 * vyp)


def solar_system_3d(n_test: int = 10_000, seed: int = 1, G: float = 1.0):
    """The solar system in space, in units G = 1, AU, solar masses (time
    unit 1 yr / 2 pi): star of mass 1 at rest, the 8 planets with
    e ~ U[0, 0.1], i ~ U[0, 3 deg], n_test massless test particles with
    a ~ U[2.1, 3.3] AU, e ~ U[0, 0.2], i ~ U[0, 10 deg]; nodes, periapses
    and phases uniform.  Returns (x, y, z, vx, vy, vz, mass, visible,
    is_ast)."""
    rng = np.random.default_rng(seed)
    m = np.array([1.0] + [pm for pm, _ in PLANETS] + [0.0] * n_test)
    a = np.array([pa for _, pa in PLANETS] + list(rng.uniform(2.1, 3.3, n_test)))
    e = np.concatenate([rng.uniform(0.0, 0.1, 8), rng.uniform(0.0, 0.2, n_test)])
    inc = np.radians(np.concatenate([rng.uniform(0.0, 3.0, 8),
                                     rng.uniform(0.0, 10.0, n_test)]))
    node, peri, nu = (rng.uniform(0.0, 2 * np.pi, 8 + n_test) for _ in range(3))
    st = _kepler_3d(a, e, inc, node, peri, nu, G)
    x, y, z, vx, vy, vz = (np.concatenate([[0.0], c]) for c in st)
    n = len(m)
    visible = np.ones(n, np.uint8)
    is_ast = np.zeros(n, np.uint8)
    is_ast[9:] = 1
    return x, y, z, vx, vy, vz, m, visible, is_ast


def asteroid_cloud(n: int = 50_000, seed: int = 4):
    """Star (mass 1) + Mercury..Mars + n asteroids in a Gaussian cloud
    (sigma 0.05 AU) at 2.5 AU moving with the circular velocity, dispersion
    0.02 of it; radii log-uniform in [1e-7, 1e-5] AU, densities U[1, 3]
    (masses from radius and density in solar masses per AU^3 scaled to
    g/cm^3).  Returns (x, y, vx, vy, mass, radius, is_ast)."""
    rng = np.random.default_rng(seed)
    vc = np.sqrt(G_AU / 2.5)
    ax = 2.5 + rng.normal(0.0, 0.05, n)
    ay = rng.normal(0.0, 0.05, n)
    avx = rng.normal(0.0, 0.02 * vc, n)
    avy = vc + rng.normal(0.0, 0.02 * vc, n)
    rad = np.exp(rng.uniform(np.log(1e-7), np.log(1e-5), n))
    rho = rng.uniform(1.0, 3.0, n)                       # g / cm^3
    au_cm, msun_g = 1.495978707e13, 1.98847e33
    am = 4.0 / 3.0 * np.pi * (rad * au_cm) ** 3 * rho / msun_g
    px, py, pvx, pvy = _kepler_2d(np.array([pa for _, pa in PLANETS[:4]]),
                                  np.zeros(4), rng.uniform(0, 2 * np.pi, 4),
                                  np.zeros(4), G_AU)
    x = np.concatenate([[0.0], px, ax])
    y = np.concatenate([[0.0], py, ay])
    vx = np.concatenate([[0.0], pvx, avx])
    vy = np.concatenate([[0.0], pvy, avy])
    mass = np.concatenate([[1.0], [pm for pm, _ in PLANETS[:4]], am])
    radius = np.concatenate([[0.00465], [1.6e-5, 4.0e-5, 4.3e-5, 2.3e-5], rad])
    is_ast = np.concatenate([np.zeros(5, np.uint8), np.ones(n, np.uint8)])
    return x, y, vx, vy, mass, radius, is_ast


def _star_first(ax, ay, avx, avy) -> dict:
    """Sun (mass 1, radius 0.00465 AU) followed by the asteroids (mass
    1e-14, radius 2e-5 AU), the film benchmark's layout (bench_film.py:
    30-42)."""
    n_ast = len(ax)
    one = np.ones(n_ast + 1, np.uint8)
    return {"x": np.concatenate(([0.0], ax)), "y": np.concatenate(([0.0], ay)),
            "vx": np.concatenate(([0.0], avx)), "vy": np.concatenate(([0.0], avy)),
            "mass": np.concatenate(([1.0], np.full(n_ast, 1e-14))),
            "real_r": np.concatenate(([0.00465], np.full(n_ast, 2e-5))),
            "visible": one,
            "is_ast": np.concatenate(([0], np.ones(n_ast, np.uint8))),
            "is_star_bh": np.concatenate(([1], np.zeros(n_ast, np.uint8)))}


def film_belt(n_ast: int, seed: int = 20260720) -> dict:
    """The film benchmark's relaxed belt (bench_film.py:45-54): radii
    1.5 + 2.5 sqrt(U) AU, uniform angles, circular speed with 3 % scatter."""
    rng = np.random.default_rng(seed)
    r = 1.5 + 2.5 * np.sqrt(rng.random(n_ast))
    th = rng.random(n_ast) * 2 * np.pi
    v = np.sqrt(G_AU / r) * (1.0 + rng.normal(0, 0.03, n_ast))
    return _star_first(r * np.cos(th), r * np.sin(th), -v * np.sin(th), v * np.cos(th))


def film_clumps(n_ast: int, clumps: int = 5, seed: int = 20260720) -> dict:
    """The film benchmark's interpenetrating clouds (bench_film.py:57-83):
    `clumps` Gaussian clouds (sigma 0.12 AU, velocity scatter 0.4 AU/yr)
    with centres at r = 2 + 0.05 i AU, angle 0.35 i, moving with the
    circular speed; the dense case of the collision search."""
    rng = np.random.default_rng(seed)
    each = n_ast // clumps
    parts = [[], [], [], []]
    for i in range(clumps):
        m = each if i < clumps - 1 else n_ast - each * (clumps - 1)
        r0, th0 = 2.0 + 0.05 * i, 0.35 * i
        cx, cy = r0 * np.cos(th0), r0 * np.sin(th0)
        v0 = np.sqrt(G_AU / r0)
        parts[0].append(rng.normal(cx, 0.12, m))
        parts[1].append(rng.normal(cy, 0.12, m))
        parts[2].append(-v0 * np.sin(th0) + rng.normal(0, 0.4, m))
        parts[3].append(v0 * np.cos(th0) + rng.normal(0, 0.4, m))
    return _star_first(*(np.concatenate(p) for p in parts))
