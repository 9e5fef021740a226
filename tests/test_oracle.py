"""Pin the CPU oracle before trusting it (no GPU needed).

Each check is chosen so that a plausible mistake in the oracle — a dropped
softening term, a wrong sign, a swapped source/target index, a wrong
Yoshida weight — fails at least one of them.
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle
from oracle import selfgrav_np as snp

getcontext().prec = 60
G_AU = 4 * np.pi ** 2


def _exact_accel(pos, m, G, eps2):
    """Decimal brute force of a_i = G sum_j m_j d/(|d|^2+eps2)^1.5."""
    dim, n = len(pos), len(m)
    P = [[Decimal(float(pos[c][i])) for c in range(dim)] for i in range(n)]
    M = [Decimal(float(v)) for v in m]
    E, GG = Decimal(float(eps2)), Decimal(float(G))
    out = np.zeros((dim, n))
    for i in range(n):
        s = [Decimal(0)] * dim
        for j in range(n):
            d = [P[j][c] - P[i][c] for c in range(dim)]
            r2 = sum(x * x for x in d) + E
            f = M[j] / (r2 * r2.sqrt())
            s = [s[c] + f * d[c] for c in range(dim)]
        out[:, i] = [float(GG * v) for v in s]
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_two_body_closed_form(dim):
    d = np.array([0.3, -0.2, 0.7])[:dim]
    pos = [np.array([0.1, 0.1 + d[c]]) for c in range(dim)]
    m = np.array([2.0, 5.0])
    eps2 = 1e-3
    a = np.array(oracle.allpairs(pos, m, pos, 1.7, eps2))
    r2 = float(d @ d) + eps2
    a0 = 1.7 * 5.0 * d / r2 ** 1.5      # body 0 pulled toward body 1
    a1 = -1.7 * 2.0 * d / r2 ** 1.5
    np.testing.assert_allclose(a[:, 0], a0, rtol=1e-15)
    np.testing.assert_allclose(a[:, 1], a1, rtol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_decimal_bruteforce(dim):
    rng = np.random.default_rng(11 + dim)
    n = 9
    pos = [rng.uniform(-1, 1, n) for _ in range(dim)]
    m = rng.uniform(0.1, 2.0, n)
    got = np.array(oracle.allpairs(pos, m, pos, G_AU, 1e-4))
    ref = _exact_accel(pos, m, G_AU, 1e-4)
    rel = np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0)
    assert rel.max() < 4e-16


def test_momentum_antisymmetry():
    rng = np.random.default_rng(5)
    n = 300
    pos = [rng.uniform(-1, 1, n) for _ in range(3)]
    m = rng.uniform(0.5, 3.0, n)
    a = np.array(oracle.allpairs(pos, m, pos, 1.0, 1e-4))
    p = (m * a).sum(axis=1)
    scale = (m * np.linalg.norm(a, axis=0)).sum()
    assert np.linalg.norm(p) / scale < 1e-15


def test_symmetric_cube_center():
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)],
                       float)
    pts = np.vstack([corners, [[0.5, 0.5, 0.5]]])
    pos = [pts[:, c] for c in range(3)]
    m = np.ones(9)
    a = np.array(oracle.allpairs(pos, m, pos, 1.0, 1e-2))
    assert np.abs(a[:, 8]).max() < 1e-16 * np.abs(a[:, 0]).max()
    # a corner is pulled toward the centre along the diagonal (+,+,+)
    assert np.all(a[:, 0] > 0)


def test_targets_are_not_sources():
    """Tracers (targets only) feel the sources but do not source."""
    pos = [np.array([0.0]), np.array([0.0])]
    tgt = [np.array([2.0, 3.0]), np.array([0.0, 0.0])]
    a = np.array(oracle.allpairs(pos, np.array([1.0]), tgt, 1.0, 1e-8))
    np.testing.assert_allclose(a[0], [-1 / 4, -1 / 9], rtol=1e-7)


def test_energy_two_body_closed_form():
    pos = [np.array([0.0, 1.0]), np.array([0.0, 0.5]), np.array([0.0, -0.25])]
    vel = [np.array([0.1, -0.2]), np.array([0.3, 0.0]), np.array([0.0, 0.4])]
    m = np.array([3.0, 0.5])
    eps2 = 0.01
    e = oracle.energy(pos, vel, m, 2.0, eps2)
    ekin = 0.5 * 3.0 * (0.01 + 0.09) + 0.5 * 0.5 * (0.04 + 0.16)
    epot = -2.0 * 3.0 * 0.5 / np.sqrt(1 + 0.25 + 0.0625 + eps2)
    np.testing.assert_allclose(e, [ekin, epot, ekin + epot], rtol=1e-15)


def test_energy_matches_reference_energie():
    x, y, vx, vy, m = snp.wolke(400, 0.4)
    eps2 = snp.EPS ** 2
    e = oracle.energy([x, y], [vx, vy], m, G_AU, eps2)[2]
    ref = snp.energie(x, y, vx, vy, m, eps2)
    assert abs(e - ref) < 1e-12 * abs(ref)


def test_kdk_matches_reference_referenz():
    """The oracle integrator equals test_selfgrav.referenz up to summation
    order (1e-16 per force evaluation) over 40 KDK steps."""
    x, y, vx, vy, m = snp.wolke(64, 0.5, seed=3)
    eps2 = snp.EPS ** 2
    rx, ry, rvx, rvy = snp.referenz(x, y, vx, vy, m, eps2, 0.0055, 40)
    (ox, oy), (ovx, ovy), _ = oracle.selfgrav_run([x, y], [vx, vy], m, None,
                                                  None, G_AU, eps2, 0.0055, 40)
    scale = np.abs(x).max()
    assert np.abs(ox - rx).max() < 1e-11 * scale
    assert np.abs(oy - ry).max() < 1e-11 * scale
    vs = np.abs(rvx).max()
    assert np.abs(ovx - rvx).max() < 1e-10 * vs
    assert np.abs(ovy - rvy).max() < 1e-10 * vs


def _circular_pair(steps, scheme, periods=1.0):
    m, r = 1.0, 1.0
    v = np.sqrt(G_AU * m / (4 * r))
    period = 2 * np.pi * r / v
    pos = [np.array([-r, r]), np.zeros(2)]
    vel = [np.zeros(2), np.array([-v, v])]
    (x, y), _, _ = oracle.selfgrav_run(pos, vel, np.array([m, m]), None, None,
                                       G_AU, 1e-8, periods * period / steps,
                                       steps, scheme=scheme)
    return float(np.hypot(x[0] + r, y[0]))


def test_reference_A_two_body_orbit_closes():
    """test_selfgrav.py:101-123: one period in 20,000 KDK steps closes to
    < 1e-4 * 2r (there eps = 1e-4 AU)."""
    m, r = 1.0, 1.0
    v = np.sqrt(G_AU * m / (4 * r))
    period = 2 * np.pi * r / v
    pos = [np.array([-r, r]), np.zeros(2)]
    vel = [np.zeros(2), np.array([-v, v])]
    (x, y), _, _ = oracle.selfgrav_run(pos, vel, np.array([m, m]), None, None,
                                       G_AU, 1e-8, period / 20000, 20000)
    assert np.hypot(x[0] + r, y[0]) < 1e-4 * 2 * r


def test_reference_F_tracer_orbit_and_fixed_mass():
    """test_selfgrav.py:227-258: a tracer on a circular orbit returns within
    1e-3 r after one period and the central mass does not move at all."""
    m, r = 1.0, 5.0
    v = np.sqrt(G_AU * m / r)
    period = 2 * np.pi * r / v
    (x, y), _, _ = oracle.selfgrav_run(
        [np.array([0.0]), np.array([0.0])], [np.array([0.0]), np.array([0.0])],
        np.array([m]), [np.array([r]), np.array([0.0])],
        [np.array([0.0]), np.array([v])], G_AU, 1e-8, period / 20000, 20000)
    assert np.hypot(x[1] - r, y[1]) < 1e-3 * r
    assert x[0] == 0.0 and y[0] == 0.0


def test_yoshida_is_fourth_order_and_kdk_second():
    """Halving dt cuts the orbit-closure error ~16x for Yoshida-4 and ~4x
    for KDK; a wrong composition weight drops Yoshida to 2nd order."""
    e1 = _circular_pair(400, 1)
    e2 = _circular_pair(800, 1)
    assert 11.0 < e1 / e2 < 22.0
    k1 = _circular_pair(400, 0)
    k2 = _circular_pair(800, 0)
    assert 3.0 < k1 / k2 < 5.5


@pytest.mark.parametrize("v_faktor", [0.4, 0.7])
def test_reference_B_energy_conservation(v_faktor):
    """test_selfgrav.py:126-152: N=400, dt=0.0055, 8000 KDK steps,
    |dE/E| < 5e-4 for the collapsing and the virialised cloud."""
    x, y, vx, vy, m = snp.wolke(400, v_faktor)
    eps2 = snp.EPS ** 2
    e0 = snp.energie(x, y, vx, vy, m, eps2)
    (ox, oy), (ovx, ovy), _ = oracle.selfgrav_run([x, y], [vx, vy], m, None,
                                                  None, G_AU, eps2, 0.0055, 8000)
    e1 = snp.energie(ox, oy, ovx, ovy, m, eps2)
    assert abs((e1 - e0) / e0) < snp.E_SCHRANKE
