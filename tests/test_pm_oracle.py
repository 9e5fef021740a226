"""Pin the particle-mesh oracle (CPU only): closed forms, brute force and
the reference's own test_pm.py bounds."""
import math

import numpy as np
import pytest

import oracle
from oracle import pm_np

G_AU = 4 * np.pi ** 2


def test_cic_cells_brute_force():
    rng = np.random.default_rng(0)
    n = [16, 8, 32]
    origin = [-1.0, 0.5, 2.0]
    h = 0.0625
    pos = [rng.uniform(o - 3 * h, o + (k + 2) * h, 500) for o, k in zip(origin, n)]
    pos[0][:4] = [origin[0], origin[0] + h, origin[0] + (n[0] - 1) * h, -1e3]
    for periodic in (False, True):
        cells, ts = pm_np.cic_cells(pos, origin, h, n, periodic)
        for c in range(3):
            for p in range(len(pos[c])):
                f = (pos[c][p] - origin[c]) / h
                i = math.floor(f)
                if not periodic:
                    i = min(max(i, 0), n[c] - 2)
                    assert ts[c][p] == f - i
                else:
                    assert ts[c][p] == f - i
                    i %= n[c]
                assert cells[c][p] == i


@pytest.mark.parametrize("periodic", [False, True])
def test_deposit_conserves_mass_and_node_hit(periodic):
    rng = np.random.default_rng(1)
    pos = [rng.uniform(0.1, 0.9, 1000) for _ in range(3)]
    m = rng.uniform(0.5, 2.0, 1000)
    rho = pm_np.cic_deposit(pos, m, [16, 16, 16], [0, 0, 0], 1 / 16, periodic)
    assert abs(rho.sum() - m.sum()) < 1e-13 * m.sum()
    # a body sitting exactly on a node puts all its mass there
    one = pm_np.cic_deposit([np.array([0.25]), np.array([0.5]), np.array([0.75])],
                            np.array([2.0]), [16, 16, 16], [0, 0, 0], 1 / 16, periodic)
    assert one[4, 8, 12] == 2.0 and one.sum() == 2.0
    # gather of a linear field is exact for CIC (interior points)
    gx = np.arange(16)[:, None, None] * (1 / 16) + np.zeros((16, 16, 16))
    inner = [rng.uniform(0.2, 0.8, 50) for _ in range(3)]
    got = pm_np.cic_gather(gx, inner, [16] * 3, [0, 0, 0], 1 / 16, periodic)
    np.testing.assert_allclose(got, inner[0], rtol=1e-14)


def test_isolated_convolution_equals_direct_sum():
    """FFT convolution on the zero-padded mesh == brute-force sum over node
    pairs with the force Green's function (checks padding and wrap)."""
    rng = np.random.default_rng(2)
    N, h, eps2, G = 6, 0.3, 0.05, 1.3
    for dim in (2, 3):
        rho = rng.uniform(0, 1, [N] * dim)
        kern = pm_np.build_force_kernels(N, h, eps2, G, dim)
        P = 2 * N
        pad = np.zeros([P] * dim)
        pad[tuple(slice(0, N) for _ in range(dim))] = rho
        fr = np.fft.rfftn(pad)
        sl = tuple(slice(0, N) for _ in range(dim))
        fields = [np.fft.irfftn(fr * k, s=[P] * dim, axes=list(range(dim)))[sl]
                  for k in kern]
        nodes = np.array(np.meshgrid(*[np.arange(N)] * dim, indexing="ij")).reshape(dim, -1).T
        r = rho.ravel()
        for pi in range(0, len(nodes), 7):
            d = (nodes[pi] - nodes) * h              # field - source
            r2 = (d * d).sum(1) + eps2
            acc = (-G * d * (r / (r2 * np.sqrt(r2)))[:, None]).sum(0)
            got = [f.ravel()[pi] for f in fields]
            np.testing.assert_allclose(got, acc, rtol=1e-10, atol=1e-12)


def test_far_pair_matches_newton():
    """test_pm.py:61-78: two masses 60 cells apart, PM within 2 % of G m/r^2."""
    grid_n, h = 512, 100.0
    eps2 = (2 * h) ** 2
    m_val, r = 1e10, 60 * h
    cx = grid_n * h / 2
    x = np.array([cx - r / 2, cx + r / 2])
    y = np.array([cx, cx])
    ax, _ = pm_np.pm_accelerations([x, y], np.array([m_val, m_val]), grid_n,
                                   [0.0, 0.0], h, eps2)
    analytic = G_AU * m_val / r ** 2
    assert abs(abs(ax[0]) - analytic) / analytic < 0.02


def test_momentum_conservation():
    """test_pm.py:44-58: |sum m a| / sum m |a| < 1e-3."""
    rng = np.random.default_rng(42)
    n = 5000
    x = rng.uniform(-40000, 40000, n)
    y = rng.uniform(-40000, 40000, n)
    m = rng.uniform(1e6, 2e7, n)
    origin, h = pm_np.grid_fuer([x, y], 512)
    ax, ay = pm_np.pm_accelerations([x, y], m, 512, origin, h, (2 * h) ** 2)
    rel = np.hypot((m * ax).sum(), (m * ay).sum()) / (m * np.hypot(ax, ay)).sum()
    assert rel < 1e-3


@pytest.mark.parametrize("eps_in_h,bound", [(4.0, 0.03), (1.0, 0.12)])
def test_cloud_vs_allpairs(eps_in_h, bound):
    """test_pm.py:81-100: median relative error vs all-pairs."""
    rng = np.random.default_rng(7)
    n = 4000
    x = rng.uniform(-40000, 40000, n)
    y = rng.uniform(-40000, 40000, n)
    m = rng.uniform(1e6, 2e7, n)
    origin, h = pm_np.grid_fuer([x, y], 512)
    eps2 = (eps_in_h * h) ** 2
    ax, ay = pm_np.pm_accelerations([x, y], m, 512, origin, h, eps2)
    rx, ry = oracle.allpairs([x, y], m, [x, y], G_AU, eps2)
    rel = np.hypot(ax - rx, ay - ry) / np.hypot(rx, ry)
    assert np.median(rel) < bound


@pytest.mark.parametrize("dim", [2, 3])
def test_periodic_poisson_single_mode(dim):
    """rho = rho0 (1 + a cos(2 pi q x / L)): the spectral solve gives phi
    exactly on the nodes, so the central difference must equal
    -4 pi G rho0 a sin(k x) / k * sin(k h)/(k h)."""
    n, L, G, q, a = 32, 2.0, 0.7, 3, 0.25
    h = L / n
    xs = np.arange(n) * h
    grids = np.meshgrid(*[xs] * dim, indexing="ij")
    rho0 = 1.5
    dens = rho0 * (1 + a * np.cos(2 * np.pi * q * grids[0] / L))
    # bodies exactly on nodes carry node mass dens * h^dim
    pos = [g.ravel() for g in grids]
    m = (dens * h ** dim).ravel()
    f = pm_np.periodic_field(pos, m, n, L, [0.0] * dim, G)
    k = 2 * np.pi * q / L
    expect = -4 * np.pi * G * rho0 * a / k * np.sin(k * grids[0]) \
        * np.sin(k * h) / (k * h)
    # phi = -4 pi G rho0 a cos(kx) / k^2, a_x = -dphi/dx points back to the
    # overdensity at x = 0 (negative for small x > 0)
    assert f[0].ravel()[dim == 2 and n or n * n] < 0
    np.testing.assert_allclose(f[0], expect, atol=1e-12 * np.abs(expect).max())
    for c in range(1, dim):
        assert np.abs(f[c]).max() < 1e-12 * np.abs(expect).max()


def test_periodic_kdk_reversible_and_second_order():
    """The periodic KDK restatement is time-reversible (a wrong kick weight
    or order breaks that), conserves momentum, and its error falls ~4x when
    dt halves (second order); Yoshida weights give a much smaller error."""
    rng = np.random.default_rng(5)
    npart, n = 800, 16
    pos = [rng.uniform(0.0, 1.0, npart) for _ in range(3)]
    vel = [rng.normal(0.0, 0.05, npart) for _ in range(3)]
    m = np.full(npart, 1.0 / npart)
    o = [0.0, 0.0, 0.0]
    x1, v1 = pm_np.periodic_kdk(pos, vel, m, n, 1.0, o, 1.0, 1e-2, 4)
    x2, v2 = pm_np.periodic_kdk(x1, [-c for c in v1], m, n, 1.0, o, 1.0, 1e-2, 4)
    dx = np.array(x2) - np.array(pos)
    dx -= np.round(dx)
    assert np.abs(dx).max() < 1e-12
    assert np.abs(np.array(v2) + np.array(vel)).max() < 1e-12
    p0 = np.sum(m * np.array(vel), axis=1)
    assert np.abs(np.sum(m * np.array(v1), axis=1) - p0).max() < 1e-12
    ref, _ = pm_np.periodic_kdk(pos, vel, m, n, 1.0, o, 1.0, 0.02 / 64, 64)

    def err(dt, k, w=(1.0,)):
        x, _ = pm_np.periodic_kdk(pos, vel, m, n, 1.0, o, 1.0, dt, k, w)
        d = np.array(x) - np.array(ref)
        return np.abs(d - np.round(d)).max()
    e1, e2 = err(0.01, 2), err(0.005, 4)
    assert 3.0 < e1 / e2 < 5.0
    w1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
    w0 = -(2.0 ** (1.0 / 3.0)) / (2.0 - 2.0 ** (1.0 / 3.0))
    assert err(0.01, 2, (w1, w0, w1)) < 0.2 * e1


def test_nbodypm_grid_limits_are_clear_errors():
    """NBodyPM beyond the longest FFT line (grid_n = 8192: 16384-point padded
    lines) or with a grid that is not a power of two raises a ValueError
    that names the limit, before any device work; the reference's default
    grid_n = 4096 (8192-point lines) is within it (tests/test_pm_gpu.py)."""
    import pytest
    from solarsystemsimulator_amd import pm as P
    with pytest.raises(ValueError, match="grid_n <= 4096"):
        P.NBodyPM([0], grid_n=8192)
    with pytest.raises(ValueError, match="power of two"):
        P.NBodyPM([0], grid_n=1000)
