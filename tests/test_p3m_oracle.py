"""Pin the cell-list, split-function and TSC oracle (CPU, no GPU)."""
import math

import numpy as np
from scipy import integrate

import oracle
from oracle import p3m_np as S


def test_cell_list_matches_brute_force():
    rng = np.random.default_rng(1)
    pos = [rng.uniform(-0.3, 1.4, 500) for _ in range(3)]
    origin, cell, nc = [0.0, -0.1, 0.05], 0.13, [7, 9, 6]
    start, order = S.cell_list(pos, origin, cell, nc)
    keys = []
    for i in range(500):
        k = 0
        for c in range(3):
            v = int(math.floor((pos[c][i] - origin[c]) / cell))
            k = k * nc[c] + min(max(v, 0), nc[c] - 1)
        keys.append(k)
    expect = sorted(range(500), key=lambda i: (keys[i], i))
    assert order.tolist() == expect
    for c in range(int(np.prod(nc))):
        assert start[c + 1] - start[c] == keys.count(c)


def test_near_field_with_huge_rs_is_allpairs():
    rng = np.random.default_rng(2)
    src = [rng.uniform(0, 1, 300) for _ in range(3)]
    m = rng.uniform(0.5, 1.5, 300)
    tgt = [rng.uniform(0, 1, 50) for _ in range(3)]
    got = S.near_field(src, m, tgt, 1.0, 1e-4, rs=1e5, rcut=3e5)
    ref = np.array(oracle.allpairs(src, m, tgt, 1.0, 1e-4))
    np.testing.assert_allclose(got, ref, rtol=1e-9)


def test_long_range_part_is_gaussian_enclosed_mass():
    """1 - g(r) = fraction of a Gaussian cloud exp(-s^2 / 4 rs^2) inside r
    (the long-range force is that of the smoothed mass)."""
    rs = 0.7
    norm = integrate.quad(lambda s: 4 * math.pi * s * s * math.exp(-s * s / (4 * rs * rs)),
                          0, 60 * rs)[0]
    for r in (0.05, 0.3, 1.0, 2.5, 4.0):
        enc = integrate.quad(lambda s: 4 * math.pi * s * s *
                             math.exp(-s * s / (4 * rs * rs)), 0, r)[0] / norm
        assert abs((1 - S.split_g(np.array([r]), rs)[0]) - enc) < 1e-10


def test_taper_is_c2_and_bounded():
    rs, rcut = 0.4, 2.0
    assert S.taper(np.array([0.5, rcut - rs]), rs, rcut).tolist() == [1.0, 1.0]
    assert abs(S.taper(np.array([rcut]), rs, rcut)[0]) < 1e-12
    r = np.linspace(rcut - rs - 0.05, rcut + 0.05, 2001)
    t = S.taper(r, rs, rcut)
    d1 = np.gradient(t, r)
    d2 = np.gradient(d1, r)
    assert np.all(np.diff(t) <= 1e-15) and t.min() >= 0.0 and t.max() <= 1.0
    # no jumps in the slope or the curvature at either end
    assert np.abs(np.diff(d1)).max() < 0.02 and np.abs(np.diff(d2)).max() < 0.5
    assert np.all(S.split_gs(np.array([rcut, rcut + 0.1]), rs, rcut) == 0.0)


def test_tsc_mass_node_weights_linear_and_adjoint():
    n, h, origin = 12, 0.1, [0.2, -0.3, 0.0]
    rng = np.random.default_rng(3)
    pos = [origin[a] + rng.uniform(2.0, n - 3.0, 400) * h for a in range(3)]
    m = rng.uniform(0.5, 2.0, 400)
    rho = S.tsc_deposit(pos, m, n, origin, h)
    assert abs(rho.sum() - m.sum()) < 1e-12 * m.sum()
    # a body on node (4, 5, 6): 3/4 per axis there, 1/8 on each side
    one = [np.array([origin[0] + 4 * h]), np.array([origin[1] + 5 * h]),
           np.array([origin[2] + 6 * h])]
    r1 = S.tsc_deposit(one, [1.0], n, origin, h)
    assert abs(r1[4, 5, 6] - 0.75 ** 3) < 1e-12
    assert abs(r1[3, 5, 6] - 0.125 * 0.75 ** 2) < 1e-12
    assert abs(r1[3, 4, 6] - 0.125 ** 2 * 0.75) < 1e-12
    assert abs(r1[3, 4, 7] - 0.125 ** 3) < 1e-12
    # linear fields are interpolated exactly
    I, J, K = np.meshgrid(*[origin[a] + np.arange(n) * h for a in range(3)],
                          indexing="ij")
    lin = 0.3 + 2.0 * I - 1.5 * J + 0.25 * K
    got = S.tsc_gather(lin, pos, n, origin, h)
    np.testing.assert_allclose(got, 0.3 + 2.0 * pos[0] - 1.5 * pos[1] + 0.25 * pos[2],
                               rtol=1e-12, atol=1e-12)
    # gather is the adjoint of deposit
    f = rng.normal(size=(n, n, n))
    assert abs((rho * f).sum() - (m * S.tsc_gather(f, pos, n, origin, h)).sum()) < 1e-10
