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


def test_split_bins_brute_force_and_tsc_consistency():
    """split_bins against a brute-force binning (nearest node found by
    search, Python's sort of (block, node in block, index) tuples), and the
    deposit summed bin by bin over every node's 27 neighbouring bins equals
    tsc_deposit (so a body's bin is always its stencil's centre)."""
    rng = np.random.default_rng(11)
    n, h, B, nbody = 12, 0.1, 5, 2000
    origin = [0.2, -0.3, 0.05]
    pos = [origin[c] + rng.uniform(1.0, n - 2.0, nbody) * h for c in range(3)]
    m = rng.uniform(0.5, 2.0, nbody)
    start, cstart, order = S.split_bins(pos, n, origin, h, B)
    nb = -(-n // B)
    nodes = np.arange(n)
    recs = []
    for i in range(nbody):
        r = [int(nodes[np.argmin(np.abs((pos[c][i] - origin[c]) / h - nodes))])
             for c in range(3)]
        recs.append((tuple(rc // B for rc in r), tuple(rc % B for rc in r), i))
    recs.sort()
    assert [rec[2] for rec in recs] == list(order)
    # bin sizes in key order, block sizes in block order
    bins = {}
    for blk, loc, _ in recs:
        bins[(blk, loc)] = bins.get((blk, loc), 0) + 1
    keys = [(bi, bj, bk, li, lj, lk) for bi in range(nb) for bj in range(nb)
            for bk in range(nb) for li in range(B) for lj in range(B) for lk in range(B)]
    sizes = [bins.get((k[:3], k[3:]), 0) for k in keys]
    assert np.array_equal(np.diff(start), sizes)
    assert np.array_equal(np.diff(cstart), np.add.reduceat(sizes, np.arange(0, len(keys), B ** 3)))
    # the blocks alone: Python's sort of (block, index) tuples, same starts
    none, cstart_b, order_b = S.split_bins(pos, n, origin, h, B, nodes=False)
    assert none is None and np.array_equal(cstart_b, cstart)
    assert [i for _, i in sorted((blk, i) for blk, _, i in recs)] == list(order_b)
    # deposit from the bins
    mesh = np.zeros((n, n, n))
    key_of = {k: q for q, k in enumerate(keys)}
    for i in range(n):
        for j in range(n):
            for k in range(n):
                acc = 0.0
                for a in (-1, 0, 1):
                    for b in (-1, 0, 1):
                        for c in (-1, 0, 1):
                            node = (i + a, j + b, k + c)
                            if min(node) < 0 or max(node) >= n:
                                continue
                            q = key_of[tuple(x // B for x in node) + tuple(x % B for x in node)]
                            for s in order[start[q]:start[q + 1]]:
                                w = 1.0
                                for ax, nd in zip(range(3), (i, j, k)):
                                    f = (pos[ax][s] - origin[ax]) / h
                                    cc, wts = S.tsc_weights(np.array([f]))
                                    off = nd - int(cc[0])
                                    assert -1 <= off <= 1
                                    w *= wts[off + 1][0]
                                acc += m[s] * w
                mesh[i, j, k] = acc
    np.testing.assert_allclose(mesh, S.tsc_deposit(pos, m, n, origin, h), rtol=1e-12,
                               atol=1e-12)
