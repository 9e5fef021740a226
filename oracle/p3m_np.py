"""NumPy oracle of the cell list and the near-field (split) force — TEST
INFRASTRUCTURE ONLY.

* `cell_list` — the uniform-grid binning: floor((x - o)/cell) clamped to
  the grid, stable argsort of the flat cell key (ascending body index within
  a cell), exclusive prefix of the counts.
* `near_field` — brute force over all source/target pairs of
  G m d gS(r) / (r^2 + eps2)^(3/2) for r < rcut with the Gaussian split
  g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2) tapered to zero over
  [rcut - rs, rcut] by S(u) = 1 - u^3 (10 - 15u + 6u^2).
* `tsc_deposit` / `tsc_gather` — triangular-shaped-cloud assignment on an
  isolated 3-D mesh (3 nodes per axis around the nearest node).
* `split_bins` — the masses binned by TSC stencil centre, nodes grouped in
  blocks (the near-field cells), stable order.
Pinned in tests/test_p3m_oracle.py: the cell list against a brute-force
binning; g -> 1 as rs -> infinity turns the near field into the plain
all-pairs sum (checked against the long-double oracle); the taper is 1
below rcut - rs, 0 at rcut, with continuous first and second derivatives
(finite differences); 1 - g(r) equals
the mass fraction of a Gaussian cloud (variance 2 rs^2) inside r, computed
by quadrature; TSC conserves mass, puts 3/4 per axis on the node a body
sits on, interpolates linear fields exactly and is the adjoint of its
gather; split_bins matches a brute-force binning (nearest node by search,
lexicographic sort of (block, node in block, index)), and summing every
node's 27 neighbouring bins with the TSC weights gives tsc_deposit.
"""
from __future__ import annotations

import math

import numpy as np


def cell_list(pos, origin, cell, nc):
    keys = np.zeros(len(pos[0]), np.int64)
    for c in range(3):
        k = np.floor((np.asarray(pos[c]) - origin[c]) * (1.0 / cell)).astype(np.int64)
        k = np.clip(k, 0, nc[c] - 1)
        keys = keys * nc[c] + k
    order = np.argsort(keys, kind="stable")
    counts = np.bincount(keys, minlength=int(np.prod(nc)))
    start = np.concatenate([[0], np.cumsum(counts)])
    return start, order


def split_bins(pos, n, origin, h, block, nodes=True):
    """The split force's mass binning (ssim_split_bin, a binning of this
    build; its inputs are the reference's body arrays): per axis the TSC
    stencil centre r = floor((x - o)/h + 1/2) clamped to [0, n - 1]; nodes
    grouped in blocks of block^3 (nb = ceil(n / block) per axis), key
    ((r0//B nb + r1//B) nb + r2//B) B^3 + ((r0%B) B + r1%B) B + r2%B; stable
    order; bin starts over all nb^3 B^3 keys; block starts.  nodes=False:
    the key is the block index alone (bin starts None)."""
    nb = -(-n // block)
    r = [np.clip(np.floor((np.asarray(pos[c]) - origin[c]) / h + 0.5), 0,
                 n - 1).astype(np.int64) for c in range(3)]
    blk = [rc // block for rc in r]
    loc = [rc % block for rc in r]
    b3 = block ** 3
    if not nodes:
        cell = (blk[0] * nb + blk[1]) * nb + blk[2]
        order = np.argsort(cell, kind="stable")
        cstart = np.concatenate([[0], np.cumsum(np.bincount(cell, minlength=nb ** 3))])
        return None, cstart, order
    key = (((blk[0] * nb + blk[1]) * nb + blk[2]) * b3 +
           (loc[0] * block + loc[1]) * block + loc[2])
    order = np.argsort(key, kind="stable")
    start = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=nb ** 3 * b3))])
    return start, start[::b3], order


def split_g(r, rs):
    x = r / (2.0 * rs)
    erfc = np.vectorize(math.erfc)
    return erfc(x) + r / (rs * math.sqrt(math.pi)) * np.exp(-x * x)


def taper(r, rs, rcut):
    u = np.clip((np.asarray(r, np.float64) - (rcut - rs)) / rs, 0.0, 1.0)
    return np.maximum(1.0 - u ** 3 * (10.0 - 15.0 * u + 6.0 * u * u), 0.0)


def split_gs(r, rs, rcut):
    """The tapered split function the near field sums."""
    return np.where(np.asarray(r) < rcut, split_g(r, rs) * taper(r, rs, rcut), 0.0)


def near_field(src, m, tgt, G, eps2, rs, rcut, block=2048):
    src = [np.asarray(c, np.float64) for c in src]
    out = np.zeros((3, len(tgt[0])))
    for b0 in range(0, len(tgt[0]), block):
        t = [np.asarray(c[b0:b0 + block], np.float64) for c in tgt]
        d = [src[c][None, :] - t[c][:, None] for c in range(3)]
        r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
        r = np.sqrt(r2)
        w = np.where(r2 < rcut * rcut,
                     m[None, :] * split_gs(r, rs, rcut) / (r2 + eps2) ** 1.5, 0.0)
        for c in range(3):
            out[c, b0:b0 + block] = G * (w * d[c]).sum(1)
    return out


def tsc_weights(f):
    """Nearest node c = floor(f + 1/2) and the weights of c-1, c, c+1."""
    c = np.floor(f + 0.5)
    d = f - c
    return c.astype(np.int64), np.stack([0.5 * (0.5 - d) ** 2, 0.75 - d * d,
                                         0.5 * (0.5 + d) ** 2])


def _tsc_stencil(pos, n, origin, h):
    cw = [tsc_weights((np.asarray(pos[a], np.float64) - origin[a]) / h)
          for a in range(3)]
    for a in range(3):
        for b in range(3):
            for e in range(3):
                idx = [cw[0][0] + a - 1, cw[1][0] + b - 1, cw[2][0] + e - 1]
                ok = np.ones(len(idx[0]), bool)
                for q in range(3):
                    ok &= (idx[q] >= 0) & (idx[q] < n)
                w = cw[0][1][a] * cw[1][1][b] * cw[2][1][e]
                yield idx, ok, w


def tsc_deposit(pos, m, n, origin, h):
    rho = np.zeros((n, n, n))
    m = np.asarray(m, np.float64)
    for idx, ok, w in _tsc_stencil(pos, n, origin, h):
        np.add.at(rho, tuple(i[ok] for i in idx), (m * w)[ok])
    return rho


def tsc_gather(field, pos, n, origin, h):
    out = np.zeros(len(pos[0]))
    for idx, ok, w in _tsc_stencil(pos, n, origin, h):
        out[ok] += w[ok] * field[tuple(i[ok] for i in idx)]
    return out
