"""NumPy oracle of the cell list and the near-field (split) force — TEST
INFRASTRUCTURE ONLY.

* `cell_list` — the uniform-grid binning: floor((x - o)/cell) clamped to
  the grid, stable argsort of the flat cell key (ascending body index within
  a cell), exclusive prefix of the counts.
* `near_field` — brute force over all source/target pairs of
  G m d g(r) / (r^2 + eps2)^(3/2) for r < rcut with the Gaussian split
  g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2).
Pinned in tests/test_p3m_oracle.py: g -> 1 as rs -> infinity turns the
near field into the plain all-pairs sum (checked against the long-double
oracle), and mesh part + near part of a pair reproduce the Plummer force.
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


def split_g(r, rs):
    x = r / (2.0 * rs)
    erfc = np.vectorize(math.erfc)
    return erfc(x) + r / (rs * math.sqrt(math.pi)) * np.exp(-x * x)


def near_field(src, m, tgt, G, eps2, rs, rcut, block=2048):
    src = [np.asarray(c, np.float64) for c in src]
    out = np.zeros((3, len(tgt[0])))
    for b0 in range(0, len(tgt[0]), block):
        t = [np.asarray(c[b0:b0 + block], np.float64) for c in tgt]
        d = [src[c][None, :] - t[c][:, None] for c in range(3)]
        r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
        r = np.sqrt(r2)
        w = np.where(r2 < rcut * rcut,
                     m[None, :] * split_g(r, rs) / (r2 + eps2) ** 1.5, 0.0)
        for c in range(3):
            out[c, b0:b0 + block] = G * (w * d[c]).sum(1)
    return out
