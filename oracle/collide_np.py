"""NumPy restatement of the uniform-grid collision search — TEST
INFRASTRUCTURE ONLY.

`hits` follows `Erkennungskarte._hits` (film_producer.py:223-317) step by
step on host arrays: alive mask with FP32 stripe compares, cell keys
`_zellschluessel` (320-321) in int64, a stable argsort of the keys (the
build pins CuPy's tie order as ascending row), searchsorted ranges per
neighbour offset (NACHBARN, 155), the `keep` rule, the FP32 prefilter with
one rounding per operation (numpy float32 arithmetic), ownership and the
FP64 recheck; the result is the pair list in the reference's order and the
candidate count.  `brute_force` is an independent O(n^2) restatement of the
same reach rule and predicate used to pin `hits`.
"""
from __future__ import annotations

import numpy as np

NACHBARN = ((0, 0), (1, 0), (0, 1), (1, 1), (1, -1))
F32_TOL = np.float32(1e-6)


def cell_key(cx, cy):
    with np.errstate(over="ignore"):
        return (np.asarray(cx, np.int64) * np.int64(73856093)) ^ \
               (np.asarray(cy, np.int64) * np.int64(19349663))


def _prefilter(gx, gy, gvx, gvy, rr, a, b, dt):
    dpx = gx[b] - gx[a]
    dpy = gy[b] - gy[a]
    dvx = gvx[b] - gvx[a]
    dvy = gvy[b] - gvy[a]
    dv2 = dvx * dvx + dvy * dvy
    t = -(dpx * dvx + dpy * dvy) / np.where(dv2 > 0, dv2, np.float32(1.0))
    t = np.clip(t, np.float32(-dt), np.float32(dt))
    cx = dpx + dvx * t
    cy = dpy + dvy * t
    rs = rr[a] + rr[b] + F32_TOL
    return cx * cx + cy * cy <= rs * rs


def _recheck(gx, gy, gvx, gvy, rr, a, b, dt):
    f = np.float64
    dpx = gx[b].astype(f) - gx[a].astype(f)
    dpy = gy[b].astype(f) - gy[a].astype(f)
    dvx = gvx[b].astype(f) - gvx[a].astype(f)
    dvy = gvy[b].astype(f) - gvy[a].astype(f)
    dv2 = dvx * dvx + dvy * dvy
    t = np.clip(-(dpx * dvx + dpy * dvy) / np.where(dv2 > 0, dv2, 1.0), -dt, dt)
    cx = dpx + dvx * t
    cy = dpy + dvy * t
    rs = rr[a].astype(f) + rr[b].astype(f)
    return cx * cx + cy * cy <= rs * rs


def hits(gx, gy, gvx, gvy, rr, is_ast, vis, dt_y, h, halo,
         lo=-np.inf, hi=np.inf):
    """((i, j) int64 arrays | None, candidate count)."""
    gx, gy, gvx, gvy, rr = (np.asarray(a, np.float32) for a in (gx, gy, gvx, gvy, rr))
    lo32, hi32 = np.float32(lo), np.float32(hi)
    alive = (np.asarray(is_ast) != 0) & (np.asarray(vis) != 0) & \
        (gx >= lo32) & (gx < hi32 + np.float32(halo))
    ai = np.flatnonzero(alive)
    if ai.size < 2:
        return None, 0
    ix = np.floor(gx[ai].astype(np.float64) / h).astype(np.int64)
    iy = np.floor(gy[ai].astype(np.float64) / h).astype(np.int64)
    key = cell_key(ix, iy)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    pa, pb, cand = [], [], 0
    for ox, oy in NACHBARN:
        k2 = cell_key(ix + ox, iy + oy)
        lo_r = np.searchsorted(ks, k2, side="left")
        hi_r = np.searchsorted(ks, k2, side="right")
        lens = hi_r - lo_r
        tot = int(lens.sum())
        cand += tot
        if tot == 0:
            continue
        rows = np.repeat(np.arange(ai.size), lens)
        starts = np.cumsum(lens) - lens
        cols = np.arange(tot) - starts[rows] + lo_r[rows]
        a = ai[rows]
        b = ai[order[cols]]
        keep = (a < b) if (ox == 0 and oy == 0) else (a != b)
        a, b = a[keep], b[keep]
        sel = _prefilter(gx, gy, gvx, gvy, rr, a, b, dt_y)
        pa.append(a[sel])
        pb.append(b[sel])
    if not pa or sum(len(p) for p in pa) == 0:
        return None, cand
    ci, cj = np.concatenate(pa), np.concatenate(pb)
    own = np.minimum(gx[ci], gx[cj]) < hi32
    ci, cj = ci[own], cj[own]
    if ci.size == 0:
        return None, cand
    sel = _recheck(gx, gy, gvx, gvy, rr, ci, cj, dt_y)
    if not sel.any():
        return None, cand
    return (ci[sel].astype(np.int64), cj[sel].astype(np.int64)), cand


def brute_force(gx, gy, gvx, gvy, rr, is_ast, vis, dt_y, h):
    """Set of pairs (a, b) over ALL alive pairs whose keys match through one
    of the neighbour offsets and that pass both predicates (no stripes)."""
    gx, gy, gvx, gvy, rr = (np.asarray(a, np.float32) for a in (gx, gy, gvx, gvy, rr))
    ai = np.flatnonzero((np.asarray(is_ast) != 0) & (np.asarray(vis) != 0))
    ix = np.floor(gx[ai].astype(np.float64) / h).astype(np.int64)
    iy = np.floor(gy[ai].astype(np.float64) / h).astype(np.int64)
    key = cell_key(ix, iy)
    out = set()
    for ox, oy in NACHBARN:
        k2 = cell_key(ix + ox, iy + oy)
        ra, rb = np.nonzero(k2[:, None] == key[None, :])
        a, b = ai[ra], ai[rb]
        keep = (a < b) if (ox == 0 and oy == 0) else (a != b)
        a, b = a[keep], b[keep]
        sel = _prefilter(gx, gy, gvx, gvy, rr, a, b, dt_y)
        a, b = a[sel], b[sel]
        sel = _recheck(gx, gy, gvx, gvy, rr, a, b, dt_y)
        out |= set(zip(a[sel].tolist(), b[sel].tolist()))
    return out
