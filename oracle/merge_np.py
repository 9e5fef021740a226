"""NumPy restatement of the massive-body contact scan, the runaway guard and
the bounce response — TEST INFRASTRUCTURE ONLY.

* `merge_scan` — the device part of `analyze_merge` (film_producer.py:
  707-779): segment-sphere tests of every body against every alive massive
  body in FP32, row-major hit order, frac = tt - 1 (previous segment) or tt
  (extrapolated segment); the runaway guard v^2 > 9 v_esc^2 about the
  massive barycentre.
* `bounce_deltas` — the host response of `bounce_deltas` (film_producer.py:
  916-1019): greedy one-contact-per-body filter in list order, entry time of
  |dp + dv t| = touch, validity, approach test, restitution impulse and the
  mass-weighted 1.1x overlap push.
* `apply_merges` — the state changes of `apply_merges` (film_producer.py:
  781-861): pairs in list order against mirrors that earlier pairs change,
  the kernel contact time replacing frac inside one raster, the shatter
  stop at |dv| >= 1.5 v_esc, the mass-weighted merge into the heavier body
  with a volume-conserving radius, the runaway removals, and the events in
  the ring's "<dIIfIff" records (`emit_event`, film_producer.py:687-705).
Pinned in tests/test_merge_oracle.py: a head-on pair conserves momentum and
reverses the approach speed by e; a body crossing a star between two
samples reports frac in (-1, 0]; the greedy filter keeps the first pair;
merges conserve mass, momentum, the mass-weighted centre and r^3; a chain
(a, b), (b, c) merges only once; a shatter stops the call.
"""
from __future__ import annotations

import struct

import numpy as np

f32 = np.float32


def _seg_hit(p0x, p0y, p1x, p1y, cx, cy, rsum2):
    sx = (p1x - p0x)[None, :]
    sy = (p1y - p0y)[None, :]
    seg2 = sx * sx + sy * sy
    tt = ((cx - p0x[None, :]) * sx + (cy - p0y[None, :]) * sy) / \
        np.where(seg2 > 0, seg2, f32(1.0))
    tt = np.clip(tt, f32(0.0), f32(1.0))
    ddx = p0x[None, :] + tt * sx - cx
    ddy = p0y[None, :] + tt * sy - cy
    return ddx * ddx + ddy * ddy <= rsum2, tt


def merge_scan(gx, gy, gvx, gvy, px, py, rr, vis, m_alive, dt):
    gx, gy, gvx, gvy, px, py, rr = (np.asarray(a, f32) for a in
                                    (gx, gy, gvx, gvy, px, py, rr))
    gm = np.asarray(m_alive, np.int64)
    cx = gx[gm][:, None]
    cy = gy[gm][:, None]
    rsum = rr[gm][:, None] + rr[None, :]
    rsum2 = rsum * rsum
    alive = np.asarray(vis) != 0
    hit_a, tt_a = _seg_hit(px, py, gx, gy, cx, cy, rsum2)
    hit_b, tt_b = _seg_hit(gx, gy, gx + gvx * f32(dt), gy + gvy * f32(dt),
                           cx, cy, rsum2)
    hit = (hit_a | hit_b) & alive[None, :]
    frac = np.where(hit_a, tt_a - f32(1.0), tt_b)
    hit[np.arange(len(gm)), gm] = False
    rows, cols = np.nonzero(hit)
    return rows, cols, frac[rows, cols]


def runaway(gx, gy, gvx, gvy, is_ast, vis, bx, by, two_gm):
    gx, gy, gvx, gvy = (np.asarray(a, f32) for a in (gx, gy, gvx, gvy))
    r = np.maximum(np.hypot(gx - f32(bx), gy - f32(by)), f32(1e-6))
    v2 = gvx * gvx + gvy * gvy
    vesc2 = f32(two_gm) / r
    flag = (v2 > f32(9.0) * vesc2) & (np.asarray(is_ast) != 0) & (np.asarray(vis) != 0)
    return np.flatnonzero(flag)


def bounce_deltas(sample, hi, hj, real_r, mass, dt_y, restitution=0.6):
    """(hi, hj, affected, deltas[k, 4]) or None for the contact pairs in
    list order (x|y|vx|vy f32 sample of n bodies)."""
    n = len(real_r)
    x, y = sample[0:n], sample[n:2 * n]
    vx, vy = sample[2 * n:3 * n], sample[3 * n:4 * n]
    used = np.zeros(n, bool)
    keep = []
    for t, (a, b) in enumerate(zip(hi, hj)):
        if used[a] or used[b]:
            continue
        used[a] = used[b] = True
        keep.append(t)
    if not keep:
        return None
    hi = np.asarray(hi)[keep].astype(np.int64)
    hj = np.asarray(hj)[keep].astype(np.int64)
    f = np.float64
    dpx = x[hj].astype(f) - x[hi].astype(f)
    dpy = y[hj].astype(f) - y[hi].astype(f)
    dvx = vx[hj].astype(f) - vx[hi].astype(f)
    dvy = vy[hj].astype(f) - vy[hi].astype(f)
    dv2 = dvx * dvx + dvy * dvy
    touch = real_r[hi] + real_r[hj]
    c = dpx * dpx + dpy * dpy - touch * touch
    b = 2.0 * (dpx * dvx + dpy * dvy)
    disc = b * b - 4.0 * dv2 * c
    sq = np.sqrt(np.maximum(disc, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        t_in = (-b - sq) / (2.0 * np.where(dv2 > 0, dv2, 1.0))
    tc = np.where(c < 0.0, 0.0, np.clip(np.nan_to_num(t_in), -dt_y, dt_y))
    valid = (c < 0.0) | (disc >= -(touch * touch) * 0.01)
    ncx = dpx + dvx * tc
    ncy = dpy + dvy * tc
    dist = np.hypot(ncx, ncy)
    dist = np.where(dist > 1e-30, dist, 1.0)
    nx, ny = ncx / dist, ncy / dist
    vrel = dvx * nx + dvy * ny
    act = valid & (vrel < 0)
    if not act.any():
        return None
    hi, hj, nx, ny, vrel = hi[act], hj[act], nx[act], ny[act], vrel[act]
    mi, mj = mass[hi], mass[hj]
    with np.errstate(divide="ignore", invalid="ignore"):
        imp = np.nan_to_num(-(1.0 + restitution) * vrel / (1.0 / mi + 1.0 / mj))
    d = np.zeros((4, n))
    d[2][hi] -= imp * nx / mi
    d[3][hi] -= imp * ny / mi
    d[2][hj] += imp * nx / mj
    d[3][hj] += imp * ny / mj
    pdx = x[hj].astype(f) - x[hi].astype(f)
    pdy = y[hj].astype(f) - y[hi].astype(f)
    pdist = np.hypot(pdx, pdy)
    overlap = real_r[hi] + real_r[hj] - pdist
    ol = (overlap > 0) & (pdist > 1e-30)
    if ol.any():
        oi, oj = hi[ol], hj[ol]
        pnx, pny = pdx[ol] / pdist[ol], pdy[ol] / pdist[ol]
        mt = mass[oi] + mass[oj]
        shift = overlap[ol] * 1.1
        d[0][oi] -= pnx * shift * (mass[oj] / mt)
        d[1][oi] -= pny * shift * (mass[oj] / mt)
        d[0][oj] += pnx * shift * (mass[oi] / mt)
        d[1][oj] += pny * shift * (mass[oi] / mt)
    aff = np.unique(np.concatenate([hi, hj]))
    return hi, hj, aff.astype(np.int64), d[:, aff].T.copy()


def apply_merges(sample, pairs, runaway, k_ev, mass, vis, real_r, is_ast, is_star_bh,
                 contact, shatter_check, G, t0_days, raster_days):
    """The loop of apply_merges on host arrays (updated in place: mass, vis,
    real_r, contact (dict j -> kernel contact time, consumed)).  Returns
    (state updates {body: (x, y, vx, vy, m)}, deactivated bodies in order,
    radius updates {body: r}, event records (bytes), merges, shatter) with
    shatter = (i, j, t_days) or None."""
    n = len(mass)
    sx, sy = sample[0:n], sample[n:2 * n]
    svx, svy = sample[2 * n:3 * n], sample[3 * n:4 * n]
    body, dead, radius, events = {}, [], {}, []
    merges = 0

    def event(a, b, m, kind, x, y, frac=0.0):
        t = t0_days + (k_ev + 1 + frac) * raster_days
        events.append(struct.pack("<dIIfIff", t, a & 0xFFFFFFFF, b, m, kind,
                                  float(x), float(y)))

    for mi, j, frac in pairs:
        tk = contact.pop(j, None)
        if tk is not None:
            f_kern = tk - (k_ev + 1)
            if -1.0 <= f_kern <= 1.0:
                frac = f_kern
        if (shatter_check and not is_ast[mi] and not is_ast[j]
                and not is_star_bh[mi] and not is_star_bh[j]
                and vis[mi] and vis[j] and mass[mi] > 0 and mass[j] > 0):
            v_imp = float(np.hypot(svx[j] - svx[mi], svy[j] - svy[mi]))
            touch = max(1e-12, real_r[mi] + real_r[j])
            v_esc = float(np.sqrt(2.0 * G * (mass[mi] + mass[j]) / touch))
            if v_imp >= 1.5 * v_esc:
                return (body, dead, radius, events, merges,
                        (int(mi), int(j), t0_days + (k_ev + 1 + frac) * raster_days))
        if not vis[mi] or not vis[j] or mass[mi] <= 0:
            continue
        a, b = (mi, j) if mass[mi] >= mass[j] else (j, mi)
        m_a, m_b = mass[a], mass[b]
        m_ges = m_a + m_b
        if m_ges <= 0:
            continue
        nx = (sx[a] * m_a + sx[b] * m_b) / m_ges
        ny = (sy[a] * m_a + sy[b] * m_b) / m_ges
        nvx = (svx[a] * m_a + svx[b] * m_b) / m_ges
        nvy = (svy[a] * m_a + svy[b] * m_b) / m_ges
        body[a] = (nx, ny, nvx, nvy, m_ges)
        dead.append(b)
        mass[a] = m_ges
        mass[b] = 0.0
        vis[b] = 0
        real_r[a] = (real_r[a] ** 3 + real_r[b] ** 3) ** (1.0 / 3.0)
        radius[a] = real_r[a]
        event(a, b, float(m_ges), 0, nx, ny, frac)
        merges += 1
    for j in runaway:
        j = int(j)
        if not vis[j]:
            continue
        dead.append(j)
        mass[j] = 0.0
        vis[j] = 0
        event(0xFFFFFFFF, j, 0.0, 2, float(sx[j]), float(sy[j]))
    return body, dead, radius, events, merges, None
