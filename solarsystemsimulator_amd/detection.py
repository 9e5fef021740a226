"""Asteroid-asteroid collision search on one MI355X: drop-in for the
detection part of the reference's film_producer.py (223-479).

`Erkennungskarte` keeps the reference's methods (`stammdaten`, `hochladen`,
`streifen`, `bounce_hits`) and returns the same pair list — bit-exact, in
the same order — from libssim's uniform-grid search (`ssim_collide_hits`).
`bounce_suche` picks the cell size from the 95th percentile of the alive
speeds like the reference and searches the stripes one after the other on
the single GPU (their union is the same list as one stripe covering
everything).  `_streifengrenzen` and `_lastanteile` are the reference's
stripe-boundary and load-share helpers.  Device arrays are torch tensors.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import device as D

NACHBARN = ((0, 0), (1, 0), (0, 1), (1, 1), (1, -1))
CHUNK = 4_000_000          # kept for API compatibility (no chunking needed)
LAST_ALPHA = 0.35
LAST_MIN_ANTEIL = 0.05


def _torch():
    import torch
    return torch


def _zellschluessel(cx, cy):
    """Cell key cx*73856093 XOR cy*19349663 in int64 (film_producer.py:320)."""
    with np.errstate(over="ignore"):
        return (np.asarray(cx, np.int64) * np.int64(73856093)) ^ \
               (np.asarray(cy, np.int64) * np.int64(19349663))


class Erkennungskarte:
    """One detection device with its resident flags and radii
    (film_producer.py:166-317)."""

    def __init__(self, dev: int, is_ast, lo: float, hi: float):
        torch = _torch()
        self.dev = dev
        self._dev = D.require_gpu(dev)
        self.lo = np.float32(lo)
        self.hi = np.float32(hi)
        self._ast_h = np.asarray(is_ast).astype(np.uint8)
        self.g_ast = D.to_dev(self._ast_h, self._dev, torch.uint8)
        self._vis_h = None
        self.g_vis = None
        self.g_rr = None
        self._cap = 1 << 16
        self._work = None
        self._n = 0

    def stammdaten(self, vis, real_r) -> None:
        torch = _torch()
        self._vis_h = np.asarray(vis).astype(np.uint8)
        self.g_vis = D.to_dev(self._vis_h, self._dev, torch.uint8)
        self.g_rr = D.to_dev(np.asarray(real_r, np.float32), self._dev, torch.float32)

    def hochladen(self, outs):
        """The whole f32 batch to the device in one transfer."""
        return D.to_dev(np.asarray(outs, np.float32), self._dev, _torch().float32)

    def streifen(self, lo: float, hi: float) -> None:
        self.lo = np.float32(lo)
        self.hi = np.float32(hi)

    def bounce_hits(self, gx, gy, gvx, gvy, dt_y: float, h: float, halo: float):
        """((i, j) host int64 | None, candidate pairs); mutates nothing."""
        return self._hits(gx, gy, gvx, gvy, dt_y, h, halo)

    def _hits(self, gx, gy, gvx, gvy, dt_y, h, halo):
        torch = _torch()
        n = gx.numel()
        lib = _lib.load()
        counts = (_lib.ctypes.c_longlong * 3)()
        # contiguous copies (if any) must outlive the launch: keep them here
        cols = [a.contiguous() for a in (gx, gy, gvx, gvy)]
        while True:
            wb = lib.ssim_collide_work_bytes(n, self._cap)
            if self._work is None or self._work.numel() < wb:
                self._work = D.workspace(wb, self._dev)
            oi = torch.empty(self._cap, dtype=torch.int64, device=self._dev)
            oj = torch.empty(self._cap, dtype=torch.int64, device=self._dev)
            _lib.call("ssim_collide_hits", n, *[_lib.ptr(a) for a in cols],
                      _lib.ptr(self.g_rr), _lib.ptr(self.g_ast),
                      _lib.ptr(self.g_vis), float(self.lo), float(self.hi),
                      float(halo), float(dt_y), float(h), self._cap, self._cap,
                      _lib.ptr(oi), _lib.ptr(oj), _lib.ctypes.addressof(counts),
                      _lib.ptr(self._work), self._work.numel(), D.stream(self._dev))
            if counts[2] >= 0:
                break
            self._cap = int(counts[1]) * 2        # more FP32 hits than room
        cand, npairs = int(counts[0]), int(counts[2])
        if npairs == 0:
            return None, cand
        return (oi[:npairs].cpu().numpy(), oj[:npairs].cpu().numpy()), cand


def merge_scan(dev, gx, gy, gvx, gvy, px, py, g_rr, g_vis, g_ast, m_alive,
               dt_years, barycentre=None, two_gm=0.0):
    """Device part of `analyze_merge` (film_producer.py:707-779).
    Returns ([(row, j, frac)], runaway indices | None); rows index
    `m_alive`.  barycentre=(bx, by) enables the runaway guard."""
    torch = _torch()
    n = gx.numel()
    nm = len(m_alive)
    lib = _lib.load()
    wb = lib.ssim_merge_work_bytes(n, nm)
    work = D.workspace(wb, dev)
    midx = D.to_dev(np.asarray(m_alive, np.int32), dev, torch.int32) if nm else None
    counts = (_lib.ctypes.c_longlong * 2)()
    cap = 4096
    while True:
        rows = torch.empty(cap, dtype=torch.int32, device=dev)
        cols = torch.empty(cap, dtype=torch.int32, device=dev)
        fr = torch.empty(cap, dtype=torch.float32, device=dev)
        run = torch.empty(cap, dtype=torch.int64, device=dev)
        bx, by = barycentre if barycentre is not None else (0.0, 0.0)
        _lib.call("ssim_merge_scan", n, *[_lib.ptr(a) for a in
                                          (gx, gy, gvx, gvy, px, py)],
                  _lib.ptr(g_rr), _lib.ptr(g_vis), _lib.ptr(g_ast), nm,
                  _lib.ptr(midx), float(np.float32(dt_years)),
                  int(barycentre is not None), float(bx), float(by),
                  float(np.float32(two_gm)), cap, _lib.ptr(rows), _lib.ptr(cols),
                  _lib.ptr(fr), cap, _lib.ptr(run),
                  _lib.ctypes.addressof(counts), _lib.ptr(work), wb,
                  D.stream(dev))
        if max(counts[0], counts[1]) <= cap:
            break
        cap = int(max(counts[0], counts[1]))
    nh, nr = int(counts[0]), int(counts[1])
    pairs = list(zip(rows[:nh].cpu().numpy().tolist(), cols[:nh].cpu().numpy().tolist(),
                     fr[:nh].cpu().numpy().tolist()))
    runaway = run[:nr].cpu().numpy() if nr else None
    return pairs, runaway


def bounce_deltas(dev, sample, hits, real_r, mass, dt_y, restitution=0.6):
    """Bounce response of a contact list on the device (the host function
    `bounce_deltas`, film_producer.py:916-1019): (hi, hj, affected,
    deltas[k, 4]) or None.  `sample` is the f32 (4n) [x|y|vx|vy] row, on
    the host or as a device tensor; `real_r` and `mass` (f64, n) likewise."""
    r = bounce_deltas_device(dev, sample, hits, real_r, mass, dt_y, restitution)
    if r is None:
        return None
    hi, hj, aff, d = r
    deltas = torch_stack_cols(d, aff).cpu().numpy()
    return hi, hj, aff.cpu().numpy().astype(np.int64), deltas


def torch_stack_cols(cols, idx):
    """(len(idx), len(cols)) from device columns gathered at idx."""
    torch = _torch()
    return torch.stack([c.index_select(0, idx) for c in cols], dim=1)


def bounce_deltas_device(dev, sample, hits, real_r, mass, dt_y, restitution=0.6):
    """bounce_deltas with the results left on the device: (hi, hj) host
    int64 of the pairs that act, the affected bodies as a device int64
    tensor and the four f64 delta columns (length n, by original index) —
    for add_body_deltas.  None when no pair acts."""
    torch = _torch()
    n = len(real_r)
    hi_h = np.asarray(hits[0], np.int64)
    hj_h = np.asarray(hits[1], np.int64)
    npairs = len(hi_h)
    if npairs == 0:
        return None
    if torch.is_tensor(sample):
        s = sample.reshape(-1)
        if s.dtype != torch.float32 or s.numel() != 4 * n:
            raise ValueError(f"sample muss float32 (4n = {4 * n}) sein")
        x, y, vx, vy = (s[k * n:(k + 1) * n] for k in range(4))
    else:
        s = np.asarray(sample, np.float32)
        x, y, vx, vy = (D.to_dev(s[k * n:(k + 1) * n], dev, torch.float32)
                        for k in range(4))
    rr = real_r if torch.is_tensor(real_r) else D.to_dev(np.asarray(real_r, np.float64), dev)
    ms = mass if torch.is_tensor(mass) else D.to_dev(np.asarray(mass, np.float64), dev)
    hi, hj = D.to_dev(hi_h, dev, torch.int64), D.to_dev(hj_h, dev, torch.int64)
    d = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4)]
    act = torch.empty(npairs, dtype=torch.int64, device=dev)
    aff = torch.empty(n, dtype=torch.int64, device=dev)
    wb = _lib.load().ssim_bounce_work_bytes(n, npairs)
    work = D.workspace(wb, dev)
    counts = (_lib.ctypes.c_longlong * 2)()
    _lib.call("ssim_bounce_deltas", n, npairs, _lib.ptr(hi), _lib.ptr(hj),
              _lib.ptr(x), _lib.ptr(y), _lib.ptr(vx), _lib.ptr(vy), _lib.ptr(rr),
              _lib.ptr(ms), float(dt_y), float(restitution),
              *[_lib.ptr(c) for c in d], _lib.ptr(act), _lib.ptr(aff),
              _lib.ctypes.addressof(counts), _lib.ptr(work), wb, D.stream(dev))
    na, nf = int(counts[0]), int(counts[1])
    if na == 0:
        return None
    a = act[:na].cpu().numpy()
    return hi_h[a], hj_h[a], aff[:nf], d


def add_body_deltas(st, aff, d, dev) -> None:
    """Add bounce deltas (device columns d, length n by original index) to
    the NBodyCuda state `st` for the affected bodies `aff` (device int64,
    distinct) — apply_deltas without the host round trip."""
    if aff.numel() == 0:
        return
    _body_index(st, dev)
    ptrs = _body_ptrs(st)
    _lib.call("ssim_add_body_deltas", int(aff.numel()), _lib.ptr(aff),
              *[_lib.ptr(c) for c in d], _lib.ptr(st["inv_kind_d"]),
              _lib.ptr(st["inv_pos_d"]), _lib.ctypes.addressof(ptrs[0]),
              _lib.ctypes.addressof(ptrs[1]), D.stream(dev))


def _body_index(st, dev) -> None:
    torch = _torch()
    if "inv_kind_d" not in st:
        st["inv_kind_d"] = torch.from_numpy(st["inv_kind"].astype(np.int32)).to(dev)
        st["inv_pos_d"] = torch.from_numpy(st["inv_pos"].astype(np.int32)).to(dev)


def _body_ptrs(st):
    """The 7 device pointers (x, y, vx, vy, m, visible, radius) of the
    massive and of the asteroid arrays of a 2-D NBodyCuda state."""
    return [(_lib.ctypes.c_void_p * 7)(*[_lib.ptr(t) for t in st[k][:5]],
                                      _lib.ptr(st[v]), _lib.ptr(st[r]))
            for k, v, r in (("M", "mvis", "mrad"), ("a", "avis", "arad"))]


def apply_merges(st, sample, pairs, runaway, k_ev, mirrors, contact_t,
                 shatter_check, G, t0_days, raster_days):
    """The host closure `apply_merges` (film_producer.py:781-861) on the
    device: the merge pairs (list of (i, j, frac)) and runaway bodies applied
    to the NBodyCuda state `st` and to the device mirrors (dict of tensors
    "mass" f64, "vis" u8, "real_r" f64, "is_ast" u8, "is_star_bh" u8, all by
    original index; updated), in list order.  `contact_t` is an f64 device
    array of kernel contact times (NaN: none; consumed).  `sample` is the f32
    (4n) [x|y|vx|vy] row (host or device).  Returns (event records as bytes
    in the ring's "<dIIfIff" format, merges, shatter) with shatter =
    (i, j, t_days) or None; a shatter stops the call like the reference's
    `return` (the caller then freezes and dumps the state)."""
    torch = _torch()
    dev = mirrors["mass"].device
    n = int(mirrors["mass"].numel())
    smp = sample if torch.is_tensor(sample) else torch.from_numpy(
        np.ascontiguousarray(sample, np.float32)).to(dev)
    npairs, nrun = len(pairs), len(runaway)
    pi = torch.tensor([int(p[0]) for p in pairs] or [0], dtype=torch.int32, device=dev)
    pj = torch.tensor([int(p[1]) for p in pairs] or [0], dtype=torch.int32, device=dev)
    pf = torch.tensor([float(p[2]) for p in pairs] or [0.0], dtype=torch.float64,
                      device=dev)
    run = torch.tensor([int(j) for j in runaway] or [0], dtype=torch.int32, device=dev)
    _body_index(st, dev)
    ptrs = _body_ptrs(st)
    events = torch.empty(32 * max(npairs + nrun, 1), dtype=torch.uint8, device=dev)
    result = torch.zeros(4, dtype=torch.int32, device=dev)
    shatter_t = torch.zeros(1, dtype=torch.float64, device=dev)
    _lib.call("ssim_apply_merges", n, _lib.ptr(smp), npairs, _lib.ptr(pi), _lib.ptr(pj),
              _lib.ptr(pf), nrun, _lib.ptr(run), _lib.ptr(mirrors["mass"]),
              _lib.ptr(mirrors["vis"]), _lib.ptr(mirrors["real_r"]),
              _lib.ptr(mirrors["is_ast"]), _lib.ptr(mirrors["is_star_bh"]),
              _lib.ptr(contact_t), int(k_ev), int(bool(shatter_check)), float(G),
              float(t0_days), float(raster_days), _lib.ptr(st["inv_kind_d"]),
              _lib.ptr(st["inv_pos_d"]), _lib.ctypes.addressof(ptrs[0]),
              _lib.ctypes.addressof(ptrs[1]), _lib.ptr(events), _lib.ptr(result),
              _lib.ptr(shatter_t), D.stream(dev))
    res = result.cpu().numpy()
    nev = int(res[0])
    shatter = None
    if res[2] >= 0:
        shatter = (int(res[2]), int(res[3]), float(shatter_t.cpu().item()))
    return bytes(events[:32 * nev].cpu().numpy()), int(res[1]), shatter


def _streifengrenzen(gx_lebt, anteile):
    """x stripes holding `anteile` of the alive asteroids (outer bounds
    infinite), film_producer.py:382-405."""
    if len(anteile) < 2:
        return [(-np.inf, np.inf)]
    cuts = 100.0 * np.cumsum(anteile[:-1])
    x = gx_lebt.cpu().numpy() if hasattr(gx_lebt, "cpu") else np.asarray(gx_lebt)
    edges = [-np.inf, *[float(v) for v in np.percentile(x, cuts)], np.inf]
    return list(zip(edges[:-1], edges[1:]))


def _lastanteile(anteile, last):
    """Update the stripe shares from the measured candidate counts,
    damped multiplicatively, with a 5 % floor (film_producer.py:418-479)."""
    k = len(anteile)
    if k < 2:
        return list(anteile)
    total = float(sum(last))
    if total <= 0 or min(last) < 0:
        return list(anteile)
    target = 1.0 / k
    new = [max(a * (target / max(c / total, 1e-9)) ** LAST_ALPHA, 1e-6)
           for a, c in zip(anteile, last)]
    s = sum(new)
    new = [a / s for a in new]
    low = [a < LAST_MIN_ANTEIL for a in new]
    if any(low):
        room = 1.0 - LAST_MIN_ANTEIL * sum(low)
        rest = sum(a for a, lo in zip(new, low) if not lo)
        if room <= 0 or rest <= 0:
            return [1.0 / k] * k
        new = [LAST_MIN_ANTEIL if lo else a / rest * room for a, lo in zip(new, low)]
    return new


def speed_percentile(karte: Erkennungskarte, gvx, gvy, q: float) -> float:
    """The q-th percentile of the alive bodies' speeds, on the device
    (ssim_speed_percentile: NumPy's linear rule on the FP32 speeds, like
    `cp.percentile(cp.hypot(gvx[lebt], gvy[lebt]), q)`, film_producer.py:357);
    one float comes back."""
    torch = _torch()
    n = gvx.numel()
    lib = _lib.load()
    wb = lib.ssim_speed_percentile_work_bytes(n)
    if getattr(karte, "_pwork", None) is None or karte._pwork.numel() < wb:
        karte._pwork = D.workspace(wb, karte._dev)
        karte._pout = torch.empty(1, dtype=torch.float32, device=karte._dev)
    cols = [a.contiguous() for a in (gvx, gvy)]
    _lib.call("ssim_speed_percentile", n, _lib.ptr(cols[0]), _lib.ptr(cols[1]),
              _lib.ptr(karte.g_ast), _lib.ptr(karte.g_vis), float(q), _lib.ptr(karte._pout),
              _lib.ptr(karte._pwork), karte._pwork.numel(), D.stream(karte._dev))
    return float(karte._pout.cpu()[0])


def bounce_suche(det, det_pool, rows, dt_y: float, rr_max: float, anteile=None):
    """Collision pairs of one sample over the stripes (film_producer.py:
    324-379): returns (pairs | None, candidates, h, new shares).  On one
    GPU every stripe is searched here in turn; `det_pool` is unused."""
    gx, gy, gvx, gvy = rows[0]
    if anteile is None or len(anteile) != len(rows):
        anteile = [1.0 / len(rows)] * len(rows)
    k0 = det[0]
    idx = np.flatnonzero((k0._ast_h != 0) & (k0._vis_h != 0))
    if idx.size < 2:
        return None, 0, 0.0, anteile
    v95 = speed_percentile(k0, gvx, gvy, 95.0)
    bounds = (_streifengrenzen(gx.cpu().numpy()[idx], anteile) if len(rows) > 1
              else [(-np.inf, np.inf)])
    h = max(1e-4, 2.0 * v95 * dt_y)
    halo = 2.0 * h + 2.0 * rr_max
    parts, per_card = [], []
    for karte, (lo, hi), row in zip(det * len(rows) if len(det) < len(rows) else det,
                                    bounds, rows):
        karte.streifen(lo, hi)
        hit, cand = karte.bounce_hits(*row, dt_y, h, halo)
        parts.append(hit)
        per_card.append(cand)
    new = _lastanteile(anteile, per_card)
    hits = [p for p in parts if p is not None]
    cand = sum(per_card)
    if not hits:
        return None, cand, h, new
    return ((np.concatenate([p[0] for p in hits]),
             np.concatenate([p[1] for p in hits])), cand, h, new)
