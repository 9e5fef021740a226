"""Few massive bodies + many asteroids on one MI355X: drop-in for the
reference's `NBodyCuda` (nbody_kernel.py:617-1060).

`load_state` takes host NumPy arrays and keeps the state resident on the GPU;
`step_batch(st, dt_raster, k)` runs k raster samples in one cooperative
launch of libssim's raster integrator and returns f32 (k, 4n)
[x|y|vx|vy] in the original body order, with `st["sub"]` (sub-samples of
hot asteroids) and `st["hits"]` (contacts) as the reference fills them;
point updates (`apply_updates`, `apply_deltas`, `apply_body_state`,
`set_radius`, `deactivate_body`) scatter on the device; `export_f64`
returns the exact f64 state.  One GPU holds every body, so the reference's
shards, replicas and host barrier are gone; a device list is accepted for
API compatibility and its first entry is used.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as D
from .constants import (G_AU, G_MAX, K_MAX, M_MAX, MAX_SUB_DT_YEARS,
                        MAX_SUB_STEPS_PER_FRAME, SOFTENING, SUB_SAMPLES,
                        SUB_SAMPLES_MAX, YOSHIDA_W0, YOSHIDA_W1)


def _torch():
    import torch
    return torch


def pick_device() -> int:
    """The GPU to use (nbody_kernel.py:575-578); one MI355X here."""
    return 0


def pick_devices() -> list[int]:
    """Physics GPUs (nbody_kernel.py:581-592): device 0."""
    return [0]


def pick_detect_devices(exclude, rank: int = 0, count: int = 1) -> list[int]:
    """GPUs outside the physics devices for collision detection
    (nbody_kernel.py:595-614).  With one GPU there are none; the caller then
    runs detection on the physics GPU, as the reference does."""
    import torch
    excl = set(exclude) if isinstance(exclude, (list, tuple, set)) else {exclude}
    cands = [i for i in range(torch.cuda.device_count()) if i not in excl]
    if not cands:
        return []
    count = min(count, len(cands))
    return [cands[(rank + i) % len(cands)] for i in range(count)]


class NBodyCuda:
    """Drop-in for nbody_kernel.NBodyCuda."""

    dim = 2
    # physical constants of the reference's scheme (nbody_kernel.py:58-65)
    G = G_AU
    soft = SOFTENING
    max_sub_dt = MAX_SUB_DT_YEARS

    def __init__(self, devices, m_sub=SUB_SAMPLES):
        if isinstance(devices, (int, np.integer)):
            devices = [int(devices)]
        devices = list(devices)
        if not devices or len(devices) > G_MAX:
            raise ValueError(f"1..{G_MAX} Devices erwartet: {devices}")
        if not 0 <= m_sub <= SUB_SAMPLES_MAX:
            raise ValueError(f"m_sub ausserhalb 0..{SUB_SAMPLES_MAX}: {m_sub}")
        self.m_sub = int(m_sub)
        self.devices = devices
        self.device = devices[0]
        self._dev = D.require_gpu(self.device)

    def name(self) -> str:
        return _torch().cuda.get_device_name(self._dev)

    # ---------------- state ----------------
    def load_state(self, x, y, vx, vy, mass, visible, is_ast,
                   real_r=None) -> dict:
        return self._load([x, y], [vx, vy], mass, visible, is_ast, real_r)

    def _load(self, pos, vel, mass, visible, is_ast, real_r) -> dict:
        torch = _torch()
        F = 2 * self.dim
        ast = np.asarray(is_ast) != 0
        m_idx = np.flatnonzero(~ast)
        a_idx = np.flatnonzero(ast)
        m = len(m_idx)
        if m > M_MAX:
            raise ValueError(f"zu viele massive Koerper: {m} > {M_MAX}")
        n = len(pos[0])
        na = len(a_idx)
        dev = self._dev
        cols = [np.asarray(c, np.float64) for c in list(pos) + list(vel) + [mass]]
        vis = np.asarray(visible).astype(np.uint8)
        rr = (np.zeros(n) if real_r is None else np.asarray(real_r, np.float64))

        def up(arr, dtype=torch.float64):
            return D.to_dev(arr, dev, dtype) if len(arr) else \
                torch.zeros(1, dtype=dtype, device=dev)
        st = {"N": n, "m": m, "n_ast": na, "m_idx_h": m_idx, "a_idx_h": a_idx,
              "dim": self.dim,
              "a": [up(c[a_idx]) for c in cols],          # positions velocities m
              "M": [up(c[m_idx]) for c in cols],
              "avis": up(vis[a_idx], torch.uint8),
              "mvis": up(vis[m_idx], torch.uint8),
              "arad": up(rr[a_idx]), "mrad": up(rr[m_idx]),
              "aacc": [torch.zeros(max(na, 1), dtype=torch.float64, device=dev)
                       for _ in range(self.dim)],
              "macc": [torch.zeros(max(m, 1), dtype=torch.float64, device=dev)
                       for _ in range(self.dim)]}
        ms = max(self.m_sub, 1)
        st["snap"] = torch.empty(K_MAX * F * max(na, 1), dtype=torch.float32, device=dev)
        st["snapM"] = torch.empty(K_MAX * F * max(m, 1), dtype=torch.float32, device=dev)
        st["subpos"] = torch.zeros(K_MAX * ms * self.dim * max(na, 1),
                                   dtype=torch.float32, device=dev)
        st["subn"] = torch.zeros(K_MAX * max(na, 1), dtype=torch.uint8, device=dev)
        st["hitt"] = torch.full((max(na, 1),), -1.0, dtype=torch.float32, device=dev)
        st["hitm"] = torch.zeros(max(na, 1), dtype=torch.int32, device=dev)
        wb = _lib.load().ssim_nbody_work_bytes(na)
        st["work"] = D.workspace(wb, dev)
        # inverse map original index -> (-1 massive / 0 asteroid, position)
        inv_kind = np.zeros(n, np.int8)
        inv_kind[m_idx] = -1
        inv_pos = np.zeros(n, np.int64)
        inv_pos[a_idx] = np.arange(na)
        inv_pos[m_idx] = np.arange(m)
        st["inv_kind"], st["inv_pos"] = inv_kind, inv_pos
        return st

    # ---------------- batch ----------------
    def step_batch(self, st: dict, dt_raster_years: float, k: int) -> np.ndarray:
        if st is None:
            raise ValueError("kein Zustand geladen — FULL-Frame noetig")
        if not 1 <= k <= K_MAX:
            raise ValueError(f"k ausserhalb 1..{K_MAX}: {k}")
        self._run_frames(st, dt_raster_years, k)
        na, m, n = st["n_ast"], st["m"], st["N"]
        F = 2 * self.dim
        out = np.empty((k, F * n), np.float32)
        snap = st["snap"][:k * F * na].cpu().numpy().reshape(k, F, na)
        snapm = st["snapM"][:k * F * m].cpu().numpy().reshape(k, F, m)
        for f in range(F):
            out[:, f * n + st["a_idx_h"]] = snap[:, f, :]
            out[:, f * n + st["m_idx_h"]] = snapm[:, f, :]
        st["sub"] = self._collect_sub(st, k)
        st["hits"] = self._collect_hits(st)
        return out

    def step_batch_slots(self, st: dict, dt_raster_years: float, k: int,
                         sub_max: int, out: np.ndarray | None = None):
        """k raster samples straight into film ring slots: the slot layout of
        film_producer.py:81-123 (`slot_bytes`, `schreibe_slot` applied to the
        batch's rows and complete tracks) is written on the device and
        copied to the host in one transfer, instead of the host scatter to
        original order plus one argsort and transpose per sample.

        Returns (slots uint8 (k, slot_bytes), overflow bool (k,)), overflow
        where more complete tracks than sub_max fell due (the lowest original
        indices are kept, as in the reference).  With `out` (uint8
        (k, slot_bytes), e.g. k consecutive slots of the shared ring) the
        slots land there and `out` is returned.  Bytes of the index and
        track fields past the slot's count are unspecified.  st["hits"] is
        collected as after step_batch; st["sub"] is None (the tracks are in
        the slots)."""
        sb = self._slots_check(st, k, sub_max, out)
        self._run_frames(st, dt_raster_years, k)
        self._slots_enqueue(st, k, sub_max)
        return self._slots_collect(st, k, sb, sub_max, out)

    def _slots_check(self, st, k, sub_max, out) -> int:
        if st is None:
            raise ValueError("kein Zustand geladen — FULL-Frame noetig")
        if self.dim != 2:
            raise ValueError("film ring slots carry the plane (x, y) only")
        if not 1 <= k <= K_MAX:
            raise ValueError(f"k ausserhalb 1..{K_MAX}: {k}")
        if sub_max < 0:
            raise ValueError(f"sub_max < 0: {sub_max}")
        sb = int(_lib.load().ssim_ring_slot_bytes(st["N"], self.m_sub, int(sub_max)))
        if out is not None and (out.shape != (k, sb) or out.dtype != np.uint8):
            raise ValueError(f"out muss uint8 ({k}, {sb}) sein, ist {out.dtype} {out.shape}")
        return sb

    def _idx_dev(self, st) -> None:
        """Original indices of the integrator's slots as u32 device arrays
        (ascending for the asteroids, which the slot order relies on)."""
        torch = _torch()
        if "a_idx_d" not in st:
            for key, h in (("a_idx_d", st["a_idx_h"]), ("m_idx_d", st["m_idx_h"])):
                h = np.asarray(h, np.int64)
                st[key] = torch.as_tensor(h.astype(np.uint32).view(np.int32)).to(self._dev)
            assert np.all(np.diff(st["a_idx_h"]) > 0)      # slot order = index order

    def _slots_enqueue(self, st, k, sub_max) -> None:
        """The ring-slot kernels of the last k frames and their copy to
        pinned host memory, queued on the current stream."""
        torch = _torch()
        na, m, n = st["n_ast"], st["m"], st["N"]
        lib = _lib.load()
        sb = int(lib.ssim_ring_slot_bytes(n, self.m_sub, int(sub_max)))
        dev = self._dev
        self._idx_dev(st)
        if st.get("ring_dev") is None or st["ring_dev"].numel() < k * sb:
            st["ring_dev"] = torch.empty(k * sb, dtype=torch.uint8, device=dev)
            st["ring_host"] = torch.empty(k * sb, dtype=torch.uint8, pin_memory=True)
            st["ring_tot"] = torch.empty(K_MAX, dtype=torch.int32, device=dev)
            st["ring_tot_h"] = torch.empty(K_MAX, dtype=torch.int32, pin_memory=True)
        wb = lib.ssim_ring_work_bytes(k, na)
        if st.get("ring_work") is None or st["ring_work"].numel() < wb:
            st["ring_work"] = D.workspace(wb, dev)
        s = D.stream(dev)
        _lib.call("ssim_ring_slots", k, n, na, m, self.m_sub, int(sub_max),
                  _lib.ptr(st["snap"]), _lib.ptr(st["snapM"]), _lib.ptr(st["subpos"]),
                  _lib.ptr(st["subn"]), _lib.ptr(st["a_idx_d"]), _lib.ptr(st["m_idx_d"]),
                  _lib.ptr(st["ring_dev"]), _lib.ptr(st["ring_tot"]),
                  _lib.ptr(st["ring_work"]), st["ring_work"].numel(), s)
        st["ring_host"][:k * sb].copy_(st["ring_dev"][:k * sb], non_blocking=True)
        st["ring_tot_h"][:k].copy_(st["ring_tot"][:k], non_blocking=True)

    def _slots_collect(self, st, k, sb, sub_max, out):
        """Wait for the queued slots and hand them out (see step_batch_slots)."""
        torch = _torch()
        torch.cuda.current_stream(self._dev).synchronize()
        host, tot = st["ring_host"][:k * sb], st["ring_tot_h"][:k]
        slots = host.numpy().reshape(k, sb)
        if out is not None:
            out[...] = slots
        else:
            out = slots.copy()
        st["sub"] = None
        st["hits"] = self._collect_hits(st)
        # no index field at all (sub_max 0): nothing overflows, as in
        # schreibe_slot's `not sub_max` branch (film_producer.py:107-109)
        return out, (tot.numpy() > sub_max) & (sub_max > 0)

    def frame_rows(self, st: dict, k: int, rows) -> None:
        """After a batch: its k samples as device rows f32 (k, 4n)
        [x|y|vx|vy] in original order (libssim's ssim_frame_rows), the
        array step_batch returns, without leaving the device.  `rows` is a
        float32 device tensor of at least k * 4n elements."""
        if self.dim != 2:
            raise ValueError("frame rows carry the plane (x, y) only")
        n, na, m = st["N"], st["n_ast"], st["m"]
        if rows.numel() < k * 4 * n or rows.dtype != _torch().float32:
            raise ValueError(f"rows muss float32 mit >= {k * 4 * n} Elementen sein")
        self._idx_dev(st)
        _lib.call("ssim_frame_rows", int(k), n, na, m, _lib.ptr(st["snap"]),
                  _lib.ptr(st["snapM"]), _lib.ptr(st["a_idx_d"]), _lib.ptr(st["m_idx_d"]),
                  _lib.ptr(rows), D.stream(self._dev))

    def _run_frames(self, st: dict, dt_raster_years: float, k: int) -> None:
        na, m = st["n_ast"], st["m"]
        F = 2 * self.dim
        s = D.stream(self._dev)
        _lib.call("ssim_fill_f32", max(na, 1), _lib.ptr(st["hitt"]), -1.0, s)
        a, M_ = st["a"], st["M"]
        entry = "ssim_nbody_frame" if self.dim == 2 else "ssim_nbody_frame3d"
        _lib.call(entry, na, *[_lib.ptr(c) for c in a[:F]],
                  *[_lib.ptr(c) for c in st["aacc"]],
                  _lib.ptr(a[F]), _lib.ptr(st["avis"]), _lib.ptr(st["arad"]),
                  m, *[_lib.ptr(c) for c in M_[:F]],
                  *[_lib.ptr(c) for c in st["macc"]],
                  _lib.ptr(M_[F]), _lib.ptr(st["mvis"]),
                  _lib.ptr(st["mrad"]), float(self.G), float(self.soft),
                  float(dt_raster_years), int(k), float(self.max_sub_dt),
                  MAX_SUB_STEPS_PER_FRAME, YOSHIDA_W1, YOSHIDA_W0, self.m_sub,
                  _lib.ptr(st["snap"]), _lib.ptr(st["snapM"]),
                  _lib.ptr(st["subpos"]), _lib.ptr(st["subn"]),
                  _lib.ptr(st["hitt"]), _lib.ptr(st["hitm"]),
                  _lib.ptr(st["work"]), st["work"].numel(), s)

    def _collect_hits(self, st):
        """(idx, t, partner): contacts found by the fine loop
        (nbody_kernel.py:861-892)."""
        na = st["n_ast"]
        if not na:
            return (np.empty(0, np.int64), np.empty(0, np.float32),
                    np.empty(0, np.int64))
        ht = st["hitt"][:na].cpu().numpy()
        sel = np.flatnonzero(ht >= 0.0)
        hm = st["hitm"][:na].cpu().numpy()
        return (st["a_idx_h"][sel].astype(np.int64), ht[sel].astype(np.float32),
                st["m_idx_h"][hm[sel]].astype(np.int64))

    def _collect_sub(self, st, k):
        """Per raster (idx, pos (mSub, dim, nh)) of asteroids with a complete
        fine track, subN == mSub (nbody_kernel.py:894-934)."""
        if self.m_sub <= 0:
            return None
        ms, na, dim = self.m_sub, st["n_ast"], self.dim
        frames = []
        if not na:
            return [(np.empty(0, np.int64), np.empty((ms, dim, 0), np.float32))] * k
        subn = st["subn"][:k * na].cpu().numpy().reshape(k, na)
        pos = None
        for ks in range(k):
            sel = np.flatnonzero(subn[ks] == ms)
            if sel.size == 0:
                frames.append((np.empty(0, np.int64), np.empty((ms, dim, 0), np.float32)))
                continue
            if pos is None:
                pos = st["subpos"][:k * ms * dim * na].cpu().numpy().reshape(k, ms, dim, na)
            frames.append((st["a_idx_h"][sel].astype(np.int64),
                           np.ascontiguousarray(pos[ks][:, :, sel])))
        return frames

    def step(self, st: dict, dt_years: float) -> np.ndarray:
        return self.step_batch(st, dt_years, 1)[0]

    # ---------------- point updates ----------------
    def _scatter(self, st, idx, vals, add):
        torch = _torch()
        idx = np.asarray(idx, np.int64)
        vals = np.asarray(vals, np.float64).reshape(len(idx), -1)
        kind = st["inv_kind"][idx]
        pos = st["inv_pos"][idx]
        for which, sel in (("a", kind == 0), ("M", kind == -1)):
            if not sel.any():
                continue
            for f in range(vals.shape[1]):
                dst = st[which][f]
                gi = torch.as_tensor(pos[sel]).to(self._dev)
                gv = torch.as_tensor(np.ascontiguousarray(vals[sel, f])).to(self._dev)
                _lib.call("ssim_scatter_f64", int(sel.sum()), _lib.ptr(gi),
                          _lib.ptr(gv), _lib.ptr(dst), int(add),
                          D.stream(self._dev))

    def apply_updates(self, st, idx, vals) -> None:
        """Set x/y/vx/vy (x/y/z/vx/vy/vz in space) of bodies (original
        indices) absolutely."""
        self._scatter(st, idx, np.asarray(vals)[:, :2 * self.dim], add=False)

    def apply_deltas(self, st, idx, deltas) -> None:
        """Add position/velocity corrections; idx must be unique."""
        self._scatter(st, idx, np.asarray(deltas)[:, :2 * self.dim], add=True)

    def apply_body_state(self, st, idx, x, y, vx, vy, mass) -> None:
        """Set one body completely (merges in the film producer)."""
        self._scatter(st, [int(idx)], np.array([[x, y, vx, vy, mass]]), add=False)

    def set_radius(self, st, idx, wert) -> None:
        torch = _torch()
        kind = st["inv_kind"][int(idx)]
        pos = int(st["inv_pos"][int(idx)])
        dst = st["mrad"] if kind == -1 else st["arad"]
        gi = torch.as_tensor(np.array([pos], np.int64)).to(self._dev)
        gv = torch.as_tensor(np.array([float(wert)])).to(self._dev)
        _lib.call("ssim_scatter_f64", 1, _lib.ptr(gi), _lib.ptr(gv),
                  _lib.ptr(dst), 0, D.stream(self._dev))

    def deactivate_body(self, st, idx) -> None:
        """Freeze a body and remove its force: mass 0, visible 0."""
        torch = _torch()
        kind = st["inv_kind"][int(idx)]
        pos = int(st["inv_pos"][int(idx)])
        which, vis = ("M", st["mvis"]) if kind == -1 else ("a", st["avis"])
        gi = torch.as_tensor(np.array([pos], np.int64)).to(self._dev)
        gv = torch.zeros(1, dtype=torch.float64, device=self._dev)
        _lib.call("ssim_scatter_f64", 1, _lib.ptr(gi), _lib.ptr(gv),
                  _lib.ptr(st[which][2 * self.dim]), 0, D.stream(self._dev))
        vis[pos:pos + 1].zero_()

    def export_f64(self, st) -> np.ndarray:
        n, F = st["N"], 2 * self.dim
        out = np.empty(F * n, dtype="<f8")
        for f in range(F):
            if st["n_ast"]:
                out[f * n + st["a_idx_h"]] = st["a"][f][:st["n_ast"]].cpu().numpy()
            if st["m"]:
                out[f * n + st["m_idx_h"]] = st["M"][f][:st["m"]].cpu().numpy()
        return out


class NBodyCuda3D(NBodyCuda):
    """The raster integrator in space: the scheme of NBodyCuda with a z
    component (libssim's ssim_nbody_frame3d).  The reference integrates the
    plane only; this class serves scenes with inclinations.  Arrays gain z:
    `load_state(x, y, z, vx, vy, vz, mass, visible, is_ast, real_r)`,
    `step_batch` returns f32 (k, 6n) [x|y|z|vx|vy|vz], sub-samples are
    (idx, pos (mSub, 3, nh)), `export_f64` is <f8 (6n), updates take six
    columns.  G, the softening added to r^2 and the largest sub-step are
    parameters (defaults: the reference's G = 4 pi^2, 1e-6, half a day in
    years), so a scene can run in its own units (e.g. G = 1 with the time
    unit 1 yr / 2 pi and Plummer softening eps^2 = 1e-12 for eps = 1e-6 AU)."""

    dim = 3

    def __init__(self, devices, m_sub=SUB_SAMPLES, G=G_AU, softening=SOFTENING,
                 max_sub_dt=MAX_SUB_DT_YEARS):
        super().__init__(devices, m_sub)
        if not softening > 0.0:
            raise ValueError(f"softening muss > 0 sein: {softening}")
        if not max_sub_dt > 0.0:
            raise ValueError(f"max_sub_dt muss > 0 sein: {max_sub_dt}")
        self.G = float(G)
        self.soft = float(softening)
        self.max_sub_dt = float(max_sub_dt)

    def load_state(self, x, y, z, vx, vy, vz, mass, visible, is_ast,
                   real_r=None) -> dict:
        return self._load([x, y, z], [vx, vy, vz], mass, visible, is_ast, real_r)

    def apply_body_state(self, st, idx, x, y, z, vx, vy, vz, mass) -> None:
        self._scatter(st, [int(idx)], np.array([[x, y, z, vx, vy, vz, mass]]),
                      add=False)
