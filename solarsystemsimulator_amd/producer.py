"""The film producer's batch loop on one MI355X: film_producer.py:1152-1295
with the closures it calls (analyze_merge 707-779, apply_merges 781-861,
analyze_bounce 901-914, apply_bounce 1021-1041, analyze_batch /
analyze_sample 1055-1101, emit_event 687-705).

Per batch of K raster samples, in the reference's order:

1. the analysis of the previous batch is collected and applied to the
   resident state (merges and runaway removals, then bounce deltas, sample
   by sample), its events appended, and the head advanced to the samples
   produced so far (events are complete up to there);
2. the raster integrator runs K samples (NBodyCuda, one cooperative launch);
3. the K samples become original-order rows [x|y|vx|vy] on the device
   (ssim_frame_rows) and the analysis of this batch (massive-body contact
   sweep, runaway guard, asteroid collision search, bounce response) is
   queued on a second HIP stream from a worker thread;
4. meanwhile the K ring slots are written on the device (ssim_ring_slots)
   and copied into the ring in one transfer, and the integrator's contact
   times are recorded for the merges.

The reference moves every sample to the host (`step_batch`), scatters it,
uploads it again to its detection GPU and packs the slots on the host; here
the samples never leave HBM except as finished ring slots.  The results
are the same: slot bytes and event records equal the host route's.

Events are the ring's 32-byte records "<dIIfIff" (t_days, a, b, mass,
kind, x, y): kind 0 merge (a survivor, b absorbed), 1 bounce (heavier,
lighter), 2 runaway removal (a = 0xFFFFFFFF).
"""
from __future__ import annotations

import struct
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from . import device as D
from .constants import (BOUNCE_E, EV_BYTES, EV_FORMAT, G_AU, K_MAX,
                        SUB_MAX_ANTEIL, SUB_MAX_MIN)
from .detection import (Erkennungskarte, add_body_deltas, apply_merges,
                        bounce_deltas_device, bounce_suche, merge_scan)


def _torch():
    import torch
    return torch


# one event record, packed like struct "<dIIfIff" (EV_FORMAT)
EV_RECORD = np.dtype([("t", "<f8"), ("a", "<u4"), ("b", "<u4"), ("mass", "<f4"),
                      ("kind", "<u4"), ("x", "<f4"), ("y", "<f4")])
assert EV_RECORD.itemsize == EV_BYTES == struct.calcsize(EV_FORMAT)


def sub_max_fuer(n: int) -> int:
    """Track slots per ring slot for n bodies (film_producer.py:77-78)."""
    return min(n, max(SUB_MAX_MIN, n // SUB_MAX_ANTEIL))


class FilmProducer:
    """The producer loop around one NBodyCuda state and a film ring.

    ring: uint8 host array (capacity, slot_bytes) — e.g. a view of the
    shared-memory ring; slot k of the film lands in row k % capacity.
    mass / real_r / visible / is_ast / is_star_bh: the scene's flags by
    original index (the state `st` was loaded from the same scene).
    """

    def __init__(self, sim, st, ring, raster_days: float, *, mass, real_r,
                 visible, is_ast, is_star_bh=None, t0_days: float = 0.0,
                 ast_bounce: bool = True, sub_max: int | None = None, K: int = 8,
                 shatter_check: bool = False, G: float = G_AU):
        torch = _torch()
        if sim.dim != 2:
            raise ValueError("the film ring carries the plane (x, y) only")
        if not 1 <= K <= K_MAX:
            raise ValueError(f"K ausserhalb 1..{K_MAX}: {K}")
        n = int(st["N"])
        self.sim, self.st, self.n = sim, st, n
        self.dev = sim._dev
        self.K = int(K)
        self.raster_days = float(raster_days)
        self.t0_days = float(t0_days)
        self.dt_years = self.raster_days / 365.25
        self.G = float(G)
        self.ast_bounce = bool(ast_bounce)
        self.shatter_check = bool(shatter_check)
        if sub_max is None:
            sub_max = sub_max_fuer(n) if sim.m_sub else 0
        self.sub_max = int(sub_max)
        self.slot_bytes = int(_lib.load().ssim_ring_slot_bytes(n, sim.m_sub, self.sub_max))
        if ring.dtype != np.uint8 or ring.ndim != 2 or ring.shape[1] != self.slot_bytes:
            raise ValueError(f"ring muss uint8 (capacity, {self.slot_bytes}) sein, "
                             f"ist {ring.dtype} {ring.shape}")
        if ring.shape[0] < self.K:
            raise ValueError(f"ring capacity {ring.shape[0]} < K = {self.K}")
        self.ring = ring
        self.capacity = ring.shape[0]
        # host mirrors (read by the analysis, updated after each apply)
        self.mass = np.array(mass, np.float64, copy=True)
        self.real_r = np.array(real_r, np.float64, copy=True)
        self.vis = np.array(visible, np.uint8, copy=True)
        self.is_ast = np.asarray(is_ast).astype(np.uint8) != 0
        self.is_star_bh = (np.zeros(n, bool) if is_star_bh is None
                           else np.asarray(is_star_bh).astype(np.uint8) != 0)
        dev = self.dev
        # device mirrors, changed by the merges kernel
        self.mirrors = {
            "mass": D.to_dev(self.mass, dev),
            "vis": D.to_dev(self.vis, dev, torch.uint8),
            "real_r": D.to_dev(self.real_r, dev),
            "is_ast": D.to_dev(self.is_ast.astype(np.uint8), dev, torch.uint8),
            "is_star_bh": D.to_dev(self.is_star_bh.astype(np.uint8), dev, torch.uint8)}
        self.det = [Erkennungskarte(sim.device, self.is_ast, -np.inf, np.inf)]
        self.rr_max = 0.0
        self._stammdaten()
        self._det_dirty = False
        self.anteile = [1.0]
        # kernel contact times by original index, in samples (NaN: none)
        self.contact = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
        self.rows = torch.empty(self.K * 4 * max(n, 1), dtype=torch.float32, device=dev)
        self._prev_buf = torch.empty((2, max(n, 1)), dtype=torch.float32, device=dev)
        self._have_prev = False
        self.side = torch.cuda.Stream(dev)
        self._pool = ThreadPoolExecutor(max_workers=1)
        self._future = None
        self.k = 0                   # samples produced
        self.head = 0                # samples whose events are complete
        self.events = bytearray()
        self.collisions = 0
        self.bounce_count = 0
        self.sub_overflow = 0
        self.candidates = 0
        self.shatter = None          # (i, j, t_days) once a shatter stopped the loop
        self.times = {"step": 0.0, "wait": 0.0, "apply": 0.0, "analysis": 0.0}
        D.on_shutdown(self)

    # ---------------- public ----------------
    @property
    def n_events(self) -> int:
        return len(self.events) // EV_BYTES

    def run(self, batches: int) -> float:
        """Produce `batches` batches and apply their analyses; returns the
        production rate in simulated days per second of wall time."""
        t, k0 = time.perf_counter(), self.k
        for _ in range(int(batches)):
            if not self.batch():
                break
        self.finish()
        dt = time.perf_counter() - t
        return (self.k - k0) * self.raster_days / dt if dt > 0 else 0.0

    def batch(self) -> bool:
        """One pass of the loop: apply the previous analysis, produce K
        samples into the ring, queue their analysis.  False once a shatter
        has stopped the producer."""
        torch = _torch()
        main = torch.cuda.current_stream(self.dev)
        self._collect(main)
        if self.shatter is not None:
            return False
        K, k, st, sim = self.K, self.k, self.st, self.sim
        t = time.perf_counter()
        sim._run_frames(st, self.dt_years, K)
        sim.frame_rows(st, K, self.rows)
        ready = torch.cuda.Event()
        ready.record(main)
        self._future = self._pool.submit(self._analyze_batch, k, ready)
        sim._slots_enqueue(st, K, self.sub_max)
        c0 = k % self.capacity
        if c0 + K <= self.capacity:
            _, over = sim._slots_collect(st, K, self.slot_bytes, self.sub_max,
                                         self.ring[c0:c0 + K])
        else:
            tmp, over = sim._slots_collect(st, K, self.slot_bytes, self.sub_max, None)
            first = self.capacity - c0
            self.ring[c0:] = tmp[:first]
            self.ring[:K - first] = tmp[first:]
        self.sub_overflow += int(np.count_nonzero(over))
        idx, t_rel, _ = st["hits"]
        if len(idx):
            when = k + t_rel.astype(np.float64) / self.dt_years
            gi = torch.as_tensor(np.asarray(idx, np.int64)).to(self.dev)
            gv = torch.as_tensor(when).to(self.dev)
            _lib.call("ssim_scatter_f64", len(idx), _lib.ptr(gi), _lib.ptr(gv),
                      _lib.ptr(self.contact), 0, D.stream(self.dev))
        self.k += K
        self.times["step"] += time.perf_counter() - t
        return True

    def finish(self) -> None:
        """Apply the outstanding analysis (the last batch's events)."""
        self._collect(_torch().cuda.current_stream(self.dev))

    def close(self) -> None:
        if self._future is not None:
            self._future.result()
            self._future = None
        self._pool.shutdown(wait=True)

    # ---------------- analysis (worker thread, side stream) ----------------
    def _stammdaten(self) -> None:
        rr32 = self.real_r.astype(np.float32)
        for karte in self.det:
            karte.stammdaten(self.vis, rr32)
        self.rr_max = float(self.real_r[self.is_ast].max()) if self.is_ast.any() else 0.0

    def _analyze_batch(self, k0: int, ready) -> list:
        torch = _torch()
        t = time.perf_counter()
        with torch.cuda.stream(self.side):
            self.side.wait_event(ready)
            if self._det_dirty:
                self._stammdaten()
                self._det_dirty = False
            rows = self.rows[:self.K * 4 * self.n].view(self.K, 4, self.n)
            res = [self._analyze_sample(rows[i], k0 + i) for i in range(self.K)]
        self.times["analysis"] += time.perf_counter() - t
        return res

    def _analyze_sample(self, row, k_ev: int) -> dict:
        torch = _torch()
        gx, gy, gvx, gvy = row[0], row[1], row[2], row[3]
        karte = self.det[0]
        if self._have_prev:
            px, py = self._prev_buf[0, :self.n], self._prev_buf[1, :self.n]
        else:
            px, py = gx, gy
        pairs, runaway = [], None
        m_all = self.st["m_idx_h"]
        m_alive = m_all[(self.vis[m_all] != 0) & (self.mass[m_all] > 0)]
        if len(m_alive):
            mw = self.mass[m_alive]
            msum = float(mw.sum())
            bary, two_gm = None, 0.0
            if msum > 0:
                sel = torch.as_tensor(np.asarray(m_alive, np.int64)).to(self.dev)
                sxy = row[0:2].index_select(1, sel).cpu().numpy()
                bary = (float((sxy[0] * mw).sum() / msum), float((sxy[1] * mw).sum() / msum))
                two_gm = 2.0 * self.G * msum
            hit, runaway = merge_scan(self.dev, gx, gy, gvx, gvy, px, py, karte.g_rr,
                                      karte.g_vis, karte.g_ast, m_alive, self.dt_years,
                                      barycentre=bary, two_gm=two_gm)
            pairs = [(int(m_alive[r]), int(j), float(fr)) for r, j, fr in hit]
        self._prev_buf[0, :self.n].copy_(gx)
        self._prev_buf[1, :self.n].copy_(gy)
        self._have_prev = True
        bounce = None
        if self.ast_bounce:
            hits, cand, _, self.anteile = bounce_suche(
                self.det, None, [(gx, gy, gvx, gvy)], self.dt_years, self.rr_max,
                self.anteile)
            self.candidates += cand
            if hits is not None:
                bounce = bounce_deltas_device(self.dev, row.reshape(-1), hits,
                                              self.mirrors["real_r"], self.mirrors["mass"],
                                              self.dt_years, BOUNCE_E)
        return {"k": k_ev, "row": row, "pairs": pairs, "runaways": runaway,
                "bounce": bounce}

    # ---------------- apply (main thread, main stream) ----------------
    def _collect(self, main) -> None:
        if self._future is None:
            return
        t = time.perf_counter()
        results = self._future.result()
        self._future = None
        self.times["wait"] += time.perf_counter() - t
        main.wait_stream(self.side)
        t = time.perf_counter()
        for res in results:
            self._apply(res)
            if self.shatter is not None:
                break
        self.head = self.k
        self.times["apply"] += time.perf_counter() - t

    def _apply(self, res: dict) -> None:
        torch = _torch()
        run = [] if res["runaways"] is None else res["runaways"]
        if res["pairs"] or len(run):
            ev, merges, shatter = apply_merges(
                self.st, res["row"].reshape(-1), res["pairs"], run, res["k"],
                self.mirrors, self.contact, self.shatter_check, self.G,
                self.t0_days, self.raster_days)
            self.events += ev
            self.collisions += merges
            if ev:
                # only the bodies named by the records changed: a merge
                # changes a (mass, radius) and b (mass, visibility), a
                # removal its body b
                rec = np.frombuffer(ev, EV_RECORD)
                merged = rec["kind"] == 0
                touched = np.unique(np.concatenate([rec["a"][merged], rec["b"]]))
                sel = torch.as_tensor(touched.astype(np.int64)).to(self.dev)
                self.mass[touched] = self.mirrors["mass"].index_select(0, sel).cpu().numpy()
                self.vis[touched] = self.mirrors["vis"].index_select(0, sel).cpu().numpy()
                self.real_r[touched] = \
                    self.mirrors["real_r"].index_select(0, sel).cpu().numpy()
                self._det_dirty = True
            if shatter is not None:
                self.shatter = shatter
                return
        if res["bounce"] is None:
            return
        hi, hj, aff, cols = res["bounce"]
        main = torch.cuda.current_stream(self.dev)
        for t in (aff, *cols):               # made on the side stream, used here
            t.record_stream(main)
        add_body_deltas(self.st, aff, cols, self.dev)
        heavy = np.where(self.mass[hi] >= self.mass[hj], hi, hj)
        light = np.where(heavy == hi, hj, hi)
        sel = torch.as_tensor(heavy.astype(np.int64)).to(self.dev)
        xy = res["row"][0:2].index_select(1, sel).cpu().numpy()
        rec = np.empty(len(hi), EV_RECORD)
        rec["t"] = self.t0_days + (res["k"] + 1 + 0.0) * self.raster_days
        rec["a"] = heavy & 0xFFFFFFFF
        rec["b"] = light
        rec["mass"] = 0.0
        rec["kind"] = 1
        rec["x"], rec["y"] = xy[0], xy[1]
        self.events += rec.tobytes()
        self.collisions += len(hi)
        self.bounce_count += len(hi)
