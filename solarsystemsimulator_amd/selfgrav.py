"""Self-gravitating masses + massless tracers on one MI355X.

`SelfGravSystem` is the device-resident state (SoA float64 tensors, masses
first, tracers after them) with its libssim calls.  `NBodySelfGrav` wraps it
with the reference's object API (selfgrav_kernel.py:424-650) so the film
producer can call it unchanged: `load_state` / `step_batch` / `export_f64`
with the same argument meaning, array layouts and ValueError conditions.

Differences that are deliberate:
 * one GPU computes everything; a device list is accepted (and validated)
   but only its first entry is used — the reference's multi-card segments
   give bitwise-identical results anyway (selfgrav_kernel.py:49-52);
 * `kraft_f32=True` selects FP32 terms with hi/lo split positions and
   compensated accumulation instead of plain FP32 sums;
 * `dim=3` (keyword) runs the same physics in 3-D, `scheme="yoshida4"` the
   4th-order composition of three KDK sub-steps.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import device as D
from .constants import G_AU, G_MAX

SCHEMES = {"kdk": _lib.SCHEME_KDK, "yoshida4": _lib.SCHEME_YOSHIDA4}


class SelfGravSystem:
    """Device state of n masses followed by t tracers, one array per axis."""

    def __init__(self, dev, dim: int, n: int, t: int, prec: int, scheme: int,
                 G: float, eps2: float):
        if dim not in (2, 3):
            raise ValueError(f"dim must be 2 or 3: {dim}")
        if n < 1:
            raise ValueError("leerer Zustand")
        if not eps2 > 0:
            raise ValueError(f"Softening muss positiv sein: {eps2}")
        self.dev, self.dim, self.n, self.t = dev, dim, int(n), int(t)
        self.prec, self.scheme = prec, scheme
        self.G, self.eps2 = float(G), float(eps2)
        nb = self.n + self.t
        self.x = [D.zeros(nb, dev) for _ in range(dim)]
        self.v = [D.zeros(nb, dev) for _ in range(dim)]
        self.a = [D.zeros(nb, dev) for _ in range(dim)]
        self.m = D.zeros(self.n, dev)
        self.acc_valid = False
        lib = _lib.load()
        self.work = D.workspace(
            lib.ssim_selfgrav_work_bytes(dim, prec, self.n, self.t), dev)
        self.work_bytes = 0 if self.work is None else self.work.numel()
        self._ework = D.workspace(lib.ssim_energy_work_bytes(dim, self.n), dev)

    # -- helpers -----------------------------------------------------------
    def _p(self, lst, c):
        return _lib.ptr(lst[c]) if c < self.dim else None

    def set_bodies(self, pos, vel, mass, tpos=None, tvel=None):
        """Upload masses (pos/vel: dim arrays of n) and tracers (dim of t)."""
        for c in range(self.dim):
            self.x[c][:self.n].copy_(D.to_dev(pos[c], self.dev))
            self.v[c][:self.n].copy_(D.to_dev(vel[c], self.dev))
            if self.t:
                self.x[c][self.n:self.n + self.t].copy_(D.to_dev(tpos[c], self.dev))
                self.v[c][self.n:self.n + self.t].copy_(D.to_dev(tvel[c], self.dev))
        self.m[:self.n].copy_(D.to_dev(mass, self.dev))
        self.acc_valid = False

    def snap_width(self) -> int:
        return 2 * self.dim * self.n + self.dim * self.t

    # -- physics -----------------------------------------------------------
    def run(self, dt: float, steps: int, snap=None):
        """`steps` integrator steps; optional f32 snapshot tensor
        (steps, snap_width) on the device."""
        if steps < 1:
            raise ValueError(f"steps muss >= 1 sein: {steps}")
        if snap is not None:
            assert snap.dtype.itemsize == 4 and snap.numel() >= steps * self.snap_width()
        _lib.call("ssim_selfgrav_run", self.dim, self.prec, self.scheme,
                  self.n, self.t, self._p(self.x, 0), self._p(self.x, 1),
                  self._p(self.x, 2), self._p(self.v, 0), self._p(self.v, 1),
                  self._p(self.v, 2), self._p(self.a, 0), self._p(self.a, 1),
                  self._p(self.a, 2), _lib.ptr(self.m), self.G, self.eps2,
                  float(dt), int(steps), 0 if self.acc_valid else 1,
                  _lib.ptr(snap), self.snap_width(), _lib.ptr(self.work),
                  self.work_bytes, D.stream(self.dev))
        self.acc_valid = True

    def accelerations(self):
        """Recompute a(x) for all bodies; returns the dim device tensors."""
        _lib.call("ssim_allpairs", self.dim, self.prec, self.n,
                  self._p(self.x, 0), self._p(self.x, 1), self._p(self.x, 2),
                  _lib.ptr(self.m), self.n + self.t, self._p(self.x, 0),
                  self._p(self.x, 1), self._p(self.x, 2), self._p(self.a, 0),
                  self._p(self.a, 1), self._p(self.a, 2), self.G, self.eps2,
                  _lib.ptr(self.work), self.work_bytes, D.stream(self.dev))
        self.acc_valid = True
        return [a[:self.n + self.t] for a in self.a]

    def energy(self) -> np.ndarray:
        """[E_kin, E_pot, E] of the masses in FP64 (host array)."""
        import torch
        out = torch.zeros(3, dtype=torch.float64, device=self.dev)
        _lib.call("ssim_energy", self.dim, self.n, self._p(self.x, 0),
                  self._p(self.x, 1), self._p(self.x, 2), self._p(self.v, 0),
                  self._p(self.v, 1), self._p(self.v, 2), _lib.ptr(self.m),
                  self.G, self.eps2, _lib.ptr(out), _lib.ptr(self._ework),
                  self._ework.numel(), D.stream(self.dev))
        return out.cpu().numpy()

    def host_state(self):
        """(pos, vel) of all n + t bodies as float64 host arrays."""
        nb = self.n + self.t
        return ([self.x[c][:nb].cpu().numpy() for c in range(self.dim)],
                [self.v[c][:nb].cpu().numpy() for c in range(self.dim)])


def _device_of(devices) -> int:
    if isinstance(devices, (int, np.integer)):
        devices = [int(devices)]
    devices = list(devices)
    if not devices or len(devices) > G_MAX:
        raise ValueError(f"1..{G_MAX} Devices erwartet: {devices}")
    return devices


class NBodySelfGrav:
    """Drop-in for selfgrav_kernel.NBodySelfGrav (selfgrav_kernel.py:424)."""

    def __init__(self, devices, softening_au: float, kraft_f32: bool = True,
                 *, dim: int = 2, scheme: str = "kdk", G: float = G_AU):
        devices = _device_of(devices)
        if not softening_au > 0:
            raise ValueError(f"Softening muss positiv sein: {softening_au}")
        if scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r}")
        self.devices = devices
        self.device = devices[0]
        self.eps2 = float(softening_au) ** 2
        self.kraft_f32 = bool(kraft_f32)
        self.dim = int(dim)
        self.scheme = scheme
        self.G = float(G)
        self._dev = D.require_gpu(self.device)

    def name(self) -> str:
        import torch
        return torch.cuda.get_device_name(self._dev)

    def load_state(self, x, y, vx, vy, mass, *_egal, tracer=None,
                   _kalibrieren: bool = False, z=None, vz=None,
                   **_auch_egal) -> dict:
        """Take over the full state (selfgrav_kernel.py:463-558).  Extra
        positional/keyword arguments of NBodyCuda.load_state are accepted and
        ignored, as in the reference.  For dim=3 pass z=, vz= and a 6-tuple
        tracer (tx, ty, tz, tvx, tvy, tvz)."""
        n = len(x)
        if n < 1:
            raise ValueError("leerer Zustand")
        d = self.dim
        if d == 3 and (z is None or vz is None):
            raise ValueError("dim=3 needs z= and vz=")
        pos = [x, y] + ([z] if d == 3 else [])
        vel = [vx, vy] + ([vz] if d == 3 else [])
        t = 0
        tpos = tvel = None
        if tracer is not None and len(tracer[0]):
            t = len(tracer[0])
            tpos = list(tracer[:d])
            tvel = list(tracer[d:2 * d])
        sysm = SelfGravSystem(self._dev, d, n, t,
                              _lib.PREC_F32C if self.kraft_f32 else _lib.PREC_F64,
                              SCHEMES[self.scheme], self.G, self.eps2)
        sysm.set_bodies(pos, vel, np.asarray(mass, np.float64), tpos, tvel)
        return {"N": n, "T": t, "sys": sysm}

    def step_batch(self, st: dict, dt_years: float, steps: int) -> np.ndarray:
        """`steps` steps; returns f32 (steps, 4N+2T) [x|y|vx|vy|tx|ty]
        (selfgrav_kernel.py:560-627); (steps, 6N+3T) for dim=3."""
        import torch
        if st is None:
            raise ValueError("kein Zustand geladen")
        if steps < 1:
            raise ValueError(f"steps muss >= 1 sein: {steps}")
        sysm = st["sys"]
        snap = torch.empty((steps, sysm.snap_width()), dtype=torch.float32,
                           device=self._dev)
        sysm.run(dt_years, steps, snap)
        return snap.cpu().numpy()

    def export_f64(self, st: dict) -> np.ndarray:
        """Exact f64 state [x|y|vx|vy] (4N) of the masses
        (selfgrav_kernel.py:629-650); [x|y|z|vx|vy|vz] for dim=3."""
        sysm = st["sys"]
        pos, vel = sysm.host_state()
        n = sysm.n
        return np.concatenate([p[:n] for p in pos] + [v[:n] for v in vel]
                              ).astype("<f8")
