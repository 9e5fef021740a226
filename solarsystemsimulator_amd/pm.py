"""Particle-mesh gravity on one MI355X (libssim FFT, CIC and Green's
function kernels; torch tensors hold the device memory).

Two flavours share the kernels:

* the reference's 2-D isolated convolution with the force Green's function
  on a zero-padded (2N)^2 mesh and an adaptive grid (pm_kernel.py:34-293):
  `build_force_kernels`, `grid_fuer`, `pm_force_field`, `gather_accel`,
  `pm_accelerations`, `_cic_deposit`, `_cic_gather` and `NBodyPM` keep its
  names, arguments and array shapes; device arrays are torch tensors where
  the reference takes CuPy arrays.  The same convolution also runs in 3-D
  (`IsolatedPM3D`), zero-padded to (2N)^3;
* a periodic 3-D Poisson solver (`PeriodicPM3D`): CIC deposit, R2C FFT,
  phi_k = -4 pi G rho_k / k^2, C2R, central-difference gradient, CIC gather,
  with the mesh in FP64 or FP32.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as D
from .constants import G_AU

F64 = _lib.PREC_F64
F32 = _lib.PREC_F32C   # FP32 mesh
MAX_FFT_LINE = 8192    # longest FFT line (points) of ssim_fft_*


def _addr(a):
    return ctypes.addressof(a)


def _torch():
    import torch
    return torch


def _mesh_dtype(prec):
    torch = _torch()
    return torch.float64 if prec == F64 else torch.float32


class FFTPlan:
    """Twiddle tables + shapes of a real<->half-complex FFT on the device."""

    _cache: dict = {}

    @classmethod
    def get(cls, dims, prec, dev):
        key = (tuple(int(d) for d in dims), prec, str(dev))
        plan = cls._cache.get(key)
        if plan is None:
            plan = cls._cache[key] = cls(dims, prec, dev)
        return plan

    def __init__(self, dims, prec, dev):
        self.dims = [int(d) for d in dims]
        self.prec, self.dev = prec, dev
        self._c_dims = _lib.int_array(self.dims)
        lib = _lib.load()
        nd = len(self.dims)
        self.work_bytes = lib.ssim_fft_work_bytes(nd, _addr(self._c_dims), prec)
        if self.work_bytes == 0:
            raise ValueError(f"unsupported FFT dims {self.dims} (powers of two "
                             f"up to {MAX_FFT_LINE})")
        self.work = D.workspace(self.work_bytes, dev)
        _lib.call("ssim_fft_prepare", nd, _addr(self._c_dims), prec,
                  _lib.ptr(self.work), self.work_bytes, D.stream(dev))
        self.rdtype = _mesh_dtype(prec)
        self.cshape = self.dims[:-1] + [self.dims[-1] // 2 + 1, 2]

    def new_complex(self):
        return _torch().empty(self.cshape, dtype=self.rdtype, device=self.dev)

    def r2c(self, real, out=None):
        out = self.new_complex() if out is None else out
        _lib.call("ssim_fft_r2c", len(self.dims), _addr(self._c_dims), self.prec,
                  _lib.ptr(real), _lib.ptr(out), _lib.ptr(self.work),
                  self.work_bytes, D.stream(self.dev))
        return out

    def c2r(self, spec, out=None):
        """Inverse (numpy irfftn); overwrites `spec`."""
        out = (_torch().empty(self.dims, dtype=self.rdtype, device=self.dev)
               if out is None else out)
        _lib.call("ssim_fft_c2r", len(self.dims), _addr(self._c_dims), self.prec,
                  _lib.ptr(spec), _lib.ptr(out), _lib.ptr(self.work),
                  self.work_bytes, D.stream(self.dev))
        return out


def _dev_of(t):
    return t.device


def minmax(arrays):
    """[(min, max)] of 1..3 device float64 arrays (one host sync)."""
    torch = _torch()
    dev = _dev_of(arrays[0])
    n = arrays[0].numel()
    lib = _lib.load()
    wb = lib.ssim_minmax_work_bytes(n)
    w = D.workspace(wb, dev)
    out = torch.empty(6, dtype=torch.float64, device=dev)
    ps = [_lib.ptr(a) for a in arrays] + [None] * (3 - len(arrays))
    _lib.call("ssim_minmax", n, len(arrays), *ps, _lib.ptr(out), _lib.ptr(w),
              wb, D.stream(dev))
    h = out.cpu().numpy()
    return [(float(h[2 * c]), float(h[2 * c + 1])) for c in range(len(arrays))]


_DEPOSIT_WORK: dict = {}


def deposit(pos, m, grid_n, origin, h, periodic, prec=F64, mesh=None,
            cells=None, tiled=True):
    """CIC mass deposit; returns the mesh (shape grid_n, FP64/FP32).
    tiled (default): ssim_cic_deposit_tiles — int64 fixed-point sums in
    LDS tiles, every node written once, bit-identical from run to run;
    False: ssim_cic_deposit, FP atomics on the mesh (order not fixed, like
    the reference's cp.bincount)."""
    torch = _torch()
    dim = len(pos)
    dev = _dev_of(pos[0])
    if mesh is None:
        mesh = (torch.empty if tiled else torch.zeros)(list(grid_n), dtype=_mesh_dtype(prec),
                                                       device=dev)
    elif not tiled:
        mesh.zero_()
    gn = _lib.int_array(grid_n)
    org = _lib.double_array(origin)
    ps = [_lib.ptr(p) for p in pos] + [None] * (3 - dim)
    n = pos[0].numel()
    if not tiled:
        _lib.call("ssim_cic_deposit", dim, int(periodic), n, *ps,
                  _lib.ptr(m), _addr(gn), _addr(org), float(h), prec,
                  _lib.ptr(mesh), _lib.ptr(cells), D.stream(dev))
        return mesh
    key = (str(dev), dim, n, tuple(int(g) for g in grid_n))
    w = _DEPOSIT_WORK.get(key)
    if w is None:
        nb = _lib.load().ssim_cic_deposit_tiles_work_bytes(dim, n, _addr(gn))
        w = _DEPOSIT_WORK[key] = D.workspace(nb, dev)
    _lib.call("ssim_cic_deposit_tiles", dim, int(periodic), n, *ps,
              _lib.ptr(m), _addr(gn), _addr(org), float(h), prec,
              _lib.ptr(mesh), _lib.ptr(cells), _lib.ptr(w), w.numel(), D.stream(dev))
    return mesh


def gather(fields, pos, grid_n, origin, h, periodic, prec=F64, outs=None):
    """CIC interpolation of 1..3 mesh fields at pos; list of float64 arrays."""
    torch = _torch()
    dim = len(pos)
    dev = _dev_of(pos[0])
    n = pos[0].numel()
    if outs is None:
        outs = [torch.empty(n, dtype=torch.float64, device=dev) for _ in fields]
    gn = _lib.int_array(grid_n)
    org = _lib.double_array(origin)
    ps = [_lib.ptr(p) for p in pos] + [None] * (3 - dim)
    fs = [_lib.ptr(f) for f in fields] + [None] * (3 - len(fields))
    os_ = [_lib.ptr(o) for o in outs] + [None] * (3 - len(outs))
    _lib.call("ssim_cic_gather", dim, int(periodic), n, *ps, _addr(gn),
              _addr(org), float(h), prec, len(fields), *fs, *os_,
              D.stream(dev))
    return outs


# ---------------------------------------------------------------------------
# The reference's 2-D isolated API (pm_kernel.py), on torch device tensors
# ---------------------------------------------------------------------------

def build_force_kernels(grid_n: int, h: float, eps2: float, G: float = G_AU,
                        dim: int = 2, prec: int = F64, dev=None):
    """Spectra of the force Green's functions on the zero-padded (2N)^dim
    mesh (pm_kernel.py:34-60): tuple of dim half spectra, each of shape
    (2N, .., N+1, 2) (interleaved complex)."""
    torch = _torch()
    dev = dev or D.require_gpu(0)
    P = 2 * int(grid_n)
    plan = FFTPlan.get([P] * dim, prec, dev)
    real = torch.empty([P] * dim, dtype=_mesh_dtype(prec), device=dev)
    out = []
    for c in range(dim):
        _lib.call("ssim_pm_force_kernel", dim, int(grid_n), c, float(h),
                  float(eps2), float(G), 0.0, 0.0, prec, _lib.ptr(real),
                  D.stream(dev))
        out.append(plan.r2c(real))
    return tuple(out)


def _cic_deposit(gx, gy, gm, grid_n, x0, y0, h):
    """Mass mesh (N, N) of the CIC deposit (pm_kernel.py:63-90)."""
    return deposit([gx, gy], gm, [grid_n, grid_n], [x0, y0], h, False)


def _cic_gather(feld, gx, gy, grid_n, x0, y0, h):
    """Field value at (gx, gy) by CIC (pm_kernel.py:93-107)."""
    return gather([feld.contiguous()], [gx, gy], [grid_n, grid_n], [x0, y0], h,
                  False)[0]


def grid_fuer(gx, gy, grid_n: int, rand_zellen: float = 3.0):
    """Origin and cell size so every point fits with `rand_zellen` cells of
    margin (pm_kernel.py:110-124); min/max reduced on the device."""
    (xmin, xmax), (ymin, ymax) = minmax([gx, gy])
    spanne = max(xmax - xmin, ymax - ymin, 1e-30)
    h = spanne / (grid_n - 2 * rand_zellen)
    return xmin - rand_zellen * h, ymin - rand_zellen * h, h


def _isolated_field(pos, gm, grid_n, origin, h, eps2, G, kernels, prec=F64):
    """Force field on the n^dim mesh by convolution on the padded mesh."""
    torch = _torch()
    dim = len(pos)
    dev = _dev_of(pos[0])
    P = 2 * int(grid_n)
    if kernels is None:
        kernels = build_force_kernels(grid_n, h, eps2, G, dim, prec, dev)
    plan = FFTPlan.get([P] * dim, prec, dev)
    rho = deposit(pos, gm, [grid_n] * dim, origin, h, False, prec)
    pad = torch.zeros([P] * dim, dtype=_mesh_dtype(prec), device=dev)
    _lib.call("ssim_pm_pad", dim, int(grid_n), prec, _lib.ptr(rho),
              _lib.ptr(pad), 0, D.stream(dev))
    fr = plan.r2c(pad)
    ncomplex = fr.numel() // 2
    prod = plan.new_complex()
    full = torch.empty([P] * dim, dtype=_mesh_dtype(prec), device=dev)
    fields = []
    for fk in kernels:
        _lib.call("ssim_spectrum_mul", ncomplex, prec, _lib.ptr(fr),
                  _lib.ptr(fk), _lib.ptr(prod), D.stream(dev))
        plan.c2r(prod, full)
        crop = torch.empty([grid_n] * dim, dtype=_mesh_dtype(prec), device=dev)
        _lib.call("ssim_pm_pad", dim, int(grid_n), prec, _lib.ptr(full),
                  _lib.ptr(crop), 1, D.stream(dev))
        fields.append(crop)
    return fields


def pm_force_field(gx, gy, gm, grid_n, x0, y0, h, eps2, G: float = G_AU,
                   kernels=None):
    """(ax_grid, ay_grid) on the N x N mesh (pm_kernel.py:127-147)."""
    ax, ay = _isolated_field([gx, gy], gm, grid_n, [x0, y0], h, eps2, G,
                             kernels)
    return ax, ay


def gather_accel(ax_grid, ay_grid, px, py, grid_n, x0, y0, h):
    """Field at (px, py) for masses or tracers (pm_kernel.py:150-154)."""
    ax, ay = gather([ax_grid, ay_grid], [px, py], [grid_n, grid_n], [x0, y0],
                    h, False)
    return ax, ay


def pm_accelerations(gx, gy, gm, grid_n, x0, y0, h, eps2, G: float = G_AU,
                     kernels=None):
    """Field built from the masses and read at their own positions
    (pm_kernel.py:157-163)."""
    axg, ayg = pm_force_field(gx, gy, gm, grid_n, x0, y0, h, eps2, G, kernels)
    return gather_accel(axg, ayg, gx, gy, grid_n, x0, y0, h)


def _kick_snap(dim, n, t, x, v, a, h, snap, dev):
    ps = lambda lst: [_lib.ptr(c) for c in lst] + [None] * (3 - dim)  # noqa: E731
    _lib.call("ssim_kick_snap", dim, n, t, *ps(x), *ps(v), *ps(a), float(h),
              _lib.ptr(snap), D.stream(dev))


def _drift(dim, nb, x, v, h, dev):
    ps = lambda lst: [_lib.ptr(c) for c in lst] + [None] * (3 - dim)  # noqa: E731
    _lib.call("ssim_drift", dim, nb, *ps(x), *ps(v), float(h), D.stream(dev))


class NBodyPM:
    """Drop-in for pm_kernel.NBodyPM (pm_kernel.py:166-293): leapfrog KDK
    with the adaptive isolated 2-D mesh, masses and tracers read the same
    field; `load_state` / `step_batch` / `export_f64` as in the reference."""

    def __init__(self, devices, softening_au: float = 0.0,
                 kraft_f32: bool = True, grid_n: int = 4096,
                 softening_zellen: float = 1.5, rand_zellen: float = 4.0):
        if isinstance(devices, (int, np.integer)):
            devices = [int(devices)]
        self.device = list(devices)[0]
        self.devices = [self.device]
        self.grid_n = int(grid_n)
        self.softening_zellen = float(softening_zellen)
        self.rand_zellen = float(rand_zellen)
        self.softening_floor = float(softening_au)
        self.kraft_f32 = bool(kraft_f32)
        self._kernels = None
        self._kernel_h = None
        # the reference's default grid_n = 4096 (pm_kernel.py:190) pads to
        # 8192-point lines, the longest the FFT here handles; larger grids
        # are rejected before any device work
        if not (2 <= self.grid_n and 2 * self.grid_n <= MAX_FFT_LINE
                and self.grid_n & (self.grid_n - 1) == 0):
            raise ValueError(f"grid_n={grid_n}: NBodyPM needs a power of two with the "
                             f"padded mesh 2N <= {MAX_FFT_LINE} (grid_n <= "
                             f"{MAX_FFT_LINE // 2}); pass grid_n explicitly")
        self._dev = D.require_gpu(self.device)

    def name(self) -> str:
        return _torch().cuda.get_device_name(self._dev) + " (PM)"

    def _accel_into(self, st):
        n, t = st["N"], st["T"]
        gx, gy = st["x"][0][:n], st["x"][1][:n]
        x0, y0, h = grid_fuer(gx, gy, self.grid_n, self.rand_zellen)
        eps = max(self.softening_zellen * h, self.softening_floor)
        eps2 = eps * eps
        if self._kernel_h is None or abs(h - self._kernel_h) > 1e-6 * h:
            self._kernels = build_force_kernels(self.grid_n, h, eps2,
                                                dev=self._dev)
            self._kernel_h = h
        axg, ayg = pm_force_field(gx, gy, st["m"], self.grid_n, x0, y0, h,
                                  eps2, kernels=self._kernels)
        gather([axg, ayg], [st["x"][0][:n + t], st["x"][1][:n + t]],
               [self.grid_n] * 2, [x0, y0], h, False,
               outs=[st["a"][0][:n + t], st["a"][1][:n + t]])

    def load_state(self, x, y, vx, vy, mass, *_egal, tracer=None,
                   **_auch_egal) -> dict:
        torch = _torch()
        n = len(x)
        t = len(tracer[0]) if tracer is not None and len(tracer[0]) else 0

        def cat(a, b):
            return D.to_dev(np.concatenate([np.asarray(a, np.float64),
                                            np.asarray(b, np.float64)]),
                            self._dev)
        tx = tracer if t else ([], [], [], [])
        st = {"N": n, "T": t,
              "x": [cat(x, tx[0]), cat(y, tx[1])],
              "v": [cat(vx, tx[2]), cat(vy, tx[3])],
              "m": D.to_dev(mass, self._dev),
              "a": [torch.zeros(n + t, dtype=torch.float64, device=self._dev)
                    for _ in range(2)]}
        self._accel_into(st)
        return st

    def step_batch(self, st: dict, dt_years: float, steps: int) -> np.ndarray:
        """`steps` KDK steps; f32 (steps, 4N+2T) [x|y|vx|vy|tx|ty]
        (pm_kernel.py:254-281)."""
        torch = _torch()
        if steps < 1:
            raise ValueError(f"steps muss >= 1 sein: {steps}")
        n, t = st["N"], st["T"]
        dt = float(dt_years)
        hdt = 0.5 * dt
        snap = torch.empty((steps, 4 * n + 2 * t), dtype=torch.float32,
                           device=self._dev)
        for s in range(steps):
            _kick_snap(2, n, t, st["x"], st["v"], st["a"], hdt, None, self._dev)
            _drift(2, n + t, st["x"], st["v"], dt, self._dev)
            self._accel_into(st)
            _kick_snap(2, n, t, st["x"], st["v"], st["a"], hdt, snap[s],
                       self._dev)
        return snap.cpu().numpy()

    def export_f64(self, st: dict) -> np.ndarray:
        n = st["N"]
        return np.concatenate([c[:n].cpu().numpy() for c in st["x"] + st["v"]]
                              ).astype("<f8")


# ---------------------------------------------------------------------------
# 3-D solvers
# ---------------------------------------------------------------------------

class PeriodicPM3D:
    """Periodic Poisson PM on an n^3 mesh over the box [origin, origin+L)^3:
    CIC deposit -> R2C -> -4 pi G rho_k / k^2 -> C2R -> central difference
    -> CIC gather.  Mesh in FP64 (prec=F64) or FP32 (prec=F32)."""

    def __init__(self, dev, n: int, box: float = 1.0, origin=(0.0, 0.0, 0.0),
                 G: float = 1.0, prec: int = F64):
        torch = _torch()
        self.dev, self.n, self.G, self.prec = dev, int(n), float(G), prec
        self.h = float(box) / self.n
        self.origin = [float(o) for o in origin]
        self.grid = [self.n] * 3
        self.plan = FFTPlan.get(self.grid, prec, dev)
        dt = _mesh_dtype(prec)
        self.rho = torch.zeros(self.grid, dtype=dt, device=dev)
        self.phi = torch.empty(self.grid, dtype=dt, device=dev)
        self.spec = self.plan.new_complex()
        self.g = [torch.empty(self.grid, dtype=dt, device=dev) for _ in range(3)]
        self._gn = _lib.int_array(self.grid)

    # True: the potential by ssim_pm_potential_periodic (the x axis's forward
    # transform, Poisson factor and inverse transform in one pass); False:
    # R2C, ssim_pm_poisson, C2R as three calls (the same result bit for bit)
    fused = True

    def field(self, pos, m):
        """Acceleration field (3 meshes) of masses m at pos."""
        deposit(pos, m, self.grid, self.origin, self.h, True, self.prec,
                mesh=self.rho)
        if self.fused:
            _lib.call("ssim_pm_potential_periodic", _addr(self._gn), self.prec,
                      _lib.ptr(self.rho), _lib.ptr(self.spec), _lib.ptr(self.phi),
                      self.h, self.G, _lib.ptr(self.plan.work), self.plan.work_bytes,
                      D.stream(self.dev))
        else:
            self.plan.r2c(self.rho, self.spec)
            _lib.call("ssim_pm_poisson", 3, _addr(self._gn), self.h, self.G,
                      self.prec, _lib.ptr(self.spec), D.stream(self.dev))
            self.plan.c2r(self.spec, self.phi)
        _lib.call("ssim_pm_gradient", 3, _addr(self._gn), self.h, self.prec,
                  _lib.ptr(self.phi), *[_lib.ptr(g) for g in self.g],
                  D.stream(self.dev))
        return self.g

    def accelerations(self, pos, m, targets, outs=None):
        self.field(pos, m)
        return gather(self.g, targets, self.grid, self.origin, self.h, True,
                      self.prec, outs)


class PeriodicPMSystem:
    """n masses in the periodic box [origin, origin + box)^3 advanced with
    leapfrog KDK (scheme SCHEME_KDK) or Yoshida-4 (SCHEME_YOSHIDA4: three
    KDK sub-steps weighted w1, w0, w1) on PeriodicPM3D forces, FP64 state,
    mesh in FP64 or FP32.  A sub-step of weight w is: v += a w dt/2;
    x += v w dt wrapped into the box (ssim_drift_periodic); a = PM force;
    v += a w dt/2 — the KDK of the reference's NBodyPM.step_batch
    (pm_kernel.py:254-281) in a periodic box."""

    def __init__(self, dev, n_mesh: int, n: int, box: float = 1.0,
                 origin=(0.0, 0.0, 0.0), G: float = 1.0, prec: int = F64,
                 scheme: int = _lib.SCHEME_KDK):
        torch = _torch()
        if scheme not in (_lib.SCHEME_KDK, _lib.SCHEME_YOSHIDA4):
            raise ValueError(f"unknown scheme {scheme}")
        if n < 1:
            raise ValueError("leerer Zustand")
        self.dev, self.n, self.box, self.scheme = dev, int(n), float(box), scheme
        self.solver = PeriodicPM3D(dev, n_mesh, box, origin, G, prec)
        self._org = _lib.double_array(self.solver.origin)
        f64 = dict(dtype=torch.float64, device=dev)
        self.x = [torch.zeros(self.n, **f64) for _ in range(3)]
        self.v = [torch.zeros(self.n, **f64) for _ in range(3)]
        self.a = [torch.zeros(self.n, **f64) for _ in range(3)]
        self.m = torch.zeros(self.n, **f64)
        self.acc_valid = False

    def set_bodies(self, pos, vel, m):
        for c in range(3):
            self.x[c].copy_(D.to_dev(pos[c], self.dev))
            self.v[c].copy_(D.to_dev(vel[c], self.dev))
        self.m.copy_(D.to_dev(m, self.dev))
        self.acc_valid = False

    def accelerations(self):
        """a = PM acceleration at the current positions."""
        self.solver.accelerations(self.x, self.m, self.x, outs=self.a)
        self.acc_valid = True
        return self.a

    def weights(self):
        if self.scheme == _lib.SCHEME_KDK:
            return (1.0,)
        from .constants import YOSHIDA_W0, YOSHIDA_W1
        return (YOSHIDA_W1, YOSHIDA_W0, YOSHIDA_W1)

    def run(self, dt: float, steps: int):
        if steps < 1:
            raise ValueError(f"steps muss >= 1 sein: {steps}")
        if not self.acc_valid:
            self.accelerations()
        st = D.stream(self.dev)
        ps = lambda lst: [_lib.ptr(c) for c in lst]  # noqa: E731
        for _ in range(steps):
            for w in self.weights():
                full = w * dt
                half = 0.5 * full
                _lib.call("ssim_kick", 3, self.n, *ps(self.v), *ps(self.a), half, st)
                _lib.call("ssim_drift_periodic", 3, self.n, *ps(self.x), *ps(self.v),
                          full, _addr(self._org), self.box, st)
                self.accelerations()
                _lib.call("ssim_kick", 3, self.n, *ps(self.v), *ps(self.a), half, st)

    def potential(self):
        """Mesh potential phi gathered at the bodies with the same CIC
        stencil (device float64 array); E_pot = 1/2 sum m phi."""
        s = self.solver
        s.field(self.x, self.m)
        return gather([s.phi], self.x, s.grid, s.origin, s.h, True, s.prec)[0]

    def host_state(self):
        return ([c.cpu().numpy() for c in self.x], [c.cpu().numpy() for c in self.v])


class IsolatedPM3D:
    """Isolated 3-D PM: convolution with the Plummer force Green's function
    on the (2n)^3 zero-padded mesh (the reference's 2-D method in 3-D) over
    a fixed grid [origin, origin + n h)^3."""

    def __init__(self, dev, n: int, origin, h: float, eps2: float,
                 G: float = 1.0, prec: int = F64):
        self.dev, self.n, self.prec = dev, int(n), prec
        self.origin = [float(o) for o in origin]
        self.h, self.eps2, self.G = float(h), float(eps2), float(G)
        self.kernels = build_force_kernels(self.n, self.h, self.eps2, self.G,
                                           3, prec, dev)

    def field(self, pos, m):
        return _isolated_field(pos, m, self.n, self.origin, self.h, self.eps2,
                               self.G, self.kernels, self.prec)

    def accelerations(self, pos, m, targets, outs=None):
        g = self.field(pos, m)
        return gather(g, targets, [self.n] * 3, self.origin, self.h, False,
                      self.prec, outs)
