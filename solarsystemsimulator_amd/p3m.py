"""Split particle-mesh force for large self-gravitating systems (3-D).

a = a_mesh + a_near (+ a_learned, see learned.py):

* a_mesh — isolated PM on an n^3 mesh zero-padded to (2n)^3: CIC deposit of
  the masses, convolution with the long-range Plummer force
  F(d) (1 - g(|d|)), g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2),
  whose spectrum is divided by the squared CIC window, CIC gather at every
  body (masses and tracers);
* a_near — the complementary F(d) g(|d|) summed directly over sources
  within rcut through a cell list.

The reference's PM (pm_kernel.py:127-229) convolves with the full softened
force on an adaptive mesh; this class keeps that adaptive, isolated mesh
and adds the near-field term so that the sum approaches the direct sum.
The mesh follows the bodies every evaluation (min/max on the device); its
cell size is rounded up to a ladder of 2^(1/64) steps so the Green's
function spectra are rebuilt only when the ladder step changes.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from . import device as D
from . import pm as P


class SplitForce:
    """Mesh + near-field force of n masses on n + t bodies (3-D, FP64 state).

    mesh_n     mesh nodes per axis (padded FFT size 2 mesh_n)
    rs_cells   split scale rs in mesh cells
    rcut_rs    near-field cutoff in units of rs
    margin     empty mesh cells kept around the bodies
    prec       mesh precision (pm.F64 or pm.F32)
    """

    def __init__(self, dev, mesh_n: int = 128, eps2: float = 1e-6,
                 G: float = 1.0, rs_cells: float = 1.25, rcut_rs: float = 6.0,
                 margin: float = 2.0, prec: int = P.F64):
        import torch
        self.dev, self.n, self.eps2, self.G = dev, int(mesh_n), float(eps2), float(G)
        self.rs_cells, self.rcut_rs = float(rs_cells), float(rcut_rs)
        self.margin, self.prec = float(margin), prec
        self.Pn = 2 * self.n
        self.plan = P.FFTPlan.get([self.Pn] * 3, prec, dev)
        dt = P._mesh_dtype(prec)
        self.rho = torch.zeros([self.n] * 3, dtype=dt, device=dev)
        self.pad = torch.zeros([self.Pn] * 3, dtype=dt, device=dev)
        self.full = torch.empty([self.Pn] * 3, dtype=dt, device=dev)
        self.spec = self.plan.new_complex()
        self.prod = self.plan.new_complex()
        self.fields = [torch.empty([self.n] * 3, dtype=dt, device=dev)
                       for _ in range(3)]
        self.kspec = None
        self.level = None
        self.h = None
        self.origin = None
        self._cells_cap = 0

    # -- grid ---------------------------------------------------------------
    def _place(self, x):
        """Mesh origin and ladder cell size covering every body."""
        box = P.minmax(x)
        span = max(hi - lo for lo, hi in box)
        need = max(span, 1e-30) / (self.n - 1 - 2 * self.margin)
        level = math.ceil(64.0 * math.log2(need))
        h = 2.0 ** (level / 64.0)
        origin = [0.5 * (lo + hi) - 0.5 * (self.n - 1) * h for lo, hi in box]
        return level, h, origin, box

    def _rebuild_kernels(self, h):
        import torch
        rs = self.rs_cells * h
        real = self.full
        self.kspec = []
        for c in range(3):
            _lib.call("ssim_pm_force_kernel", 3, self.n, c, h, self.eps2,
                      self.G, rs, self.prec, _lib.ptr(real), D.stream(self.dev))
            spec = self.plan.r2c(real)
            _lib.call("ssim_pm_deconvolve", 3, self.Pn, self.prec,
                      _lib.ptr(spec), D.stream(self.dev))
            self.kspec.append(spec)
        del torch

    # -- force --------------------------------------------------------------
    def mesh_accel(self, x, n_mass, m, a):
        """a = long-range mesh acceleration at all bodies (overwrites a)."""
        level, h, origin, box = self._place(x)
        if level != self.level:
            self._rebuild_kernels(h)
            self.level = level
        self.h, self.origin, self.box = h, origin, box
        src = [c[:n_mass] for c in x]
        P.deposit(src, m[:n_mass], [self.n] * 3, origin, h, False, self.prec,
                  mesh=self.rho)
        self.pad.zero_()
        st = D.stream(self.dev)
        _lib.call("ssim_pm_pad", 3, self.n, self.prec, _lib.ptr(self.rho),
                  _lib.ptr(self.pad), 0, st)
        self.plan.r2c(self.pad, self.spec)
        nc = self.spec.numel() // 2
        for c in range(3):
            _lib.call("ssim_spectrum_mul", nc, self.prec, _lib.ptr(self.spec),
                      _lib.ptr(self.kspec[c]), _lib.ptr(self.prod), st)
            self.plan.c2r(self.prod, self.full)
            _lib.call("ssim_pm_pad", 3, self.n, self.prec, _lib.ptr(self.full),
                      _lib.ptr(self.fields[c]), 1, st)
        P.gather(self.fields, x, [self.n] * 3, origin, h, False, self.prec,
                 outs=a)
        return a

    def near_accel(self, x, n_mass, m, a):
        """a += near-field term (cell list of the masses)."""
        import torch
        h = self.h
        rs = self.rs_cells * h
        rcut = self.rcut_rs * rs
        lo = [b[0] for b in self.box]
        span = [b[1] - b[0] for b in self.box]
        nc = [max(1, int(s // rcut)) for s in span]
        cell = max(max(s / k for s, k in zip(span, nc)), rcut)
        ncells = nc[0] * nc[1] * nc[2]
        if self._cells_cap < ncells or self._np != n_mass:
            self.cell_start = torch.empty(ncells + 1, dtype=torch.int32,
                                          device=self.dev)
            self.order = torch.empty(max(n_mass, 1), dtype=torch.int32,
                                     device=self.dev)
            self.packed = torch.empty((max(n_mass, 1), 4), dtype=torch.float32,
                                      device=self.dev)
            wb = _lib.load().ssim_cells_work_bytes(n_mass, ncells)
            self.cwork = D.workspace(wb, self.dev)
            self._cells_cap, self._np = ncells, n_mass
        gn = _lib.int_array(nc)
        org = _lib.double_array(lo)
        st = D.stream(self.dev)
        _lib.call("ssim_cells_build", n_mass, *[_lib.ptr(c) for c in x],
                  _lib.ptr(m), P._addr(org), cell, P._addr(gn),
                  _lib.ptr(self.cell_start), _lib.ptr(self.order),
                  _lib.ptr(self.packed), _lib.ptr(self.cwork),
                  self.cwork.numel(), st)
        nt = x[0].numel()
        _lib.call("ssim_pp_short", nt, *[_lib.ptr(c) for c in x], P._addr(org),
                  cell, P._addr(gn), _lib.ptr(self.cell_start),
                  _lib.ptr(self.packed), self.G, self.eps2, rs, rcut,
                  *[_lib.ptr(c) for c in a], st)
        return a

    _np = -1

    def accelerations(self, x, n_mass, m, a):
        """a = mesh + near-field acceleration of all bodies in x (n_mass
        masses first, tracers after them)."""
        self.mesh_accel(x, n_mass, m, a)
        self.near_accel(x, n_mass, m, a)
        return a


def relative_errors(a, ref):
    """Per-body |a - ref| / |ref| for lists of 3 host arrays."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    return np.linalg.norm(a - ref, axis=0) / np.linalg.norm(ref, axis=0)
