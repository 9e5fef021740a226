"""Split particle-mesh force for large self-gravitating systems (3-D).

a = a_mesh + a_near (+ a_learned, see learned.py):

* a_mesh — isolated PM on an n^3 mesh zero-padded to (2n)^3: CIC or TSC
  deposit of the masses, convolution with the long-range Plummer force
  F(d) (1 - g(|d|)), g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2),
  whose spectrum is divided by the squared assignment window, gather with
  the same assignment at every body (masses and tracers).  TSC's sinc^3
  window aliases far less than CIC's sinc^2: at the 250k-mass cube's mesh
  spacing the median force error drops about five-fold.  The kernel is odd, so its spectrum is kept as
  its imaginary part, and the convolution runs as pruned FFT passes that
  never touch the padding (ssim_pm_conv_isolated3);
* a_near — the complementary F(d) g(|d|) summed directly over sources
  within rcut through a cell list.  Beyond rcut the mesh kernel is the
  full F(d), so no part of a pair force is dropped (a hard cut would leave
  out F g(rcut) ~ 0.6 % of the pair force at 5 rs, always attractive, a
  bias that shows up in the energy of a contracting system).

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
    freeze     keep the mesh placement while every body stays at least one
               cell inside it and the bodies still span 3/4 of the usable
               mesh (the force field is then a fixed function of the
               positions between re-placements, which keeps the method's
               energy error down)
    assign     "cic" (the reference's assignment) or "tsc"
    """

    def __init__(self, dev, mesh_n: int = 128, eps2: float = 1e-6,
                 G: float = 1.0, rs_cells: float = 1.25, rcut_rs: float = 6.0,
                 margin: float = 2.0, prec: int = P.F64, freeze: bool = False,
                 assign: str = "cic"):
        if assign not in ("cic", "tsc"):
            raise ValueError(f"assign must be 'cic' or 'tsc', not {assign!r}")
        import torch
        self.dev, self.n, self.eps2, self.G = dev, int(mesh_n), float(eps2), float(G)
        self.rs_cells, self.rcut_rs = float(rs_cells), float(rcut_rs)
        self.margin, self.prec, self.freeze = float(margin), prec, bool(freeze)
        self.assign = assign
        self.assign_order = 2 if assign == "cic" else 3
        self.Pn = 2 * self.n
        self.plan = P.FFTPlan.get([self.Pn] * 3, prec, dev)
        dt = P._mesh_dtype(prec)
        self.rho = torch.zeros([self.n] * 3, dtype=dt, device=dev)
        self.spec = self.plan.new_complex()
        self.prod = self.plan.new_complex()
        self.fields = [torch.empty([self.n] * 3, dtype=dt, device=dev)
                       for _ in range(3)]
        self.kim = None
        self.fused = True     # False: pad + rfftn + product + irfftn + crop
        # TSC on an FP32 mesh gathers the three fields as one float4 mesh
        self.f4 = (torch.empty((self.n ** 3, 4), dtype=torch.float32, device=dev)
                   if assign == "tsc" and prec == P.F32 else None)
        self.level = None
        self.h = None
        self.origin = None
        self._cells_cap = 0
        # near-field cells of TSC runs: blocks of block^3 mesh nodes, edge
        # block * h >= rcut = rcut_rs * rs_cells * h
        self.block = max(1, math.ceil(self.rcut_rs * self.rs_cells - 1e-9))
        # TSC deposit from the block bins by LDS tiles (block edge 5); False:
        # the node-bin gather (ssim_tsc_deposit_blocked)
        self.tiled_deposit = True

    # -- grid ---------------------------------------------------------------
    def _place(self, box):
        """Mesh origin and ladder cell size covering the box [(lo, hi)]."""
        span = max(hi - lo for lo, hi in box)
        if self.freeze and self.level is not None:
            h, o = self.h, self.origin
            usable = (self.n - 1 - 2 * self.margin) * h
            if span >= 0.75 * usable and all(
                    lo >= o[c] + h and hi <= o[c] + (self.n - 2) * h
                    for c, (lo, hi) in enumerate(box)):
                return self.level, h, o, box
        need = max(span, 1e-30) / (self.n - 1 - 2 * self.margin)
        level = math.ceil(64.0 * math.log2(need))
        h = 2.0 ** (level / 64.0)
        origin = [0.5 * (lo + hi) - 0.5 * (self.n - 1) * h for lo, hi in box]
        return level, h, origin, box

    def _rebuild_kernels(self, h):
        """Imaginary spectra of the 3 deconvolved long-range kernels."""
        import torch
        rs = self.rs_cells * h
        st = D.stream(self.dev)
        real = torch.empty([self.Pn] * 3, dtype=self.rho.dtype, device=self.dev)
        nc = self.spec.numel() // 2
        self.kim = []
        for c in range(3):
            _lib.call("ssim_pm_force_kernel", 3, self.n, c, h, self.eps2,
                      self.G, rs, self.rcut_rs * rs, self.prec, _lib.ptr(real), st)
            self.plan.r2c(real, self.prod)
            _lib.call("ssim_pm_deconvolve", 3, self.Pn, self.prec,
                      self.assign_order, _lib.ptr(self.prod), st)
            kim = torch.empty(nc, dtype=self.rho.dtype, device=self.dev)
            _lib.call("ssim_spectrum_imag", nc, self.prec, _lib.ptr(self.prod),
                      _lib.ptr(kim), st)
            self.kim.append(kim)

    # -- force --------------------------------------------------------------
    def place(self, x):
        """Mesh placement for the bodies x (rebuilds the kernel spectra when
        the cell size changes)."""
        return self.place_box(P.minmax(x))

    def place_box(self, box):
        """Mesh placement for bodies inside box [(lo, hi)] x 3; returns True
        when the placement (cell size or origin) changed."""
        level, h, origin, box = self._place(box)
        changed = level != self.level or origin != self.origin
        if level != self.level:
            self._rebuild_kernels(h)
            self.level = level
        self.h, self.origin, self.box = h, origin, box
        return changed

    def placement(self):
        """(ladder level, origin) of the current mesh placement."""
        return self.level, tuple(self.origin) if self.origin is not None else None

    def visit_order(self):
        """Cell-sorted body order of the last bin_bodies (None before)."""
        return self.torder if self._grid is not None else None

    def deposit(self, x, n_mass, m, binned=False):
        """rho = the mass mesh.  TSC: `binned` = the masses were binned
        (mass_bins) for these positions."""
        src = [c[:n_mass] for c in x]
        if self.assign == "cic":
            P.deposit(src, m[:n_mass], [self.n] * 3, self.origin, self.h, False,
                      self.prec, mesh=self.rho)
        else:
            if not binned:
                self.mass_bins(x, n_mass, m)
            org = _lib.double_array(self.origin)
            if self.block == 5 and self.tiled_deposit:
                _lib.call("ssim_tsc_deposit_tiles", self.n, P._addr(org), self.h,
                          self.block, _lib.ptr(self.cell_start), _lib.ptr(self.body),
                          self.mass_bound(m[:n_mass]), self.prec, _lib.ptr(self.rho),
                          D.stream(self.dev))
            else:
                _lib.call("ssim_tsc_deposit_blocked", self.n, P._addr(org), self.h,
                          self.block, _lib.ptr(self.bin_start), _lib.ptr(self.body),
                          self.prec, _lib.ptr(self.rho), D.stream(self.dev))

    def set_mass_bound(self, m):
        """Fix the bound sum |m| of the tiled deposit's fixed point (host
        sync: call it outside graph capture, e.g. when the masses are set)."""
        self._mass_key = (m.data_ptr(), m.numel(), m._version)
        self._mass_bound = float(m.abs().sum()) * (1.0 + 1e-9) + 1e-300

    def mass_bound(self, m):
        """sum |m| (slightly enlarged), cached while the mass tensor is
        unchanged; computed here only outside graph capture."""
        import torch
        key = (m.data_ptr(), m.numel(), m._version)
        if getattr(self, "_mass_key", None) != key:
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError("mass bound unknown inside graph capture: call "
                                   "set_mass_bound() first")
            self.set_mass_bound(m)
        return self._mass_bound

    def mesh_fields(self, x, n_mass, m, binned=False):
        """Deposit the masses and convolve: the acceleration meshes."""
        self.deposit(x, n_mass, m, binned)
        self.convolve(interleaved=self.f4 is not None)

    def can_fold_correction(self):
        """True when the gather can add a learned correction itself (TSC
        with the interleaved FP32 fields)."""
        return self.assign == "tsc" and self.f4 is not None

    def mesh_gather(self, x, a, order=None, accumulate=False, corr=None, box=None):
        """a (+)= the meshes interpolated at the bodies x; `corr` (a dict of
        out, scale, relative, vel, n_proj) folds a learned correction in.
        `box` (a dict of vel, half, full, out, work; can_fold_correction()
        only): the gather also writes the box of the positions the next
        kicks and drift give (ssim_tsc_gather_box) to out."""
        if box is not None:
            if not self.can_fold_correction():
                raise ValueError("the box folds into the TSC float4 gather only")
            org = _lib.double_array(self.origin)
            c = corr or {}
            _lib.call("ssim_tsc_gather_box", x[0].numel(), _lib.ptr(order),
                      *[_lib.ptr(q) for q in x], self.n, P._addr(org), self.h,
                      _lib.ptr(self.f4), *[_lib.ptr(q) for q in a], int(accumulate),
                      _lib.ptr(c.get("out")), float(c.get("scale", 0.0)),
                      int(c.get("relative", 0)), *[_lib.ptr(q) for q in box["vel"]],
                      int(c.get("n_proj", 0)), float(box["half"]), float(box["full"]),
                      box["out"], _lib.ptr(box["work"]), box["work"].numel(),
                      D.stream(self.dev))
            return a
        if corr is not None:
            if not self.can_fold_correction():
                raise ValueError("the correction folds into the TSC float4 gather only")
            org = _lib.double_array(self.origin)
            v = corr.get("vel") or [None, None, None]
            _lib.call("ssim_tsc_gather_corrected", x[0].numel(), _lib.ptr(order),
                      *[_lib.ptr(c) for c in x], self.n, P._addr(org), self.h,
                      _lib.ptr(self.f4), *[_lib.ptr(c) for c in a], int(accumulate),
                      _lib.ptr(corr["out"]), float(corr["scale"]),
                      int(corr["relative"]), *[_lib.ptr(c) for c in v],
                      int(corr.get("n_proj", 0)), D.stream(self.dev))
            return a
        if self.assign == "cic":
            if accumulate:
                raise ValueError("the CIC gather overwrites")
            P.gather(self.fields, x, [self.n] * 3, self.origin, self.h, False,
                     self.prec, outs=a)
            return a
        org = _lib.double_array(self.origin)
        fs = ([self.f4, None, None] if self.f4 is not None else self.fields)
        _lib.call("ssim_tsc_gather", x[0].numel(), _lib.ptr(order),
                  *[_lib.ptr(c) for c in x], self.n, P._addr(org), self.h,
                  self.prec, *[_lib.ptr(f) for f in fs],
                  *[_lib.ptr(c) for c in a], int(accumulate), D.stream(self.dev))
        return a

    def mesh_accel(self, x, n_mass, m, a, order=None, placed=False):
        """a = long-range mesh acceleration at all bodies (overwrites a);
        `order`: the order the gather visits the bodies in."""
        if not placed:
            self.place(x)
        self.mesh_fields(x, n_mass, m)
        return self.mesh_gather(x, a, order)

    def mass_bins(self, x, n_mass, m):
        """TSC runs: one stable binning of the masses (ssim_split_bin) by
        the block of nodes their stencils are centred on; the blocks are the
        near field's cells and the tiled deposit's input.  Node bins inside
        the blocks (bin_start) only for the node-bin deposit
        (tiled_deposit False, or a block edge other than 5)."""
        import torch
        nm = n_mass
        if getattr(self, "_bin_n", -1) != nm:
            self._all_key = None
            nb = -(-self.n // self.block)
            K = nb ** 3 * self.block ** 3
            i32 = dict(dtype=torch.int32, device=self.dev)
            self.bin_start = torch.empty(K + 1, **i32)
            self.cell_start = torch.empty(nb ** 3 + 1, **i32)
            self.order = torch.empty(max(nm, 1), **i32)
            self.packed = torch.empty((max(nm, 1), 4), dtype=torch.float32,
                                      device=self.dev)
            self.near = torch.empty((max(nm, 1), 4), dtype=torch.float32,
                                    device=self.dev)
            self.body = torch.empty((max(nm, 1), 4), dtype=torch.float64,
                                    device=self.dev)
            self.bwork = D.workspace(
                _lib.load().ssim_split_bin_work_bytes(nm, self.n, self.block), self.dev)
            # allocated with the rest so that switching near_hilo on never
            # allocates inside a captured step
            self.near_lo = torch.empty((max(nm, 1), 4), dtype=torch.float32,
                                       device=self.dev)
            self._bin_n = nm
        lo = _lib.ptr(self.near_lo) if (self.near_hilo or self.near_bits) else None
        org = _lib.double_array(self.origin)
        # node bins only for the node-bin gather deposit
        bins = None if self.block == 5 and self.tiled_deposit else _lib.ptr(self.bin_start)
        _lib.call("ssim_split_bin", nm, *[_lib.ptr(c) for c in x], _lib.ptr(m), self.n,
                  P._addr(org), self.h, self.block, bins,
                  _lib.ptr(self.cell_start), _lib.ptr(self.order),
                  _lib.ptr(self.packed), _lib.ptr(self.body), _lib.ptr(self.near), lo,
                  _lib.ptr(self.bwork), self.bwork.numel(), D.stream(self.dev))

    def convolve(self, interleaved=False):
        """fields = long-range acceleration meshes of self.rho (or, with
        `interleaved`, self.f4 = the three as float4 [n^3])."""
        import torch
        st = D.stream(self.dev)
        if self.fused:
            outs = [self.f4, None, None] if interleaved else self.fields
            _lib.call("ssim_pm_conv_isolated3", self.n, self.prec,
                      _lib.ptr(self.rho), *[_lib.ptr(k) for k in self.kim],
                      *[_lib.ptr(f) for f in outs], _lib.ptr(self.spec),
                      _lib.ptr(self.prod), _lib.ptr(self.plan.work),
                      self.plan.work_bytes, st)
            return self.f4 if interleaved else self.fields
        pad = torch.zeros([self.Pn] * 3, dtype=self.rho.dtype, device=self.dev)
        full = torch.empty_like(pad)
        _lib.call("ssim_pm_pad", 3, self.n, self.prec, _lib.ptr(self.rho),
                  _lib.ptr(pad), 0, st)
        self.plan.r2c(pad, self.spec)
        kc = torch.zeros_like(self.prod)
        for c in range(3):
            kc[..., 1] = self.kim[c].view(kc.shape[:-1])
            _lib.call("ssim_spectrum_mul", self.spec.numel() // 2, self.prec,
                      _lib.ptr(self.spec), _lib.ptr(kc), _lib.ptr(self.prod), st)
            self.plan.c2r(self.prod, full)
            _lib.call("ssim_pm_pad", 3, self.n, self.prec, _lib.ptr(full),
                      _lib.ptr(self.fields[c]), 1, st)
        return self.fields

    def bin_bodies(self, x, n_mass, m):
        """Cell lists of the sources (stable, packed) and of all targets on
        a grid of cells >= rcut over the mesh (a function of the placement
        only, so the launches of a step repeat while the mesh is kept);
        returns the grid."""
        import torch
        rs = self.rs_cells * self.h
        rcut = self.rcut_rs * rs
        if self.assign == "tsc":
            return self.tsc_bins(x, n_mass, m)
        lo = list(self.origin)
        extent = (self.n - 1) * self.h
        k = max(1, int(extent // rcut))
        nc = [k, k, k]
        cell = max(extent / k, rcut)
        ncells = nc[0] * nc[1] * nc[2]
        nt = x[0].numel()
        if self._cells_cap < ncells or self._np != n_mass or self._nt != nt:
            i32 = dict(dtype=torch.int32, device=self.dev)
            self.cell_start = torch.empty(ncells + 1, **i32)
            self.order = torch.empty(max(n_mass, 1), **i32)
            self.packed = torch.empty((max(n_mass, 1), 4), dtype=torch.float32,
                                      device=self.dev)
            self.tcell_start = torch.empty(ncells + 1, **i32)
            self.torder = torch.empty(max(nt, 1), **i32)
            lib = _lib.load()
            wb = max(lib.ssim_cells_work_bytes(n_mass, ncells),
                     lib.ssim_cells_bin_work_bytes(nt, ncells),
                     lib.ssim_pp_work_bytes(nt, ncells))
            self.cwork = D.workspace(wb, self.dev)
            self._cells_cap, self._np, self._nt = ncells, n_mass, nt
        gn = _lib.int_array(nc)
        org = _lib.double_array(lo)
        st = D.stream(self.dev)
        _lib.call("ssim_cells_build", n_mass, *[_lib.ptr(c) for c in x],
                  _lib.ptr(m), P._addr(org), cell, P._addr(gn),
                  _lib.ptr(self.cell_start), _lib.ptr(self.order),
                  _lib.ptr(self.packed), _lib.ptr(self.cwork),
                  self.cwork.numel(), st)
        _lib.call("ssim_cells_bin", nt, *[_lib.ptr(c) for c in x],
                  P._addr(org), cell, P._addr(gn), _lib.ptr(self.tcell_start),
                  _lib.ptr(self.torder), _lib.ptr(self.cwork),
                  self.cwork.numel(), st)
        self._grid = (gn, org, cell, rs, rcut)
        return self._grid

    def tsc_bins(self, x, n_mass, m):
        """TSC runs: the mass bins and the target lists, by one binning of
        all bodies (ssim_split_bin_all) when the deposit needs no node bins,
        else by mass_bins and bin_targets; returns the grid."""
        if not (self.unified_bins and self.block == 5 and self.tiled_deposit):
            self.mass_bins(x, n_mass, m)
            return self.bin_targets(x)
        import torch
        nt = x[0].numel()
        nb = -(-self.n // self.block)
        ncells = nb ** 3
        if getattr(self, "_all_key", None) != (n_mass, nt, ncells):
            i32 = dict(dtype=torch.int32, device=self.dev)
            f32 = dict(dtype=torch.float32, device=self.dev)
            nm = max(n_mass, 1)
            self.cell_start = torch.empty(ncells + 1, **i32)
            self.order = torch.empty(nm, **i32)
            self.packed = torch.empty((nm, 4), **f32)
            self.near = torch.empty((nm, 4), **f32)
            self.near_lo = torch.empty((nm, 4), **f32)
            self.body = torch.empty((nm, 4), dtype=torch.float64, device=self.dev)
            self.tcell_start = torch.empty(ncells + 1, **i32)
            self.torder = torch.empty(max(nt, 1), **i32)
            lib = _lib.load()
            self.awork = D.workspace(
                lib.ssim_split_bin_all_work_bytes(nt, self.n, self.block), self.dev)
            self.cwork = D.workspace(lib.ssim_pp_work_bytes(nt, ncells), self.dev)
            self._all_key = (n_mass, nt, ncells)
            self._bin_n = -1          # mass_bins / bin_targets reallocate
            self._nt = -1
        org = _lib.double_array(self.origin)
        lo = _lib.ptr(self.near_lo) if (self.near_hilo or self.near_bits) else None
        _lib.call("ssim_split_bin_all", nt, n_mass, *[_lib.ptr(c) for c in x], _lib.ptr(m),
                  self.n, P._addr(org), self.h, self.block, _lib.ptr(self.cell_start),
                  _lib.ptr(self.order), _lib.ptr(self.packed), _lib.ptr(self.body),
                  _lib.ptr(self.near), lo, _lib.ptr(self.tcell_start),
                  _lib.ptr(self.torder), _lib.ptr(self.awork), self.awork.numel(),
                  D.stream(self.dev))
        rs = self.rs_cells * self.h
        gn = _lib.int_array([nb, nb, nb])
        corner = _lib.double_array([o - 0.5 * self.h for o in self.origin])
        self._grid = (gn, corner, self.block * self.h, rs, self.rcut_rs * rs)
        return self._grid

    def bin_targets(self, x):
        """TSC runs: cell list of all targets on the grid of node blocks
        (origin o - h/2, edge block * h); returns the grid."""
        import torch
        rs = self.rs_cells * self.h
        rcut = self.rcut_rs * rs
        nb = -(-self.n // self.block)
        nc = [nb, nb, nb]
        ncells = nb ** 3
        cell = self.block * self.h
        lo = [o - 0.5 * self.h for o in self.origin]
        nt = x[0].numel()
        if self._nt != nt or self._cells_cap != ncells:
            self._all_key = None
            i32 = dict(dtype=torch.int32, device=self.dev)
            self.tcell_start = torch.empty(ncells + 1, **i32)
            self.torder = torch.empty(max(nt, 1), **i32)
            lib = _lib.load()
            wb = max(lib.ssim_cells_bin_work_bytes(nt, ncells),
                     lib.ssim_pp_work_bytes(nt, ncells))
            self.cwork = D.workspace(wb, self.dev)
            self._cells_cap, self._nt = ncells, nt
        gn = _lib.int_array(nc)
        org = _lib.double_array(lo)
        _lib.call("ssim_cells_bin", nt, *[_lib.ptr(c) for c in x],
                  P._addr(org), cell, P._addr(gn), _lib.ptr(self.tcell_start),
                  _lib.ptr(self.torder), _lib.ptr(self.cwork),
                  self.cwork.numel(), D.stream(self.dev))
        self._grid = (gn, org, cell, rs, rcut)
        return self._grid

    def near_sum(self, x, a, accumulate=True):
        """a += (or, accumulate=False, a =) the near-field term over the
        lists of the last bin_bodies.  TSC runs (masses binned by
        mass_bins) use the block-relative kernel (ssim_pp_near) unless
        near_precise is False, and its hi/lo form (ssim_pp_near_hilo) when
        near_hilo is set."""
        gn, org, cell, rs, rcut = self._grid
        if self.assign == "tsc" and self.near_precise and self.near_bits:
            _lib.call("ssim_pp_near_bits", x[0].numel(),
                      *[_lib.ptr(c) for c in x], P._addr(org), cell, P._addr(gn),
                      _lib.ptr(self.cell_start), _lib.ptr(self.near),
                      _lib.ptr(self.near_lo), _lib.ptr(self.tcell_start),
                      _lib.ptr(self.torder), self.G, self.eps2, rs, rcut,
                      *[_lib.ptr(c) for c in a], int(accumulate), int(self.near_hilo),
                      _lib.ptr(self.cwork), self.cwork.numel(), D.stream(self.dev))
            return a
        if self.assign == "tsc" and self.near_precise:
            hl = self.near_hilo
            _lib.call("ssim_pp_near_hilo" if hl else "ssim_pp_near", x[0].numel(),
                      *[_lib.ptr(c) for c in x], P._addr(org), cell, P._addr(gn),
                      _lib.ptr(self.cell_start), _lib.ptr(self.near),
                      *([_lib.ptr(self.near_lo)] if hl else []),
                      _lib.ptr(self.tcell_start),
                      _lib.ptr(self.torder), self.G, self.eps2, rs, rcut,
                      *[_lib.ptr(c) for c in a], int(accumulate), _lib.ptr(self.cwork),
                      self.cwork.numel(), D.stream(self.dev))
            return a
        _lib.call("ssim_pp_short", x[0].numel(), *[_lib.ptr(c) for c in x],
                  P._addr(org), cell, P._addr(gn), _lib.ptr(self.cell_start),
                  _lib.ptr(self.packed), _lib.ptr(self.tcell_start),
                  _lib.ptr(self.torder), self.G, self.eps2, rs, rcut,
                  *[_lib.ptr(c) for c in a], int(accumulate), _lib.ptr(self.cwork),
                  self.cwork.numel(), D.stream(self.dev))
        return a

    def near_pair_counts(self, x):
        """(pairs tested, pairs inside rcut) of the last bin_bodies."""
        import torch
        gn, org, cell, rs, rcut = self._grid
        counts = torch.zeros(2, dtype=torch.int64, device=self.dev)
        _lib.call("ssim_pp_count", x[0].numel(), *[_lib.ptr(c) for c in x],
                  P._addr(org), cell, P._addr(gn), _lib.ptr(self.cell_start),
                  _lib.ptr(self.packed), _lib.ptr(self.tcell_start),
                  _lib.ptr(self.torder), self.eps2, rs, rcut, _lib.ptr(counts),
                  _lib.ptr(self.cwork), self.cwork.numel(), D.stream(self.dev))
        c = counts.cpu().numpy()
        return int(c[0]), int(c[1])

    def near_accel(self, x, n_mass, m, a):
        """a += near-field term (cell list of the masses)."""
        self.bin_bodies(x, n_mass, m)
        return self.near_sum(x, a)

    _np = -1
    _nt = -1
    _grid = None

    def accelerations(self, x, n_mass, m, a, placed=False, correction=None, box=None):
        """a = mesh + near-field acceleration of all bodies in x (n_mass
        masses first, tracers after them); `placed`: the caller has placed
        the mesh for x already (place_box), so nothing waits for the host.
        `correction` (can_fold_correction() only): a callable, launched
        once the mass mesh is deposited, that enqueues a learned correction
        on the current stream and returns mesh_gather's `corr` dict; the
        gather then adds it.  A pair (prepare, finish) splits it: prepare
        (the features, which read the mass mesh) runs right after the
        deposit on the current stream, finish (the network) after it on the
        side stream behind the near field and returns the dict.  `box`
        (TSC with interleaved FP32 fields): the gather also writes the box
        of the next drift (mesh_gather).

        With TSC the work runs on two streams: the current one bins all
        bodies once (the deposit's mass bins, the near field's sources and
        targets: tsc_bins); then the side stream adds the near field (bound
        by the vector ALU) and, once the mass mesh is deposited, runs the
        correction, while the current one deposits and convolves (bound by
        memory and latency); the near field is stored into a, and the
        current stream waits for the side stream and adds the mesh gather
        (in index order: each body's sum and the box do not depend on the
        visit order, and index order reads the bodies' arrays contiguously;
        measured as fast as the cell order on randomly ordered bodies and
        faster in the step on the lattice-ordered cube).  With CIC
        everything runs in order on one stream."""
        import torch
        if not placed:
            self.place(x)
        if (correction is not None or box is not None) and not self.can_fold_correction():
            raise ValueError("correction / box need TSC with interleaved FP32 fields")
        if self.assign == "tsc" and not self.overlap:
            self.tsc_bins(x, n_mass, m)
            self.near_sum(x, a, accumulate=False)
            self.deposit(x, n_mass, m, binned=True)
            if isinstance(correction, tuple):
                correction[0]()
                corr = correction[1]()
            else:
                corr = correction() if correction is not None else None
            self.convolve(interleaved=self.f4 is not None)
            self.mesh_gather(x, a, order=None, accumulate=True, corr=corr, box=box)
            return a
        if self.assign != "tsc":
            self.bin_bodies(x, n_mass, m)
            self.mesh_accel(x, n_mass, m, a, order=self.torder, placed=True)
            self.near_sum(x, a)
            return a
        main = torch.cuda.current_stream(self.dev)
        start = torch.cuda.Event()
        start.record(main)
        self.run_segments(self.tsc_segments(x, n_mass, m, a, correction, box),
                          {"start": start})
        return a

    def side_stream(self):
        import torch
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.dev)  # default priority
        return self._side

    def tsc_segments(self, x, n_mass, m, a, correction=None, box=None):
        """The overlapped TSC force as ordered segments (stream, fn, waits,
        signal): stream "main" (the current one) or "side"; before a
        segment runs, its stream waits for the named events, and afterwards
        records `signal`.  The caller records "start" on main before the
        first segment.  The binning comes first (one binning of all
        bodies, tsc_bins; without unified bins the mass binning on main and
        the target binning on the side stream, side_first putting the
        target binning first, the order a captured graph needs); the near
        field runs on the side stream beside the deposit and the
        convolution, and so does the correction once the mass mesh exists;
        the gather waits for the side stream."""
        held = {}

        split = isinstance(correction, tuple)

        def corr():
            held["corr"] = correction[1]() if split else correction()

        if self.unified_bins and self.block == 5 and self.tiled_deposit:
            segs = [("main", lambda: self.tsc_bins(x, n_mass, m), (), "bins")]
        else:
            mass = ("main", lambda: self.mass_bins(x, n_mass, m), (), "bins")
            targets = ("side", lambda: self.bin_targets(x), ("start",), None)
            segs = [targets, mass] if self.side_first else [mass, targets]
        segs += [("side", lambda: self.near_sum(x, a, accumulate=False), ("bins",), None),
                 ("main", lambda: self.deposit(x, n_mass, m, binned=True), (), "deposited")]
        if correction is not None and split:
            # the features behind the deposit on the current stream, the
            # network behind the near field on the side stream
            segs.append(("main", correction[0], (), "prepared"))
            segs.append(("side", corr, ("prepared",), None))
        elif correction is not None:
            segs.append(("side", corr, ("deposited",), None))
        last = segs[-1] if correction is not None else \
            next(sg for sg in segs if sg[0] == "side" and sg[2] == ("bins",))
        segs[segs.index(last)] = last[:3] + ("side_done",)
        segs += [("main", lambda: self.convolve(interleaved=self.f4 is not None), (), None),
                 ("main", lambda: self.mesh_gather(x, a, order=None, accumulate=True,
                                                   corr=held.get("corr"), box=box),
                  ("side_done",), None)]
        return segs

    def run_segments(self, segs, events):
        """Launch segments (see tsc_segments) on the current and the side
        stream; `events` maps names to torch events (created as needed)."""
        import torch
        main = torch.cuda.current_stream(self.dev)
        side = self.side_stream()
        for sname, fn, waits, sig in segs:
            st = main if sname == "main" else side
            for w in waits:
                st.wait_event(events[w])
            with torch.cuda.stream(st):
                fn()
            if sig:
                if sig not in events:
                    events[sig] = torch.cuda.Event()
                events[sig].record(st)

    _side = None
    near_precise = True   # TSC: the near field on block-relative coordinates
    near_bits = True      # TSC: the near field as MFMA filter + sparse sum
                          # (ssim_pp_near_bits); False: the cell walk
                          # (ssim_pp_near / ssim_pp_near_hilo)
    near_hilo = False     # TSC: FP64 positions (hi + lo) and pair terms in
                          # the near field (p99 force error <= 1e-6, about
                          # 1.35x its time)
    overlap = True        # TSC: near field on a second stream
    unified_bins = True   # TSC: masses and targets from one binning
    side_first = False    # enqueue the target binning before the mass binning


def relative_errors(a, ref):
    """Per-body |a - ref| / |ref| for lists of 3 host arrays."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    return np.linalg.norm(a - ref, axis=0) / np.linalg.norm(ref, axis=0)
