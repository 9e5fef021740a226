"""Learned correction of the split particle-mesh force (device side).

A 32-64-64-64-3 SiLU network (libssim, BF16 matrix cores, FP32
accumulation) maps local features of a body — the 27 CIC node masses of
the 3x3x3 mesh stencil around its nearest node and its offset from that
node, and whether the body is a mass — to the residual between the exact
direct sum and the split mesh + near-field force, in units of 1e-3 of the
split force's magnitude at that body (targets clipped at 3 %; an absolute
mode in units of G m_body / h^2 is kept).  It is trained on
snapshots whose exact forces the FP64 all-pairs kernel computes, with the
project's own forward/backward/Adam kernels, and at run time its output is
added to the split force; the direct sum stays available as the fallback
and as the judge of the error.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as D

NIN = 32


def _torch():
    import torch
    return torch


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform fan-in initialisation (float32 block, output rows 3..15 = 0)."""
    rng = np.random.default_rng(seed)
    blocks = []
    for rows, cols, scale in ((64, 32, 1.0), (64, 64, 1.0), (64, 64, 1.0),
                              (16, 64, 0.5)):
        lim = scale * np.sqrt(6.0 / cols)
        w = rng.uniform(-lim, lim, (rows, cols)).astype(np.float32)
        if rows == 16:
            w[3:] = 0
        blocks += [w.ravel(), np.zeros(rows, np.float32)]
    p = np.concatenate(blocks)
    assert p.size == _lib.load().ssim_mlp_param_count()
    return p


class LearnedCorrection:
    """Parameters, Adam state and the libssim calls of the network."""

    def __init__(self, dev, params: np.ndarray | None = None, seed: int = 0):
        torch = _torch()
        self.dev = dev
        self.nparam = _lib.load().ssim_mlp_param_count()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.as_tensor(p).to(dev)
        self.wbf = torch.empty(self.nparam, dtype=torch.int16, device=dev)
        self.m = torch.zeros(self.nparam, dtype=torch.float32, device=dev)
        self.v = torch.zeros(self.nparam, dtype=torch.float32, device=dev)
        self.grad = torch.zeros(self.nparam, dtype=torch.float32, device=dev)
        self.step = 0
        self.a_unit = 1.0      # set by the caller: G m_body / h^2 of training
        self.mass_unit = 1.0   # m_body
        # output units: relative -> out_scale * |a| of the corrected force
        # (the residual in units of 1e-3 of the force), else a_unit
        self.relative = True
        self.out_scale = 1e-3
        self.clip = 30.0       # training targets clipped to +-3 % of |a|
        self.refresh()

    def refresh(self):
        _lib.call("ssim_mlp_to_bf16", _lib.ptr(self.params), _lib.ptr(self.wbf),
                  D.stream(self.dev))

    def host_params(self) -> np.ndarray:
        return self.params.cpu().numpy().copy()

    # -- features / inference ---------------------------------------------
    def features(self, x, mesh, origin, h, mesh_prec, out=None, n_mass=0,
                 order=None):
        """BF16 feature rows [nb][32] (int16 storage) of the bodies x, the
        first n_mass of which are masses; `order` (int32 device array) is
        the order the threads visit the bodies in (a cell-sorted order reads
        the mesh with locality), None for index order."""
        torch = _torch()
        nb = x[0].numel()
        if out is None:
            out = torch.empty((nb, NIN), dtype=torch.int16, device=self.dev)
        org = _lib.double_array(origin)
        _lib.call("ssim_mlp_features", nb, int(n_mass), _lib.ptr(order),
                  *[_lib.ptr(c) for c in x],
                  mesh.shape[0], ctypes.addressof(org), float(h), mesh_prec,
                  _lib.ptr(mesh), float(self.mass_unit), _lib.ptr(out),
                  D.stream(self.dev))
        return out

    def predict(self, feat, out=None, add_to=None, scale=None, zs=None, hs=None,
                relative=False, vel=None, n_proj=0):
        """Network outputs [nb][3] (FP32) and/or u*out added to add_to, with
        u = scale, or scale * |add_to| per body when `relative`; the first
        n_proj bodies get only the part orthogonal to their velocity `vel`."""
        nb = feat.shape[0]
        s = self.a_unit if scale is None else scale
        a = add_to if add_to is not None else [None, None, None]
        v = vel if vel is not None else [None, None, None]
        _lib.call("ssim_mlp_forward", nb, _lib.ptr(feat), _lib.ptr(self.wbf),
                  _lib.ptr(self.params), float(s), int(relative), _lib.ptr(out),
                  *[_lib.ptr(c) for c in a], *[_lib.ptr(c) for c in v],
                  int(n_proj), _lib.ptr(zs), _lib.ptr(hs), D.stream(self.dev))
        return out

    # -- training ----------------------------------------------------------
    def gradient(self, feat, target):
        """Loss (mean squared error over the batch) and self.grad."""
        torch = _torch()
        nb = feat.shape[0]
        out = torch.empty((nb, 3), dtype=torch.float32, device=self.dev)
        zs = torch.empty((3, nb, 64), dtype=torch.float32, device=self.dev)
        hs = torch.empty((3, nb, 64), dtype=torch.int16, device=self.dev)
        self.predict(feat, out=out, zs=zs, hs=hs)
        gout = torch.empty((nb, 3), dtype=torch.float32, device=self.dev)
        nblk = (nb + 255) // 256
        part = torch.empty(nblk, dtype=torch.float32, device=self.dev)
        _lib.call("ssim_mlp_mse_grad", nb, _lib.ptr(out), _lib.ptr(target),
                  _lib.ptr(gout), _lib.ptr(part), D.stream(self.dev))
        wb = _lib.load().ssim_mlp_train_work_bytes(nb)
        work = D.workspace(wb, self.dev)
        _lib.call("ssim_mlp_backward", nb, _lib.ptr(feat), _lib.ptr(self.wbf),
                  _lib.ptr(zs), _lib.ptr(hs), _lib.ptr(gout), _lib.ptr(self.grad),
                  _lib.ptr(work), wb, D.stream(self.dev))
        return float(part.cpu().numpy().astype(np.float64).sum() / nb), out

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.call("ssim_mlp_adam", _lib.ptr(self.params), _lib.ptr(self.grad),
                  _lib.ptr(self.m), _lib.ptr(self.v), float(lr), float(beta1),
                  float(beta2), float(eps), self.step, _lib.ptr(self.wbf),
                  D.stream(self.dev))

    def train(self, feat, target, epochs=4, batch=8192, lr=2e-3, seed=0,
              lr_decay=0.5, log=None):
        """Minibatch Adam over a pre-shuffled dataset (contiguous batches in
        a per-epoch random order).  Returns the loss history."""
        rng = np.random.default_rng(seed)
        n = feat.shape[0]
        nbat = n // batch
        hist = []
        for ep in range(epochs):
            order = rng.permutation(nbat)
            tot = 0.0
            for k in order:
                sl = slice(int(k) * batch, int(k) * batch + batch)
                loss, _ = self.gradient(feat[sl], target[sl])
                self.adam_step(lr)
                tot += loss
            hist.append(tot / max(nbat, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e} lr {lr:.2e}")
            lr *= lr_decay
        return hist
