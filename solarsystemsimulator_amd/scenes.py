"""Seeded synthetic scenes of the benchmark configurations (G = 1 units).

* `perturbed_cube` — equal-mass bodies in the unit cube: one uniformly
  jittered point per cell of an nx*ny*nz lattice, displaced by a Zel'dovich
  field of a Gaussian random density with P(k) ~ k^-2, scaled so the rms
  density contrast is `amp`.  Cold start.  `periodic=True` wraps positions
  into [0, 1).  Used for the 65,536-body all-pairs case (64x32x32, seed 2),
  the 262,144-body particle-mesh case (64^3, seed 3, periodic) and the
  250k-mass + 500k-tracer case (100x50x50 seed 5 / 100x100x50 seed 55).
"""
from __future__ import annotations

import numpy as np


def _zeldovich(shape, rng, amp):
    """Displacement field psi (3, *shape) in lattice units with rms div = amp."""
    kx = np.fft.fftfreq(shape[0]) * 2 * np.pi
    ky = np.fft.fftfreq(shape[1]) * 2 * np.pi
    kz = np.fft.rfftfreq(shape[2]) * 2 * np.pi
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    white = np.fft.rfftn(rng.standard_normal(shape))
    delta_k = white / np.sqrt(k2)          # P(k) ~ k^-2
    delta_k[0, 0, 0] = 0.0
    delta = np.fft.irfftn(delta_k, s=shape, axes=(0, 1, 2))
    delta_k *= amp / delta.std()
    # div psi = -delta  =>  psi_k = i k delta_k / k^2
    psi = [np.fft.irfftn(1j * K * delta_k / k2, s=shape, axes=(0, 1, 2))
           for K in (KX, KY, KZ)]
    return np.stack(psi)


def perturbed_cube(nx: int, ny: int, nz: int, seed: int, amp: float = 0.05,
                   periodic: bool = False):
    """(x, y, z) float64 arrays of nx*ny*nz bodies in the unit cube."""
    rng = np.random.default_rng(seed)
    shape = (nx, ny, nz)
    idx = np.indices(shape).reshape(3, -1).astype(np.float64)
    jit = rng.uniform(0.0, 1.0, idx.shape)
    psi = _zeldovich(shape, rng, amp).reshape(3, -1)
    scale = np.array([nx, ny, nz], np.float64)[:, None]
    pos = (idx + jit + psi) / scale
    if periodic:
        pos = np.mod(pos, 1.0)
    return pos[0].copy(), pos[1].copy(), pos[2].copy()


def selfgrav_cube(n_shape=(64, 32, 32), seed=2, amp=0.05):
    """All-pairs case: equal masses (total 1) on a perturbed lattice."""
    x, y, z = perturbed_cube(*n_shape, seed=seed, amp=amp)
    m = np.full(x.size, 1.0 / x.size)
    return (x, y, z), m


def learned_scene(mass_shape=(100, 50, 50), tracer_shape=(100, 100, 50),
                  seed=5, tracer_seed=55, amp=0.05):
    """250k masses + 500k tracers (default sizes), cold start."""
    pos = perturbed_cube(*mass_shape, seed=seed, amp=amp)
    tpos = perturbed_cube(*tracer_shape, seed=tracer_seed, amp=amp)
    m = np.full(pos[0].size, 1.0 / pos[0].size)
    return pos, m, tpos
