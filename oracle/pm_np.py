"""NumPy restatement of the particle-mesh path — TEST INFRASTRUCTURE ONLY.

* `cic_cells`, `cic_deposit`, `cic_gather` — pm_kernel.py:63-107
  (isolated: floor + clip to [0, n-2]; periodic: floor wrapped mod n),
  any dimension, x index slowest.
* `build_force_kernels`, `pm_force_field`, `pm_accelerations`, `grid_fuer`
  — pm_kernel.py:34-163, for 2-D and 3-D.
* `periodic_field` — the periodic Poisson flavour: rfftn, -4 pi G rho_k/k^2,
  irfftn, central difference.
* `periodic_kdk` — KDK / Yoshida sub-steps on `periodic_accelerations`
  (the reference's NBodyPM.step_batch order, pm_kernel.py:254-281, with
  the drift wrapped into the box).
* `energie_impuls` — test_pm.py:103-118.

Pinned by tests/test_pm_oracle.py: mass conservation, brute-force CIC
indices, a brute-force real-space convolution on a tiny mesh, the far-pair
1/r^2 limit and momentum test of test_pm.py, an analytic single Fourier
mode for the Poisson solve, and the cloud-vs-all-pairs error bounds of
test_pm.py.
"""
from __future__ import annotations

import numpy as np

G_AU = 4 * np.pi * np.pi


def cic_cells(pos, origin, h, n, periodic):
    """Integer cells (dim, np) and fractional weights t (dim, np)."""
    cells, ts = [], []
    for c, x in enumerate(pos):
        f = (np.asarray(x, np.float64) - origin[c]) / h
        i = np.floor(f).astype(np.int64)
        if not periodic:
            i = np.clip(i, 0, n[c] - 2)
        ts.append(f - i)
        cells.append(np.mod(i, n[c]) if periodic else i)
    return np.array(cells), np.array(ts)


def _corners(cells, ts, n, periodic):
    """(flat node index, per-axis weights) of each of the 2^dim corners, x
    index varying fastest — the reference's corner order (i,j), (i+1,j),
    (i,j+1), (i+1,j+1) (pm_kernel.py:77-88)."""
    dim = len(n)
    out = []
    for bits in range(2 ** dim):
        flat = np.zeros(cells.shape[1], np.int64)
        ws = []
        for c in range(dim):
            up = (bits >> c) & 1
            idx = cells[c] + up
            if periodic:
                idx = np.mod(idx, n[c])
            flat = flat * n[c] + idx
            ws.append(ts[c] if up else 1.0 - ts[c])
        out.append((flat, ws))
    return out


def _weighted(v, ws):
    """((v * w_x) * w_y) * w_z — the reference's operand order."""
    for w in ws:
        v = v * w
    return v


def cic_deposit(pos, m, n, origin, h, periodic=False):
    cells, ts = cic_cells(pos, origin, h, n, periodic)
    rho = np.zeros(int(np.prod(n)))
    for flat, ws in _corners(cells, ts, n, periodic):
        np.add.at(rho, flat, _weighted(np.asarray(m, np.float64), ws))
    return rho.reshape(n)


def cic_gather(field, pos, n, origin, h, periodic=False):
    cells, ts = cic_cells(pos, origin, h, n, periodic)
    f = np.asarray(field, np.float64).ravel()
    out = None
    for flat, ws in _corners(cells, ts, n, periodic):
        term = _weighted(f[flat], ws)
        out = term if out is None else out + term
    return out


def grid_fuer(pos, grid_n, rand_zellen=3.0):
    spanne = max(max(float(p.max() - p.min()) for p in pos), 1e-30)
    h = spanne / (grid_n - 2 * rand_zellen)
    return [float(p.min()) - rand_zellen * h for p in pos], h


def build_force_kernels(grid_n, h, eps2, G=G_AU, dim=2):
    P = 2 * grid_n
    off = np.arange(P, dtype=np.float64)
    off = np.where(off < grid_n, off, off - P) * h
    grids = np.meshgrid(*([off] * dim), indexing="ij")
    r2 = sum(g * g for g in grids) + eps2
    inv_r3 = 1.0 / (r2 * np.sqrt(r2))
    return tuple(np.fft.rfftn((-G) * g * inv_r3) for g in grids)


def pm_force_field(pos, m, grid_n, origin, h, eps2, G=G_AU, kernels=None):
    dim = len(pos)
    if kernels is None:
        kernels = build_force_kernels(grid_n, h, eps2, G, dim)
    P = 2 * grid_n
    rho = cic_deposit(pos, m, [grid_n] * dim, origin, h, False)
    pad = np.zeros([P] * dim)
    pad[tuple(slice(0, grid_n) for _ in range(dim))] = rho
    fr = np.fft.rfftn(pad)
    sl = tuple(slice(0, grid_n) for _ in range(dim))
    return [np.fft.irfftn(fr * k, s=[P] * dim, axes=list(range(dim)))[sl]
            for k in kernels]


def pm_accelerations(pos, m, grid_n, origin, h, eps2, G=G_AU, kernels=None):
    fields = pm_force_field(pos, m, grid_n, origin, h, eps2, G, kernels)
    n = [grid_n] * len(pos)
    return [cic_gather(f, pos, n, origin, h) for f in fields]


def periodic_field(pos, m, n, box, origin, G=1.0):
    """Acceleration meshes of the periodic Poisson solve (any dim)."""
    dim = len(pos)
    h = box / n
    rho = cic_deposit(pos, m, [n] * dim, origin, h, True) / h ** dim
    rk = np.fft.rfftn(rho)
    ks = [np.fft.fftfreq(n) * n for _ in range(dim - 1)] + [np.fft.rfftfreq(n) * n]
    K = np.meshgrid(*[2 * np.pi * k / (n * h) for k in ks], indexing="ij")
    k2 = sum(k * k for k in K)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(k2 > 0, -4 * np.pi * G / k2, 0.0)
    phi = np.fft.irfftn(rk * g, s=[n] * dim, axes=list(range(dim)))
    return [-(np.roll(phi, -1, axis=c) - np.roll(phi, 1, axis=c)) / (2 * h)
            for c in range(dim)]


def periodic_accelerations(pos, m, targets, n, box, origin, G=1.0):
    fields = periodic_field(pos, m, n, box, origin, G)
    h = box / n
    return [cic_gather(f, targets, [n] * len(pos), origin, h, True) for f in fields]


def periodic_kdk(pos, vel, m, n, box, origin, G, dt, steps, weights=(1.0,)):
    """`steps` KDK steps in the periodic box (each a composition of
    sub-steps of the given weights: v += a w dt/2; x += v w dt wrapped as
    x - box floor((x - o)/box); a = periodic_accelerations; v += a w dt/2).
    Returns (pos, vel) lists of float64 arrays."""
    x = [np.array(c, np.float64) for c in pos]
    v = [np.array(c, np.float64) for c in vel]
    a = periodic_accelerations(x, m, x, n, box, origin, G)
    for _ in range(steps):
        for w in weights:
            full = w * dt
            half = 0.5 * full
            v = [v[c] + a[c] * half for c in range(3)]
            x = [x[c] + v[c] * full for c in range(3)]
            x = [x[c] - box * np.floor((x[c] - origin[c]) / box) for c in range(3)]
            a = periodic_accelerations(x, m, x, n, box, origin, G)
            v = [v[c] + a[c] * half for c in range(3)]
    return x, v


def energie_impuls(x, y, vx, vy, m, eps2, G=G_AU):
    ke = 0.5 * float(np.sum(m * (vx * vx + vy * vy)))
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    r = np.sqrt(dx * dx + dy * dy + eps2)
    mm = m[:, None] * m[None, :]
    pe = 0.5 * (-G * float(np.sum(mm / r)) + G * float(np.sum(m * m)) / np.sqrt(eps2))
    return ke + pe, (float(np.sum(m * vx)), float(np.sum(m * vy)))
