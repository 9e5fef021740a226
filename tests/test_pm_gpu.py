"""GPU parity of the particle-mesh path: FFT against numpy.fft, CIC cells
bit-exact, deposit/gather, the 2-D isolated force field against the NumPy
restatement of pm_kernel.py, the periodic 3-D Poisson solver, and the
reference's test_pm.py checks through the drop-in functions."""
import numpy as np
import pytest
import torch

import oracle
from oracle import pm_np
from solarsystemsimulator_amd import pm as P
from solarsystemsimulator_amd import device as D

pytestmark = pytest.mark.gpu
G_AU = 4 * np.pi ** 2


def dev_t(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dev, dtype)


@pytest.mark.parametrize("prec", [P.F64, P.F32])
@pytest.mark.parametrize("dims", [[2], [8], [256], [4096], [8192], [2, 4], [16, 8],
                                  [64, 128], [256, 16], [1024, 512], [4, 8, 2],
                                  [32, 32, 32], [128, 64, 256], [256, 4, 256],
                                  [8192, 8], [4, 8192]])
def test_fft_matches_numpy(dev, dims, prec):
    print(f"fft dims={dims} prec={prec}", flush=True)
    rng = np.random.default_rng(len(dims) * 100 + dims[-1])
    a = rng.standard_normal(dims)
    plan = P.FFTPlan.get(dims, prec, dev)
    rd = torch.float64 if prec == P.F64 else torch.float32
    spec = plan.r2c(dev_t(a, dev, rd))
    got = spec.cpu().numpy().astype(np.float64)
    ref = np.fft.rfftn(a)
    tol = 1e-13 if prec == P.F64 else 2e-6
    scale = np.abs(ref).max()
    assert np.abs(got[..., 0] - ref.real).max() < tol * scale * np.log2(a.size + 1)
    assert np.abs(got[..., 1] - ref.imag).max() < tol * scale * np.log2(a.size + 1)
    back = plan.c2r(spec).cpu().numpy()
    assert np.abs(back - a).max() < tol * np.abs(a).max() * np.log2(a.size + 1)


@pytest.mark.parametrize("tiled", [True, False])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dim", [2, 3])
def test_cic_cells_bit_exact_and_deposit(dev, dim, periodic, tiled):
    rng = np.random.default_rng(5 + dim)
    n = [32, 16, 64][:dim]
    origin = [-0.3, 0.1, 0.7][:dim]
    h = 1.0 / 37.0
    npart = 20000
    pos = [rng.uniform(o - 2 * h, o + (k + 1) * h, npart) for o, k in zip(origin, n)]
    pos[0][:3] = [origin[0], origin[0] + 5 * h, origin[0] + (n[0] - 1) * h]
    m = rng.uniform(0.5, 2.0, npart)
    cells = torch.empty(dim * npart, dtype=torch.int32, device=dev)
    tp = [dev_t(c, dev) for c in pos]
    mesh = P.deposit(tp, dev_t(m, dev), n, origin, h, periodic, cells=cells, tiled=tiled)
    ref_cells, _ = pm_np.cic_cells(pos, origin, h, n, periodic)
    got_cells = cells.cpu().numpy().reshape(npart, dim).T
    assert np.array_equal(got_cells, ref_cells)
    ref = pm_np.cic_deposit(pos, m, n, origin, h, periodic)
    np.testing.assert_allclose(mesh.cpu().numpy(), ref, rtol=1e-12,
                               atol=1e-14 * m.sum())
    if tiled:                      # every node written, bit-identical again
        again = torch.full(n, 5.0, dtype=torch.float64, device=dev)
        P.deposit(tp, dev_t(m, dev), n, origin, h, periodic, mesh=again)
        assert torch.equal(again, mesh)


@pytest.mark.parametrize("case", ["lattice", "shuffled", "clump", "tiny_periodic",
                                  "ragged", "grid2048"])
@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_cic_tiles_cases(dev, case, prec):
    """The tiled deposit against the oracle and run to run: the 262,144-body
    periodic cube in lattice order and shuffled, all masses in one tile, a
    periodic mesh smaller than a tile (its one tile is its own neighbour),
    a mesh that is not a multiple of the tile edge, and the 2-D 2048^2 mesh
    of NBodyPM with 10,000 bodies."""
    print(f"cic tiles {case} prec={prec}", flush=True)
    rng = np.random.default_rng(77)
    if case in ("lattice", "shuffled"):
        from solarsystemsimulator_amd import scenes
        pos = list(scenes.perturbed_cube(64, 64, 64, seed=3, periodic=True))
        if case == "shuffled":
            perm = rng.permutation(pos[0].size)
            pos = [c[perm] for c in pos]
        n, origin, h, periodic = [128] * 3, [0.0] * 3, 1.0 / 128, True
    elif case == "clump":
        pos = [0.5 + rng.normal(0.0, 0.003, 30000) for _ in range(3)]
        n, origin, h, periodic = [64] * 3, [0.0] * 3, 1.0 / 64, False
    elif case == "tiny_periodic":
        pos = [rng.uniform(0.0, 1.0, 5000) for _ in range(3)]
        n, origin, h, periodic = [5, 6, 7], [0.0] * 3, 1.0 / 5, True
    elif case == "ragged":
        pos = [rng.uniform(-0.1, 1.1, 40000) for _ in range(3)]
        n, origin, h, periodic = [37, 29, 45], [0.0] * 3, 1.0 / 30, False
    else:
        pos = [rng.uniform(-4e4, 4e4, 10000) for _ in range(2)]
        org, hh = pm_np.grid_fuer(pos, 2048, 4.0)
        n, origin, h, periodic = [2048, 2048], org, hh, False
    m = rng.uniform(0.5, 1.5, pos[0].size) / pos[0].size
    tp = [dev_t(c, dev) for c in pos]
    md = dev_t(m, dev)
    a = P.deposit(tp, md, n, origin, h, periodic, prec)
    b = P.deposit(tp, md, n, origin, h, periodic, prec)
    assert torch.equal(a, b)
    ref = pm_np.cic_deposit(pos, m, n, origin, h, periodic)
    got = a.cpu().numpy().astype(np.float64)
    tol = 1e-12 if prec == P.F64 else 1e-7
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert abs(got.sum() - m.sum()) <= (1e-12 if prec == P.F64 else 1e-6) * m.sum()
    # gather is deterministic: bit-exact against the oracle on the same mesh
    field = rng.standard_normal(n)
    got = P.gather([dev_t(field, dev)], tp, n, origin, h, periodic)[0]
    exp = pm_np.cic_gather(field, pos, n, origin, h, periodic)
    assert np.array_equal(got.cpu().numpy(), exp)


def test_force_field_2d_matches_oracle(dev):
    rng = np.random.default_rng(3)
    n = 3000
    x = rng.uniform(-40000, 40000, n)
    y = rng.uniform(-40000, 40000, n)
    m = rng.uniform(1e6, 2e7, n)
    gx, gy, gm = dev_t(x, dev), dev_t(y, dev), dev_t(m, dev)
    x0, y0, h = P.grid_fuer(gx, gy, 256)
    (ox0, oy0), oh = pm_np.grid_fuer([x, y], 256)
    assert (x0, y0, h) == (ox0, oy0, oh)
    eps2 = (2 * h) ** 2
    ax, ay = P.pm_accelerations(gx, gy, gm, 256, x0, y0, h, eps2)
    rx, ry = pm_np.pm_accelerations([x, y], m, 256, [x0, y0], h, eps2)
    a = np.hypot(rx, ry)
    err = np.hypot(ax.cpu().numpy() - rx, ay.cpu().numpy() - ry) / a
    assert err.max() < 1e-10


@pytest.mark.parametrize("prec", [P.F64, P.F32])
@pytest.mark.parametrize("dims", [[32, 32, 32], [128, 128, 128], [64, 32, 16],
                                  [16, 128, 64], [256, 8, 32]])
def test_fused_periodic_potential_bit_identical(dev, dims, prec):
    """ssim_pm_potential_periodic (x axis forward, Poisson factor and
    inverse in one pass) == R2C + ssim_pm_poisson + C2R, bit for bit."""
    from solarsystemsimulator_amd import _lib
    print(f"fused potential dims={dims} prec={prec}", flush=True)
    rng = np.random.default_rng(sum(dims))
    rd = torch.float64 if prec == P.F64 else torch.float32
    rho = dev_t(rng.uniform(0.0, 1.0, dims), dev, rd)
    plan = P.FFTPlan.get(dims, prec, dev)
    gn = _lib.int_array(dims)
    h, G = 1.0 / dims[0], 0.7
    spec = plan.r2c(rho)
    _lib.call("ssim_pm_poisson", 3, _lib.ctypes.addressof(gn), h, G, prec,
              _lib.ptr(spec), _lib.stream_handle(dev))
    ref = plan.c2r(spec).cpu().numpy()
    spec2 = plan.new_complex()
    phi = torch.empty(dims, dtype=rd, device=dev)
    _lib.call("ssim_pm_potential_periodic", _lib.ctypes.addressof(gn), prec, _lib.ptr(rho),
              _lib.ptr(spec2), _lib.ptr(phi), h, G, _lib.ptr(plan.work), plan.work_bytes,
              _lib.stream_handle(dev))
    got = phi.cpu().numpy()
    assert np.abs(ref).max() > 0
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("prec,tol", [(P.F64, 1e-10), (P.F32, 2e-4)])
def test_periodic_poisson_3d_matches_oracle(dev, prec, tol):
    rng = np.random.default_rng(9)
    n, npart = 32, 20000
    pos = [rng.uniform(0, 1, npart) for _ in range(3)]
    m = np.full(npart, 1.0 / npart)
    solver = P.PeriodicPM3D(dev, n, 1.0, (0, 0, 0), 1.0, prec)
    tp = [dev_t(c, dev) for c in pos]
    got = solver.accelerations(tp, dev_t(m, dev), tp)
    ref = pm_np.periodic_accelerations(pos, m, pos, n, 1.0, [0, 0, 0], 1.0)
    g = np.array([c.cpu().numpy() for c in got])
    r = np.array(ref)
    err = np.linalg.norm(g - r, axis=0) / np.linalg.norm(r, axis=0)
    assert np.median(err) < tol and np.linalg.norm(g - r) / np.linalg.norm(r) < tol


def test_isolated_3d_matches_oracle(dev):
    rng = np.random.default_rng(10)
    n, npart = 32, 5000
    pos = [rng.uniform(0.1, 0.9, npart) for _ in range(3)]
    m = np.full(npart, 1.0 / npart)
    h = 1.0 / (n - 4)
    origin = [-2 * h] * 3
    solver = P.IsolatedPM3D(dev, n, origin, h, (1.5 * h) ** 2, 1.0, P.F64)
    tp = [dev_t(c, dev) for c in pos]
    got = solver.accelerations(tp, dev_t(m, dev), tp)
    ref = pm_np.pm_accelerations(pos, m, n, origin, h, (1.5 * h) ** 2, 1.0)
    g = np.array([c.cpu().numpy() for c in got])
    r = np.array(ref)
    err = np.linalg.norm(g - r, axis=0) / np.linalg.norm(r, axis=0)
    assert err.max() < 1e-10


# --- the reference's test_pm.py through the drop-in API --------------------

def test_ref_1_momentum(dev):
    rng = np.random.default_rng(42)
    n = 5000
    x = dev_t(rng.uniform(-40000, 40000, n), dev)
    y = dev_t(rng.uniform(-40000, 40000, n), dev)
    m = dev_t(rng.uniform(1e6, 2e7, n), dev)
    x0, y0, h = P.grid_fuer(x, y, 512)
    ax, ay = P.pm_accelerations(x, y, m, 512, x0, y0, h, (2 * h) ** 2)
    px = float(torch.sum(m * ax))
    py = float(torch.sum(m * ay))
    skala = float(torch.sum(m * torch.hypot(ax, ay)))
    assert np.hypot(px, py) / skala < 1e-3


def test_ref_2_far_pair(dev):
    grid_n, h = 512, 100.0
    eps2 = (2 * h) ** 2
    m_val, r = 1e10, 60 * h
    cx = grid_n * h / 2
    x = dev_t([cx - r / 2, cx + r / 2], dev)
    y = dev_t([cx, cx], dev)
    m = dev_t([m_val, m_val], dev)
    ax, ay = P.pm_accelerations(x, y, m, grid_n, 0.0, 0.0, h, eps2)
    analytic = G_AU * m_val / (r * r)
    assert abs(abs(float(ax[0])) - analytic) / analytic < 0.02


@pytest.mark.parametrize("eps_in_h,schwelle", [(4.0, 0.03), (1.0, 0.12)])
def test_ref_3_4_cloud_vs_allpairs(dev, eps_in_h, schwelle):
    rng = np.random.default_rng(42)
    n = 4000
    xs = rng.uniform(-40000, 40000, n)
    ys = rng.uniform(-40000, 40000, n)
    ms = rng.uniform(1e6, 2e7, n)
    x, y, m = dev_t(xs, dev), dev_t(ys, dev), dev_t(ms, dev)
    x0, y0, h = P.grid_fuer(x, y, 512)
    eps2 = (eps_in_h * h) ** 2
    ax, ay = P.pm_accelerations(x, y, m, 512, x0, y0, h, eps2)
    rx, ry = oracle.allpairs([xs, ys], ms, [xs, ys], G_AU, eps2)
    rel = np.hypot(ax.cpu().numpy() - rx, ay.cpu().numpy() - ry) / np.hypot(rx, ry)
    assert np.median(rel) < schwelle


def test_nbodypm_default_grid(dev):
    """NBodyPM() with the reference's defaults (grid_n = 4096: 8192-point
    padded FFT lines, pm_kernel.py:190): the PM accelerations of a clumped
    disc against the direct sum (median error below the 2048-grid one's
    bound) and 20 KDK steps conserving momentum."""
    print("NBodyPM default grid", flush=True)
    rng = np.random.default_rng(43)
    n, R, Mtot = 4000, 30000.0, 5.6e9
    r = R * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    x, y = r * np.cos(th), r * np.sin(th)
    m = np.full(n, Mtot / n)
    kern = P.NBodyPM([0])
    assert kern.grid_n == 4096
    gx, gy = dev_t(x, dev), dev_t(y, dev)
    x0, y0, h = P.grid_fuer(gx, gy, 4096, 4.0)
    eps2 = (kern.softening_zellen * h) ** 2
    ax, ay = P.pm_accelerations(gx, gy, dev_t(m, dev), 4096, x0, y0, h, eps2)
    rx, ry = oracle.allpairs([x, y], m, [x, y], G_AU, eps2)
    rel = np.hypot(ax.cpu().numpy() - rx, ay.cpu().numpy() - ry) / np.hypot(rx, ry)
    assert np.median(rel) < 0.05
    vvir = np.sqrt(G_AU * Mtot / R)
    ang = rng.uniform(0, 2 * np.pi, n)
    vx, vy = 0.7 * vvir * np.cos(ang), 0.7 * vvir * np.sin(ang)
    vx -= np.sum(m * vx) / Mtot
    vy -= np.sum(m * vy) / Mtot
    st = kern.load_state(x, y, vx, vy, m)
    out = kern.step_batch(st, R / vvir / 1000.0, steps=20)
    assert out.shape == (20, 4 * n)
    f = kern.export_f64(st)
    _, p1 = pm_np.energie_impuls(f[0:n], f[n:2 * n], f[2 * n:3 * n], f[3 * n:4 * n],
                                 m, eps2)
    p_rel = np.hypot(*p1) / np.hypot(np.sum(m * np.abs(vx)), np.sum(m * np.abs(vy)))
    assert p_rel < 1e-6


def test_ref_5_dynamics(dev):
    """test_pm.py:121-159: 100 KDK steps of NBodyPM, energy drift < 3 %,
    residual momentum < 1e-6 (grid_n 2048 -> padded 4096)."""
    rng = np.random.default_rng(42)
    n, R, Mtot = 10000, 30000.0, 5.6e9
    r = R * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    x, y = r * np.cos(th), r * np.sin(th)
    m = np.full(n, Mtot / n)
    vvir = np.sqrt(G_AU * Mtot / R)
    ang = rng.uniform(0, 2 * np.pi, n)
    vx, vy = 0.7 * vvir * np.cos(ang), 0.7 * vvir * np.sin(ang)
    vx -= np.sum(m * vx) / Mtot
    vy -= np.sum(m * vy) / Mtot
    kern = P.NBodyPM([0], grid_n=2048, softening_zellen=1.5)
    _, _, h = P.grid_fuer(dev_t(x, dev), dev_t(y, dev), 2048, 4.0)
    eps2 = (1.5 * h) ** 2
    e0, _ = pm_np.energie_impuls(x, y, vx, vy, m, eps2)
    st = kern.load_state(x, y, vx, vy, m)
    out = kern.step_batch(st, R / vvir / 1000.0, steps=100)
    assert out.shape == (100, 4 * n)
    f = kern.export_f64(st)
    e1, p1 = pm_np.energie_impuls(f[0:n], f[n:2 * n], f[2 * n:3 * n],
                                  f[3 * n:4 * n], m, eps2)
    e_drift = abs(e1 - e0) / abs(e0)
    p_rel = np.hypot(*p1) / max(np.hypot(np.sum(m * np.abs(vx)),
                                         np.sum(m * np.abs(vy))), 1e-30)
    assert e_drift < 0.03 and p_rel < 1e-6
