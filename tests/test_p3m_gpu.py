"""GPU parity of the cell list (bit-exact order) and the near-field force
(FP32 terms, <= 1e-5 relative) against the NumPy oracle, and the accuracy
of the split mesh + near-field force against the direct FP64 sum."""
import numpy as np
import pytest
import torch

import oracle
from oracle import p3m_np
from solarsystemsimulator_amd import _lib, scenes
from solarsystemsimulator_amd import device as D
from solarsystemsimulator_amd import pm as P
from solarsystemsimulator_amd.p3m import SplitForce, relative_errors

pytestmark = pytest.mark.gpu


def cloud(n, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return [rng.uniform(lo, hi, n) for _ in range(3)], rng.uniform(0.5, 1.5, n) / n


def test_cell_list_bit_exact(dev):
    pos, m = cloud(20000, 1, -0.2, 1.3)
    origin, cell, nc = [0.0, 0.0, 0.0], 0.07, [15, 12, 17]
    ncells = int(np.prod(nc))
    start = torch.empty(ncells + 1, dtype=torch.int32, device=dev)
    order = torch.empty(20000, dtype=torch.int32, device=dev)
    packed = torch.empty((20000, 4), dtype=torch.float32, device=dev)
    wb = _lib.load().ssim_cells_work_bytes(20000, ncells)
    w = D.workspace(wb, dev)
    x = [D.to_dev(c, dev) for c in pos]
    md = D.to_dev(m, dev)
    gn, org = _lib.int_array(nc), _lib.double_array(origin)
    import ctypes
    _lib.call("ssim_cells_build", 20000, *[_lib.ptr(c) for c in x],
              _lib.ptr(md), ctypes.addressof(org), cell,
              ctypes.addressof(gn), _lib.ptr(start), _lib.ptr(order),
              _lib.ptr(packed), _lib.ptr(w), wb, D.stream(dev))
    rs, ro = p3m_np.cell_list(pos, origin, cell, nc)
    assert np.array_equal(start.cpu().numpy(), rs)
    assert np.array_equal(order.cpu().numpy(), ro)
    pk = packed.cpu().numpy()
    np.testing.assert_array_equal(pk[:, 0], (pos[0][ro] - origin[0]).astype(np.float32))
    np.testing.assert_array_equal(pk[:, 3], m[ro].astype(np.float32))


@pytest.mark.parametrize("n,nc,clump,inv_cell", [(300_000, [40, 33, 29], 0.3, 29),
                                                  (300_000, [50, 41, 37], 0.0, 37),
                                                  (5, [2, 1, 3], 0.0, 29)])
def test_cell_list_large_and_clumped(dev, n, nc, clump, inv_cell):
    """Radix passes over > 4096-entry histograms, cell scans in one workgroup
    (<= 65536 cells) and over several (75850 cells), a clump holding a third
    of the bodies in one cell, and no packed copy."""
    print(f"cell list n={n} nc={nc}", flush=True)
    rng = np.random.default_rng(4)
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    k = int(clump * n)
    for c in range(3):
        pos[c][:k] = 0.5 + rng.uniform(0.0, 1e-3, k)
    origin, cell = [0.0, 0.0, 0.0], 1.0 / inv_cell
    ncells = int(np.prod(nc))
    start = torch.empty(ncells + 1, dtype=torch.int32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    w = D.workspace(_lib.load().ssim_cells_work_bytes(n, ncells), dev)
    import ctypes
    gn, org = _lib.int_array(nc), _lib.double_array(origin)
    x = [D.to_dev(c, dev) for c in pos]
    _lib.call("ssim_cells_build", n, *[_lib.ptr(c) for c in x], None,
              ctypes.addressof(org), cell, ctypes.addressof(gn), _lib.ptr(start),
              _lib.ptr(order), None, _lib.ptr(w), w.numel(), D.stream(dev))
    rs, ro = p3m_np.cell_list(pos, origin, cell, nc)
    assert np.array_equal(start.cpu().numpy(), rs)
    assert np.array_equal(order.cpu().numpy(), ro)


def test_cell_bin_octant_order_stable(dev):
    """ssim_cells_bin: same cell_start as the stable list, the bodies of a
    cell grouped by octant and ascending by index inside an octant (a
    stable sort of cell * 8 + octant)."""
    n, nc = 200_000, [23, 31, 17]
    rng = np.random.default_rng(8)
    pos = [rng.normal(0.5, 0.2, n) for _ in range(3)]
    origin, cell = [0.0, 0.0, 0.0], 1.0 / 23
    ncells = int(np.prod(nc))
    start = torch.empty(ncells + 1, dtype=torch.int32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    w = D.workspace(_lib.load().ssim_cells_bin_work_bytes(n, ncells), dev)
    import ctypes
    gn, org = _lib.int_array(nc), _lib.double_array(origin)
    x = [D.to_dev(c, dev) for c in pos]
    _lib.call("ssim_cells_bin", n, *[_lib.ptr(c) for c in x],
              ctypes.addressof(org), cell, ctypes.addressof(gn), _lib.ptr(start),
              _lib.ptr(order), _lib.ptr(w), w.numel(), D.stream(dev))
    rs, ro = p3m_np.cell_list(pos, origin, cell, nc)
    got_s, got_o = start.cpu().numpy(), order.cpu().numpy()
    assert np.array_equal(got_s, rs)
    # sorting each cell's segment by body index gives the stable list
    keys = np.repeat(np.arange(ncells), np.diff(rs))
    assert np.array_equal(got_o[np.lexsort((got_o, keys))], ro)
    # inside a cell the bodies come grouped by octant (x, y, z halves)
    f = [(pos[c] - origin[c]) * (1.0 / cell) for c in range(3)]
    ijk = [np.clip(np.floor(f[c]).astype(np.int64), 0, nc[c] - 1) for c in range(3)]
    octant = sum(np.clip(np.floor(2.0 * f[c]).astype(np.int64) - 2 * ijk[c], 0, 1)
                 << (2 - c) for c in range(3))
    key8 = ((ijk[0] * nc[1] + ijk[1]) * nc[2] + ijk[2]) * 8 + octant
    assert np.array_equal(got_o, np.argsort(key8, kind="stable"))


def test_tsc_deposit_and_gather_match_oracle(dev):
    print("tsc deposit/gather", flush=True)
    n, h, origin = 24, 0.05, [0.1, -0.2, 0.3]
    rng = np.random.default_rng(6)
    pos = [origin[a] + rng.uniform(1.6, n - 2.6, 30000) * h for a in range(3)]
    m = rng.uniform(0.5, 2.0, 30000)
    x = [D.to_dev(c, dev) for c in pos]
    mesh = torch.zeros((n, n, n), dtype=torch.float64, device=dev)
    import ctypes
    org = _lib.double_array(origin)
    md = D.to_dev(m, dev)
    _lib.call("ssim_tsc_deposit", 30000, *[_lib.ptr(c) for c in x],
              _lib.ptr(md), n, ctypes.addressof(org), h, 0,
              _lib.ptr(mesh), D.stream(dev))
    ref = p3m_np.tsc_deposit(pos, m, n, origin, h)
    np.testing.assert_allclose(mesh.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * ref.max())
    # the atomic-free deposit from nearest-node bins, twice: bit-identical
    nn = n ** 3
    start = torch.empty(nn + 1, dtype=torch.int32, device=dev)
    order = torch.empty(30000, dtype=torch.int32, device=dev)
    w = D.workspace(_lib.load().ssim_cells_work_bytes(30000, nn), dev)
    gn = _lib.int_array([n] * 3)
    org_b = _lib.double_array([o - 0.5 * h for o in origin])
    _lib.call("ssim_cells_build", 30000, *[_lib.ptr(c) for c in x], None,
              ctypes.addressof(org_b), h, ctypes.addressof(gn), _lib.ptr(start),
              _lib.ptr(order), None, _lib.ptr(w), w.numel(), D.stream(dev))
    binned = []
    for _ in range(2):
        mb = torch.full((n, n, n), 7.0, dtype=torch.float64, device=dev)
        bw = D.workspace(_lib.load().ssim_tsc_deposit_work_bytes(30000), dev)
        _lib.call("ssim_tsc_deposit_binned", 30000, *[_lib.ptr(c) for c in x],
                  _lib.ptr(md), n, ctypes.addressof(org), h, _lib.ptr(start),
                  _lib.ptr(order), 0, _lib.ptr(mb), _lib.ptr(bw), bw.numel(),
                  D.stream(dev))
        binned.append(mb.cpu().numpy())
    assert np.array_equal(binned[0], binned[1])
    np.testing.assert_allclose(binned[0], ref, rtol=1e-12, atol=1e-12 * ref.max())
    fields = [rng.normal(size=(n, n, n)) for _ in range(3)]
    fd = [D.to_dev(f, dev) for f in fields]        # keep the device copies alive
    out = [torch.empty(30000, dtype=torch.float64, device=dev) for _ in range(3)]
    _lib.call("ssim_tsc_gather", 30000, None, *[_lib.ptr(c) for c in x], n,
              ctypes.addressof(org), h, 0, *[_lib.ptr(f) for f in fd],
              *[_lib.ptr(o) for o in out], 0, D.stream(dev))
    for c in range(3):
        np.testing.assert_allclose(out[c].cpu().numpy(),
                                   p3m_np.tsc_gather(fields[c], pos, n, origin, h),
                                   rtol=1e-12, atol=1e-12)
    # FP32 fields: three arrays and the interleaved float4 mesh agree exactly
    f32 = [D.to_dev(f, dev, torch.float32) for f in fields]
    f4 = torch.zeros((n ** 3, 4), dtype=torch.float32, device=dev)
    for c in range(3):
        f4[:, c] = f32[c].reshape(-1)
    o3 = [torch.empty(30000, dtype=torch.float64, device=dev) for _ in range(3)]
    o4 = [torch.empty(30000, dtype=torch.float64, device=dev) for _ in range(3)]
    _lib.call("ssim_tsc_gather", 30000, None, *[_lib.ptr(c) for c in x], n,
              ctypes.addressof(org), h, 1, *[_lib.ptr(f) for f in f32],
              *[_lib.ptr(o) for o in o3], 0, D.stream(dev))
    _lib.call("ssim_tsc_gather", 30000, None, *[_lib.ptr(c) for c in x], n,
              ctypes.addressof(org), h, 1, _lib.ptr(f4), None, None,
              *[_lib.ptr(o) for o in o4], 0, D.stream(dev))
    for c in range(3):
        assert torch.equal(o3[c], o4[c])
    # any iteration order gives the same per-body results
    perm = D.to_dev(np.random.default_rng(7).permutation(30000), dev, torch.int32)
    o5 = [torch.empty(30000, dtype=torch.float64, device=dev) for _ in range(3)]
    _lib.call("ssim_tsc_gather", 30000, _lib.ptr(perm), *[_lib.ptr(c) for c in x], n,
              ctypes.addressof(org), h, 1, _lib.ptr(f4), None, None,
              *[_lib.ptr(o) for o in o5], 0, D.stream(dev))
    for c in range(3):
        assert torch.equal(o3[c], o5[c])
    _lib.call("ssim_tsc_gather", 30000, None, *[_lib.ptr(c) for c in x], n,
              ctypes.addressof(org), h, 1, _lib.ptr(f4), None, None,
              *[_lib.ptr(o) for o in o5], 1, D.stream(dev))      # accumulate
    for c in range(3):
        assert torch.equal(o5[c], 2.0 * o3[c])


@pytest.mark.parametrize("assign", ["cic", "tsc"])
def test_split_force_assignments_on_cube(dev, assign):
    """Median force error of the split force on a 64k-mass cube against the
    FP64 direct sum: CIC < 1e-2, TSC below CIC's."""
    print(f"split force {assign}", flush=True)
    pos, m, tpos = scenes.learned_scene((40, 40, 40), (40, 40, 40), seed=5,
                                        tracer_seed=55)
    n = len(m)
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    idx = np.random.default_rng(0).choice(len(allx[0]), 2000, replace=False)
    ref = np.array(oracle.allpairs(pos, m, [c[idx] for c in allx], 1.0, 1e-6, plain=True))
    errs = {}
    for a_ in ("cic", assign):
        sf = SplitForce(dev, mesh_n=64, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0, assign=a_)
        a = [torch.zeros(len(allx[0]), dtype=torch.float64, device=dev) for _ in range(3)]
        sf.accelerations(x, n, md, a)
        got = np.array([c.cpu().numpy()[idx] for c in a])
        errs[a_] = np.median(relative_errors(got, ref))
        # the two-stream evaluation (TSC) equals mesh then near field in
        # order; bit for bit, as the binned TSC deposit is reproducible (the
        # CIC deposit's atomics are not, to rounding)
        b = [torch.zeros_like(c) for c in a]
        sf.mesh_accel(x, n, md, b)
        sf.near_accel(x, n, md, b)
        for c in range(3):
            if a_ == "tsc":
                assert torch.equal(a[c], b[c])
            else:
                scale = float(a[c].abs().max())
                torch.testing.assert_close(a[c], b[c], rtol=1e-12, atol=1e-12 * scale)
    assert errs["cic"] < 1e-2
    if assign == "tsc":
        assert errs["tsc"] < 0.5 * errs["cic"], errs


@pytest.mark.parametrize("B", [5, 3, 2])
def test_split_bins_and_blocked_deposit(dev, B):
    """ssim_split_bin equals the oracle's binning exactly (stable order, bin
    and block starts, both packed copies); the blocked deposit matches the
    TSC oracle and is bit-identical run to run."""
    print(f"split bins B={B}", flush=True)
    import ctypes
    n, h, nbody = 24, 0.05, 40_000
    origin = [0.1, -0.2, 0.3]
    rng = np.random.default_rng(12)
    pos = [origin[a] + rng.uniform(1.0, n - 2.0, nbody) * h for a in range(3)]
    pos[0][:500] = origin[0] + (np.arange(500) % 20 + 0.5) * h   # on bin edges
    m = rng.uniform(0.5, 2.0, nbody)
    x = [D.to_dev(c, dev) for c in pos]
    md = D.to_dev(m, dev)
    nb = -(-n // B)
    K = nb ** 3 * B ** 3
    i32 = dict(dtype=torch.int32, device=dev)
    start = torch.empty(K + 1, **i32)
    cstart = torch.empty(nb ** 3 + 1, **i32)
    order = torch.empty(nbody, **i32)
    packed = torch.empty((nbody, 4), dtype=torch.float32, device=dev)
    body = torch.empty((nbody, 4), dtype=torch.float64, device=dev)
    near = torch.empty((nbody, 4), dtype=torch.float32, device=dev)
    near_lo = torch.empty((nbody, 4), dtype=torch.float32, device=dev)
    w = D.workspace(_lib.load().ssim_split_bin_work_bytes(nbody, n, B), dev)
    org = _lib.double_array(origin)
    _lib.call("ssim_split_bin", nbody, *[_lib.ptr(c) for c in x], _lib.ptr(md), n,
              ctypes.addressof(org), h, B, _lib.ptr(start), _lib.ptr(cstart),
              _lib.ptr(order), _lib.ptr(packed), _lib.ptr(body), _lib.ptr(near),
              _lib.ptr(near_lo), _lib.ptr(w), w.numel(), D.stream(dev))
    rs, rc, ro = p3m_np.split_bins(pos, n, origin, h, B)
    assert np.array_equal(start.cpu().numpy(), rs)
    assert np.array_equal(cstart.cpu().numpy(), rc)
    assert np.array_equal(order.cpu().numpy(), ro)
    pk, bd, nr = packed.cpu().numpy(), body.cpu().numpy(), near.cpu().numpy()
    lo = near_lo.cpu().numpy()
    for c in range(3):
        assert np.array_equal(bd[:, c], (pos[c][ro] - origin[c]) / h)
        assert np.array_equal(pk[:, c], (pos[c][ro] - (origin[c] - 0.5 * h)).astype(np.float32))
        # relative to the corner of the mass's block (its TSC centre node // B)
        blk = np.clip(np.floor((pos[c][ro] - origin[c]) / h + 0.5), 0, n - 1).astype(int) // B
        corner = (origin[c] - 0.5 * h) + blk * (B * h)
        rel = pos[c][ro] - corner
        assert np.array_equal(nr[:, c], rel.astype(np.float32))
        # hi + lo: the block-relative position to the rounding of the remainder
        assert np.array_equal(lo[:, c], (rel - nr[:, c].astype(np.float64)).astype(np.float32))
        assert np.abs(nr[:, c].astype(np.float64) + lo[:, c] - rel).max() <= 2.0 ** -48 * B * h
    assert np.array_equal(bd[:, 3], m[ro])
    assert np.array_equal(nr[:, 3], m[ro].astype(np.float32))
    assert not lo[:, 3].any()
    meshes = []
    for _ in range(2):
        mb = torch.full((n, n, n), 7.0, dtype=torch.float64, device=dev)
        _lib.call("ssim_tsc_deposit_blocked", n, ctypes.addressof(org), h, B,
                  _lib.ptr(start), _lib.ptr(body), 0, _lib.ptr(mb), D.stream(dev))
        meshes.append(mb.cpu().numpy())
    assert np.array_equal(meshes[0], meshes[1])
    ref = p3m_np.tsc_deposit(pos, m, n, origin, h)
    np.testing.assert_allclose(meshes[0], ref, rtol=1e-12, atol=1e-12 * ref.max())
    if B == 5:
        _tiled_deposit_checks(dev, n, origin, h, cstart, body, m, ref)
    # the blocks alone (bin_start = NULL): stable block order, both copies
    cstart2 = torch.empty_like(cstart)
    _lib.call("ssim_split_bin", nbody, *[_lib.ptr(c) for c in x], _lib.ptr(md), n,
              ctypes.addressof(org), h, B, None, _lib.ptr(cstart2),
              _lib.ptr(order), _lib.ptr(packed), _lib.ptr(body), None, None, _lib.ptr(w),
              w.numel(), D.stream(dev))
    _, rc2, ro2 = p3m_np.split_bins(pos, n, origin, h, B, nodes=False)
    assert np.array_equal(cstart2.cpu().numpy(), rc2) and np.array_equal(rc2, rc)
    assert np.array_equal(order.cpu().numpy(), ro2)
    bd = body.cpu().numpy()
    for c in range(3):
        assert np.array_equal(bd[:, c], (pos[c][ro2] - origin[c]) / h)
    if B == 5:
        _tiled_deposit_checks(dev, n, origin, h, cstart2, body, m, ref)


def _tiled_deposit_checks(dev, n, origin, h, cstart, body, m, ref):
    """ssim_tsc_deposit_tiles: the TSC oracle in FP64 (bit-identical run to
    run: int64 fixed-point sums) and in FP32."""
    import ctypes
    org = _lib.double_array(origin)
    bound = float(np.abs(m).sum())
    meshes = []
    for prec in (0, 0, 1):
        dt = torch.float64 if prec == 0 else torch.float32
        mt = torch.full((n, n, n), 7.0, dtype=dt, device=dev)
        _lib.call("ssim_tsc_deposit_tiles", n, ctypes.addressof(org), h, 5,
                  _lib.ptr(cstart), _lib.ptr(body), bound, prec, _lib.ptr(mt),
                  D.stream(dev))
        meshes.append(mt.cpu().numpy().astype(np.float64))
    assert np.array_equal(meshes[0], meshes[1])
    np.testing.assert_allclose(meshes[0], ref, rtol=1e-12, atol=1e-12 * ref.max())
    np.testing.assert_allclose(meshes[2], ref, rtol=1e-6, atol=1e-7 * ref.max())


def _split_bin_call(dev, pos, m, n, origin, h, B):
    """ssim_split_bin (blocks only) on the masses; host copies."""
    import ctypes
    nbody = len(m)
    nb = -(-n // B)
    i32 = dict(dtype=torch.int32, device=dev)
    out = dict(cstart=torch.empty(nb ** 3 + 1, **i32), order=torch.empty(max(nbody, 1), **i32))
    for k in ("packed", "near", "near_lo"):
        out[k] = torch.empty((max(nbody, 1), 4), dtype=torch.float32, device=dev)
    out["body"] = torch.empty((max(nbody, 1), 4), dtype=torch.float64, device=dev)
    x = [D.to_dev(c, dev) for c in pos]
    md = D.to_dev(m, dev)
    w = D.workspace(_lib.load().ssim_split_bin_work_bytes(nbody, n, B), dev)
    org = _lib.double_array(origin)
    _lib.call("ssim_split_bin", nbody, *[_lib.ptr(c) for c in x], _lib.ptr(md), n,
              ctypes.addressof(org), h, B, None, _lib.ptr(out["cstart"]),
              _lib.ptr(out["order"]), _lib.ptr(out["packed"]), _lib.ptr(out["body"]),
              _lib.ptr(out["near"]), _lib.ptr(out["near_lo"]), _lib.ptr(w), w.numel(),
              D.stream(dev))
    return {k: v.cpu().numpy()[:nbody] if k not in ("cstart",) else v.cpu().numpy()
            for k, v in out.items()}


@pytest.mark.parametrize("case", ["uniform", "clump", "no_tracers", "no_masses", "empty",
                                  "one", "two"])
def test_split_bin_all_equals_two_binnings(dev, case):
    """ssim_split_bin_all (one binning of masses and tracers): its target
    lists equal the oracle's block binning of all bodies (stable, index
    order inside a block), and its mass outputs equal ssim_split_bin's on
    the masses alone bit for bit (starts, order, all packed copies)."""
    print(f"split bin all {case}", flush=True)
    import ctypes
    n, h, B = 30, 0.04, 5
    origin = [0.1, -0.2, 0.3]
    rng = np.random.default_rng(31)
    nm, nt = {"uniform": (20_000, 45_000), "clump": (20_000, 45_000),
              "no_tracers": (7_000, 0), "no_masses": (0, 9_000), "empty": (0, 0),
              "one": (1, 0), "two": (1, 1)}[case]
    pos = [origin[a] + rng.uniform(0.5, n - 1.5, nm + nt) * h for a in range(3)]
    if case == "clump":
        for a in range(3):
            pos[a][:8000] = origin[a] + 12.3 * h + rng.normal(0.0, 0.3 * h, 8000)
            pos[a][nm:nm + 9000] = origin[a] + 12.3 * h + rng.normal(0.0, 0.3 * h, 9000)
    nall = nm + nt
    k = min(300, nall)
    pos[0][:k] = origin[0] + (np.arange(k) % 25 + 0.5) * h            # on bin edges
    m = rng.uniform(0.5, 2.0, nm)
    x = [D.to_dev(c, dev) for c in pos]
    md = D.to_dev(m, dev)
    nb = -(-n // B)
    i32 = dict(dtype=torch.int32, device=dev)
    cstart = torch.empty(nb ** 3 + 1, **i32)
    tstart = torch.empty(nb ** 3 + 1, **i32)
    order = torch.empty(max(nm, 1), **i32)
    torder = torch.empty(max(nall, 1), **i32)
    bufs = {k: torch.empty((max(nm, 1), 4), dtype=torch.float32, device=dev)
            for k in ("packed", "near", "near_lo")}
    body = torch.empty((max(nm, 1), 4), dtype=torch.float64, device=dev)
    w = D.workspace(_lib.load().ssim_split_bin_all_work_bytes(nall, n, B), dev)
    org = _lib.double_array(origin)
    _lib.call("ssim_split_bin_all", nall, nm, *[_lib.ptr(c) for c in x], _lib.ptr(md), n,
              ctypes.addressof(org), h, B, _lib.ptr(cstart), _lib.ptr(order),
              _lib.ptr(bufs["packed"]), _lib.ptr(body), _lib.ptr(bufs["near"]),
              _lib.ptr(bufs["near_lo"]), _lib.ptr(tstart), _lib.ptr(torder), _lib.ptr(w),
              w.numel(), D.stream(dev))
    _, rc, ro = p3m_np.split_bins(pos, n, origin, h, B, nodes=False)
    assert np.array_equal(tstart.cpu().numpy(), rc)
    assert np.array_equal(torder.cpu().numpy()[:nall], ro)
    ref = _split_bin_call(dev, [c[:nm] for c in pos], m, n, origin, h, B)
    assert np.array_equal(cstart.cpu().numpy(), ref["cstart"])
    assert np.array_equal(order.cpu().numpy()[:nm], ref["order"])
    assert np.array_equal(body.cpu().numpy()[:nm], ref["body"])
    for k in ("packed", "near", "near_lo"):
        assert np.array_equal(bufs[k].cpu().numpy()[:nm], ref[k])
    if case == "clump":
        assert np.diff(rc).max() > 8000                   # one dense block


def test_unified_bins_same_forces(dev):
    """The split force (TSC, overlapped streams, FP32 mesh) with the unified
    binning is bit-identical to the one with the two separate binnings: the
    mass bins are the same, and each target's near-field sum and gather do
    not depend on how the targets are grouped."""
    print("unified bins forces", flush=True)
    rng = np.random.default_rng(33)
    n, t = 20_000, 30_000
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    for c in range(3):
        pos[c][:3000] = 0.4 + rng.normal(0.0, 0.02, 3000)
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = rng.uniform(0.5, 1.5, n) / n
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    out = []
    for unified in (True, False, True):
        sf = SplitForce(dev, mesh_n=64, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0,
                        prec=P.F32, assign="tsc", margin=4.0, freeze=True)
        sf.unified_bins = unified
        x = [D.to_dev(c, dev) for c in allx]
        md = D.to_dev(m, dev)
        sf.set_mass_bound(md)
        a = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
        sf.accelerations(x, n, md, a)
        out.append(np.array([c.cpu().numpy() for c in a]))
    assert np.array_equal(out[0], out[2])
    assert np.array_equal(out[0], out[1])


def test_tiled_deposit_dense_clump(dev):
    """ssim_tsc_deposit_tiles with 30000 of 40000 masses inside one node
    (the LDS list of a tile fills and is flushed many times) and masses on
    the tile edges, against the TSC oracle."""
    import ctypes
    print("tiled deposit clump", flush=True)
    n, h, nbody, B = 37, 0.02, 40_000, 5
    origin = [0.0, 0.0, 0.0]
    rng = np.random.default_rng(31)
    pos = [rng.uniform(1.0, n - 2.0, nbody) * h for _ in range(3)]
    for c in range(3):
        pos[c][:30_000] = (17.3 + rng.uniform(-0.2, 0.2, 30_000)) * h
        pos[c][30_000:31_000] = (np.arange(1000) % 3 * 10 + 9.5) * h   # tile edges
    m = rng.uniform(0.1, 1.0, nbody)
    x = [D.to_dev(c, dev) for c in pos]
    md = D.to_dev(m, dev)
    nb = -(-n // B)
    i32 = dict(dtype=torch.int32, device=dev)
    cstart = torch.empty(nb ** 3 + 1, **i32)
    order = torch.empty(nbody, **i32)
    packed = torch.empty((nbody, 4), dtype=torch.float32, device=dev)
    body = torch.empty((nbody, 4), dtype=torch.float64, device=dev)
    w = D.workspace(_lib.load().ssim_split_bin_work_bytes(nbody, n, B), dev)
    org = _lib.double_array(origin)
    _lib.call("ssim_split_bin", nbody, *[_lib.ptr(c) for c in x], _lib.ptr(md), n,
              ctypes.addressof(org), h, B, None, _lib.ptr(cstart),
              _lib.ptr(order), _lib.ptr(packed), _lib.ptr(body), None, None, _lib.ptr(w),
              w.numel(), D.stream(dev))
    ref = p3m_np.tsc_deposit(pos, m, n, origin, h)
    _tiled_deposit_checks(dev, n, origin, h, cstart, body, m, ref)


@pytest.mark.parametrize("prec,tol,n", [(0, 1e-12, 32), (1, 1e-5, 32), (1, 1e-5, 128),
                                        (0, 1e-12, 128)])
def test_pruned_convolution_matches_padded_fft(dev, prec, tol, n):
    """ssim_pm_conv_isolated3 (pruned passes, imaginary kernel spectrum)
    equals the corner of irfftn(rfftn(pad(rho)) * i kim) in NumPy, and the
    unfused device path."""
    print(f"pruned convolution prec={prec} n={n}", flush=True)
    sf = SplitForce(dev, mesh_n=n, eps2=1e-6, rs_cells=1.25, rcut_rs=4.5, prec=prec)
    sf._rebuild_kernels(0.03)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.0, 1.0, (n, n, n))
    rho[rng.uniform(size=rho.shape) < 0.5] = 0.0
    sf.rho.copy_(torch.from_numpy(rho).to(sf.rho.dtype))
    got = [f.cpu().numpy().astype(np.float64) for f in sf.convolve()]
    if prec == 1:                       # the interleaved float4 output too
        sf.f4 = torch.full((n ** 3, 4), 9.0, dtype=torch.float32, device=dev)
        f4 = sf.convolve(interleaved=True).cpu().numpy()
        for c in range(3):
            assert np.array_equal(f4[:, c].reshape(n, n, n), got[c].astype(np.float32))
        assert np.all(f4[:, 3] == (0.0 if n == 128 else 9.0))   # whole slots at n = 128
    sf.fused = False
    ref_dev = [f.cpu().numpy().astype(np.float64) for f in sf.convolve()]
    P = 2 * n
    pad = np.zeros((P, P, P))
    pad[:n, :n, :n] = sf.rho.cpu().numpy().astype(np.float64)
    spec = np.fft.rfftn(pad)
    for c in range(3):
        kim = sf.kim[c].cpu().numpy().astype(np.float64).reshape(P, P, n + 1)
        ref = np.fft.irfftn(spec * (1j * kim), s=(P, P, P), axes=(0, 1, 2))[:n, :n, :n]
        scale = np.abs(ref).max()
        assert np.abs(got[c] - ref).max() <= tol * scale
        assert np.abs(ref_dev[c] - ref).max() <= tol * scale


def test_near_field_matches_oracle(dev):
    n, t = 6000, 3000
    pos, m = cloud(n, 2)
    tpos, _ = cloud(t, 3)
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    sf = SplitForce(dev, mesh_n=32, eps2=1e-6, rs_cells=1.25, rcut_rs=4.5)
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    a = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.mesh_accel(x, n, md, a)            # places the grid (h, box)
    for c in a:
        c.zero_()
    sf.near_accel(x, n, md, a)
    rs = sf.rs_cells * sf.h
    ref = p3m_np.near_field(pos, m, allx, 1.0, 1e-6, rs, sf.rcut_rs * rs)
    got = np.array([c.cpu().numpy() for c in a])
    err = np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0)
    # FP32 terms on origin-relative FP32 positions: ~1e-6 per body
    assert np.percentile(err, 99) < 5e-5 and np.median(err) < 5e-6


def _tsc_split(dev, pos, m, tpos, mesh_n=32):
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    sf = SplitForce(dev, mesh_n=mesh_n, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0,
                    prec=P.F32, assign="tsc", margin=4.0, freeze=True)
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    sf.set_mass_bound(md)
    sf.place(x)
    sf.bin_bodies(x, len(m), md)
    return sf, x, allx


@pytest.mark.parametrize("clump", [False, True])
def test_near_precise_matches_oracle(dev, clump):
    """The near field on block-relative coordinates (ssim_pp_near) against
    the exact oracle sum (p3m_np.near_field): median of the per-body
    relative error <= 1e-6 and p99 <= 5e-6 (FP32 pair terms, 4096-interval
    table, FP64 sums over the cells' partials), at least 5x below the
    origin-relative kernel's p99; bit-identical when run twice.  The clump
    puts 4000 of the 12,000 masses in a ball well inside one cell."""
    print(f"near precise clump={clump}", flush=True)
    n, t = 12000, 6000
    rng = np.random.default_rng(21)
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    if clump:
        for c in range(3):
            pos[c][:4000] = 0.5 + rng.normal(0.0, 0.01, 4000)
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = rng.uniform(0.5, 1.5, n) / n
    sf, x, allx = _tsc_split(dev, pos, m, tpos)
    sf.near_bits = False          # the cell walk (ssim_pp_near)
    rs = sf.rs_cells * sf.h
    ref = p3m_np.near_field(pos, m, allx, 1.0, 1e-6, rs, sf.rcut_rs * rs)
    runs = []
    for _ in range(2):
        a = [torch.full((n + t,), 3.0, dtype=torch.float64, device=dev) for _ in range(3)]
        sf.near_sum(x, a, accumulate=False)
        runs.append(np.array([c.cpu().numpy() for c in a]))
    assert np.array_equal(runs[0], runs[1])
    err = np.linalg.norm(runs[0] - ref, axis=0) / np.linalg.norm(ref, axis=0)
    sf.near_precise = False
    a = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a, accumulate=False)
    sf.near_precise = True
    old = np.array([c.cpu().numpy() for c in a])
    err_old = np.linalg.norm(old - ref, axis=0) / np.linalg.norm(ref, axis=0)
    print(f"near precise: median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} "
          f"max {err.max():.2e}; origin-relative p99 {np.percentile(err_old, 99):.2e}")
    assert np.median(err) <= 1e-6 and np.percentile(err, 99) <= 5e-6
    assert np.percentile(err, 99) * 5 <= np.percentile(err_old, 99)


def _near_time(sf, x, a, reps=20):
    """Average ms of one near_sum (targets already binned)."""
    st = torch.cuda.current_stream(x[0].device)
    sf.near_sum(x, a, accumulate=False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        sf.near_sum(x, a, accumulate=False)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


@pytest.mark.parametrize("clump", [False, True])
def test_near_hilo_matches_oracle(dev, clump):
    """The hi/lo near field (ssim_pp_near_hilo: masses as FP32 hi + lo
    pairs from ssim_split_bin, targets from FP64, pair arithmetic and sums
    in FP64, the table interpolated quadratically) against the exact oracle
    sum (p3m_np.near_field): p99 of the per-body relative error <= 1e-6,
    median <= 1e-7, at least 5x below the plain block-relative kernel's
    p99; bit-identical when run twice.  Prints both kernels' times on the
    same lists (its cost)."""
    print(f"near hilo clump={clump}", flush=True)
    n, t = 12000, 6000
    rng = np.random.default_rng(21)
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    if clump:
        for c in range(3):
            pos[c][:4000] = 0.5 + rng.normal(0.0, 0.01, 4000)
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = rng.uniform(0.5, 1.5, n) / n
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    sf = SplitForce(dev, mesh_n=32, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0,
                    prec=P.F32, assign="tsc", margin=4.0, freeze=True)
    sf.near_hilo, sf.near_bits = True, False   # the cell walk's hi/lo kernel
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    sf.set_mass_bound(md)
    sf.place(x)
    sf.bin_bodies(x, n, md)
    rs = sf.rs_cells * sf.h
    ref = p3m_np.near_field(pos, m, allx, 1.0, 1e-6, rs, sf.rcut_rs * rs)
    runs = []
    for _ in range(2):
        a = [torch.full((n + t,), 3.0, dtype=torch.float64, device=dev) for _ in range(3)]
        sf.near_sum(x, a, accumulate=False)
        runs.append(np.array([c.cpu().numpy() for c in a]))
    assert np.array_equal(runs[0], runs[1])
    err = np.linalg.norm(runs[0] - ref, axis=0) / np.linalg.norm(ref, axis=0)
    t_hilo = _near_time(sf, x, a)
    sf.near_hilo = False
    a = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a, accumulate=False)
    plain = np.array([c.cpu().numpy() for c in a])
    t_plain = _near_time(sf, x, a)
    err_plain = np.linalg.norm(plain - ref, axis=0) / np.linalg.norm(ref, axis=0)
    print(f"near hilo: median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} "
          f"max {err.max():.2e} {t_hilo * 1e3:.1f} us; plain p99 "
          f"{np.percentile(err_plain, 99):.2e} {t_plain * 1e3:.1f} us")
    assert np.median(err) <= 1e-7 and np.percentile(err, 99) <= 1e-6
    assert np.percentile(err, 99) * 5 <= np.percentile(err_plain, 99)


@pytest.mark.parametrize("clump", [False, True])
@pytest.mark.parametrize("precise", [False, True])
def test_near_bits_matches_oracle(dev, clump, precise):
    """The filtered near field (ssim_pp_near_bits: FP16 MFMA distance
    filter, per-lane sums over the set bits, FP64 positions) against the
    exact oracle sum (p3m_np.near_field), per-body relative error: FP64
    positions and terms (precise) p99 <= 1e-6 and median <= 1e-7; FP32
    positions relative to the cell and FP32 terms the plain kernel's p99 <=
    5e-6 and median <= 1e-6; bit-identical when run twice.  The clumped case puts 4000 masses in a ball of sigma 0.01
    (several batches per chunk).  Prints the kernels' times."""
    print(f"near bits clump={clump} precise={precise}", flush=True)
    n, t = 12000, 6000
    rng = np.random.default_rng(23)
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    if clump:
        for c in range(3):
            pos[c][:4000] = 0.5 + rng.normal(0.0, 0.01, 4000)
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = rng.uniform(0.5, 1.5, n) / n
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    sf = SplitForce(dev, mesh_n=32, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0,
                    prec=P.F32, assign="tsc", margin=4.0, freeze=True)
    sf.near_bits, sf.near_hilo = True, precise
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    sf.set_mass_bound(md)
    sf.place(x)
    sf.bin_bodies(x, n, md)
    rs = sf.rs_cells * sf.h
    ref = p3m_np.near_field(pos, m, allx, 1.0, 1e-6, rs, sf.rcut_rs * rs)
    runs = []
    for _ in range(2):
        a = [torch.full((n + t,), 3.0, dtype=torch.float64, device=dev) for _ in range(3)]
        sf.near_sum(x, a, accumulate=False)
        runs.append(np.array([c.cpu().numpy() for c in a]))
    assert np.array_equal(runs[0], runs[1])
    err = np.linalg.norm(runs[0] - ref, axis=0) / np.linalg.norm(ref, axis=0)
    t_bits = _near_time(sf, x, a)
    sf.near_bits, sf.near_hilo = False, True
    t_hilo = _near_time(sf, x, a)
    sf.near_hilo = False
    t_plain = _near_time(sf, x, a)
    print(f"near bits: median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} "
          f"max {err.max():.2e} {t_bits * 1e3:.1f} us; hilo {t_hilo * 1e3:.1f} us, "
          f"plain {t_plain * 1e3:.1f} us")
    if precise:
        assert np.median(err) <= 1e-7 and np.percentile(err, 99) <= 1e-6
    else:
        # FP32 positions relative to the cell: the plain kernel's bounds
        assert np.median(err) <= 1e-6 and np.percentile(err, 99) <= 5e-6


def test_near_bits_accumulates(dev):
    """ssim_pp_near_bits with accumulate = 1 adds its sum to what is there,
    the same sum as a STORE."""
    rng = np.random.default_rng(24)
    n, t = 3000, 1000
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = np.full(n, 1.0 / n)
    sf, x, _ = _tsc_split(dev, pos, m, tpos)
    sf.near_bits = True
    sf.bin_bodies(x, n, D.to_dev(m, dev))
    a0 = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a0, accumulate=False)
    a1 = [torch.full((n + t,), 0.25, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a1, accumulate=True)
    for c in range(3):
        assert torch.equal(a1[c], a0[c] + 0.25)
    assert float(a0[0].abs().max()) > 0.0


def test_near_precise_accumulates(dev):
    """accumulate = 1 adds the near field to what is there."""
    rng = np.random.default_rng(22)
    n, t = 3000, 1000
    pos = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    tpos = [rng.uniform(0.0, 1.0, t) for _ in range(3)]
    m = np.full(n, 1.0 / n)
    sf, x, _ = _tsc_split(dev, pos, m, tpos)
    sf.near_bits = False          # the cell walk (ssim_pp_near)
    a0 = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a0, accumulate=False)
    a1 = [torch.full((n + t,), 0.25, dtype=torch.float64, device=dev) for _ in range(3)]
    sf.near_sum(x, a1, accumulate=True)
    for c in range(3):
        assert torch.equal(a1[c], a0[c] + 0.25)


def test_split_force_accuracy_perturbed_cube(dev):
    """Mesh + near field against the direct FP64 sum on a 64k-mass /
    64k-tracer perturbed cube: median relative force error < 1e-2 (the
    learned correction and its tests take it further)."""
    pos, m, tpos = scenes.learned_scene((40, 40, 40), (40, 40, 40), seed=5,
                                        tracer_seed=55)
    n = len(m)
    allx = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    x = [D.to_dev(c, dev) for c in allx]
    md = D.to_dev(m, dev)
    sf = SplitForce(dev, mesh_n=64, eps2=1e-6)
    a = [torch.zeros(len(allx[0]), dtype=torch.float64, device=dev) for _ in range(3)]
    sf.accelerations(x, n, md, a)
    idx = np.random.default_rng(0).choice(len(allx[0]), 2000, replace=False)
    ref = oracle.allpairs(pos, m, [c[idx] for c in allx], 1.0, 1e-6)
    err = relative_errors([c.cpu().numpy()[idx] for c in a], ref)
    print(f"split force median {np.median(err):.3e} p90 {np.percentile(err, 90):.3e}")
    assert np.median(err) < 1e-2
