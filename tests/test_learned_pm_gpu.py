"""Time stepping of the PM + learned method: the exact box of the next
drift (the mesh placement is made from it before a step), and steps
replayed from a captured HIP graph against steps launched one by one."""
import numpy as np
import pytest
import torch

from solarsystemsimulator_amd import _lib, scenes
from solarsystemsimulator_amd import device as D
from solarsystemsimulator_amd import pm as P
from solarsystemsimulator_amd.learned_pm import PMLearnedSystem, load_weights

pytestmark = pytest.mark.gpu


def _ptrs(arrs):
    return [_lib.ptr(c) for c in arrs]


def test_box_after_kdk_is_exact(dev):
    print("box after kick + drift", flush=True)
    rng = np.random.default_rng(3)
    n = 100_003
    x = [rng.normal(0.0, 1.0, n) for _ in range(3)]
    v = [rng.normal(0.0, 10.0, n) for _ in range(3)]
    a = [rng.normal(0.0, 100.0, n) for _ in range(3)]
    half, full = 1.25e-3, 2.5e-3
    xd, vd, ad = ([D.to_dev(c, dev) for c in arr] for arr in (x, v, a))
    out = torch.empty(6, dtype=torch.float64, device=dev)
    w = D.workspace(_lib.load().ssim_minmax_work_bytes(n), dev)
    st = D.stream(dev)
    _lib.call("ssim_box_after_kdk", 3, n, *_ptrs(xd), *_ptrs(vd), *_ptrs(ad), half,
              full, _lib.ptr(out), _lib.ptr(w), w.numel(), st)
    got = out.cpu().numpy()
    # the kick and drift kernels themselves, then the box of their result
    _lib.call("ssim_kick", 3, n, *_ptrs(vd), *_ptrs(ad), half, st)
    _lib.call("ssim_drift", 3, n, *_ptrs(xd), *_ptrs(vd), full, st)
    for c in range(3):
        p = xd[c].cpu().numpy()
        assert got[2 * c] == p.min() and got[2 * c + 1] == p.max()
        # NumPy's arithmetic (no fused multiply-add) gives the same positions
        assert np.array_equal(x[c] + (v[c] + a[c] * half) * full, p)


def test_fused_kicks_match_separate_launches(dev):
    """ssim_kick_drift = ssim_kick + ssim_drift and ssim_kick_box = ssim_kick
    + ssim_box_after_kdk, bit for bit, on two calls in a row with a work
    buffer that starts as garbage (no initial contents needed)."""
    print("fused kicks", flush=True)
    rng = np.random.default_rng(4)
    n = 70_001
    st = D.stream(dev)
    lib = _lib.load()
    half, full = 1.25e-3, 2.5e-3
    for call in range(2):
        x = [rng.normal(0.0, 1.0, n) for _ in range(3)]
        v = [rng.normal(0.0, 10.0, n) for _ in range(3)]
        a = [rng.normal(0.0, 100.0, n) for _ in range(3)]
        xa, va, aa = ([D.to_dev(c, dev) for c in arr] for arr in (x, v, a))
        xb, vb = [c.clone() for c in xa], [c.clone() for c in va]
        _lib.call("ssim_kick_drift", 3, n, *_ptrs(xa), *_ptrs(va), *_ptrs(aa), half, full,
                  st)
        _lib.call("ssim_kick", 3, n, *_ptrs(vb), *_ptrs(aa), half, st)
        _lib.call("ssim_drift", 3, n, *_ptrs(xb), *_ptrs(vb), full, st)
        for c in range(3):
            assert torch.equal(xa[c], xb[c]) and torch.equal(va[c], vb[c])
        if call == 0:
            kw = torch.full((lib.ssim_kick_box_work_bytes(n),), 0xAB, dtype=torch.uint8,
                            device=dev)
        out_a = torch.empty(6, dtype=torch.float64, device=dev)
        out_b = torch.empty(6, dtype=torch.float64, device=dev)
        _lib.call("ssim_kick_box", 3, n, *_ptrs(xa), *_ptrs(va), *_ptrs(aa), half, full,
                  _lib.ptr(out_a), _lib.ptr(kw), kw.numel(), st)
        _lib.call("ssim_kick", 3, n, *_ptrs(vb), *_ptrs(aa), half, st)
        w = D.workspace(lib.ssim_minmax_work_bytes(n), dev)
        _lib.call("ssim_box_after_kdk", 3, n, *_ptrs(xb), *_ptrs(vb), *_ptrs(aa), half,
                  full, _lib.ptr(out_b), _lib.ptr(w), w.numel(), st)
        for c in range(3):
            assert torch.equal(va[c], vb[c])
        assert torch.equal(out_a, out_b)
        # ssim_kick_drift_box = ssim_kick_box then ssim_kick_drift (state),
        # and its box is the one ssim_kick_box gives for that drift
        xc, vc = [c.clone() for c in xa], [c.clone() for c in va]
        out_c = torch.empty(6, dtype=torch.float64, device=dev)
        _lib.call("ssim_kick_drift_box", 3, n, *_ptrs(xc), *_ptrs(vc), *_ptrs(aa), half,
                  full, _lib.ptr(out_c), _lib.ptr(kw), kw.numel(), st)
        _lib.call("ssim_kick_box", 3, n, *_ptrs(xa), *_ptrs(va), *_ptrs(aa), half, full,
                  _lib.ptr(out_a), _lib.ptr(kw), kw.numel(), st)
        _lib.call("ssim_kick_drift", 3, n, *_ptrs(xa), *_ptrs(va), *_ptrs(aa), half, full,
                  st)
        for c in range(3):
            assert torch.equal(xa[c], xc[c]) and torch.equal(va[c], vc[c])
        assert torch.equal(out_a, out_c)


def _system(dev, pos, m, tpos, vel, tvel):
    s = PMLearnedSystem(dev, len(m), len(tpos[0]), 1e-6, 1.0,
                        net=load_weights(dev), mesh_n=32)
    s.set_bodies(pos, vel, m, tpos, tvel)
    return s


def test_replayed_steps_equal_eager_steps(dev):
    print("graph-replayed steps vs eager steps", flush=True)
    pos, m, tpos = scenes.learned_scene((20, 20, 10), (20, 20, 20), seed=7,
                                        tracer_seed=77)
    n, t = len(m), len(tpos[0])
    # a bulk drift of about one mesh cell per step (cell ~ 1/23 of the
    # cube): the kept placement runs out every few steps, so the run
    # re-places the mesh and captures the step again several times
    dt, vb = 2.5e-4, 170.0
    vel = [np.full(n, vb), np.zeros(n), np.zeros(n)]
    tvel = [np.full(t, vb), np.zeros(t), np.zeros(t)]
    eager = _system(dev, pos, m, tpos, vel, tvel)
    replay = _system(dev, pos, m, tpos, vel, tvel)
    eager.run(dt, 12, graph=False)
    replay.run(dt, 12, graph=True)
    assert eager.captures == 0 and replay.captures >= 3
    for c in range(3):
        assert torch.equal(eager.x[c], replay.x[c])
        assert torch.equal(eager.v[c], replay.v[c])
        assert torch.equal(eager.a[c], replay.a[c])
    # the eager run re-placed the mesh the same way
    assert eager.split.placement() == replay.split.placement()


@pytest.mark.parametrize("graph", [True, False])
def test_fused_step_ends_equal_single_steps(dev, graph):
    """One run of 6 steps (each step but the last ends with the next one's
    opening kick and drift, ssim_kick_drift_box) equals 6 runs of one step
    (each opens with ssim_kick_drift and closes with ssim_kick_box), bit
    for bit."""
    print(f"fused step ends graph={graph}", flush=True)
    pos, m, tpos = scenes.learned_scene((20, 20, 10), (20, 20, 20), seed=9,
                                        tracer_seed=99)
    n, t = len(m), len(tpos[0])
    rng = np.random.default_rng(10)
    vel = [rng.normal(0.0, 2.0, n) for _ in range(3)]
    tvel = [rng.normal(0.0, 2.0, t) for _ in range(3)]
    one = _system(dev, pos, m, tpos, vel, tvel)
    many = _system(dev, pos, m, tpos, vel, tvel)
    one.run(2.5e-4, 6, graph=graph)
    for _ in range(6):
        many.run(2.5e-4, 1, graph=graph)
    for c in range(3):
        assert torch.equal(one.x[c], many.x[c])
        assert torch.equal(one.v[c], many.v[c])
        assert torch.equal(one.a[c], many.a[c])


@pytest.mark.parametrize("overlap", [True, False])
def test_folded_correction_equals_separate_update(dev, overlap):
    """The learned correction added inside the mesh gather (network beside
    the convolution) equals the separate update after it, bit for bit,
    with moving masses (the velocity projection) and tracers."""
    print(f"folded correction overlap={overlap}", flush=True)
    pos, m, tpos = scenes.learned_scene((20, 20, 10), (20, 20, 20), seed=8,
                                        tracer_seed=88)
    n, t = len(m), len(tpos[0])
    rng = np.random.default_rng(9)
    vel = [rng.normal(0.0, 0.3, n) for _ in range(3)]
    tvel = [rng.normal(0.0, 0.3, t) for _ in range(3)]
    s = _system(dev, pos, m, tpos, vel, tvel)
    s.split.overlap = overlap
    s.fold = True
    folded = [c.clone() for c in s.force(learned=True)]
    s.fold = False
    separate = [c.clone() for c in s.force(learned=True)]
    plain = [c.clone() for c in s.force(learned=False)]
    for c in range(3):
        assert torch.equal(folded[c], separate[c])
        assert not torch.equal(folded[c], plain[c])


@pytest.mark.parametrize("graph", [True, False])
def test_gather_box_steps_equal_step_box(dev, graph):
    """Steps whose gather writes the next box (the host waits for it and
    not for the kicks and drift) equal steps whose closing kick writes it,
    bit for bit, with the same mesh placements."""
    print(f"gather box steps graph={graph}", flush=True)
    pos, m, tpos = scenes.learned_scene((20, 20, 10), (20, 20, 20), seed=11,
                                        tracer_seed=111)
    n, t = len(m), len(tpos[0])
    rng = np.random.default_rng(12)
    vel = [rng.normal(0.0, 2.0, n) for _ in range(3)]
    tvel = [rng.normal(0.0, 2.0, t) for _ in range(3)]
    early = _system(dev, pos, m, tpos, vel, tvel)
    late = _system(dev, pos, m, tpos, vel, tvel)
    assert early._gather_box(True)
    late._gather_box = lambda learned: False
    early.run(2.5e-4, 7, graph=graph)
    late.run(2.5e-4, 7, graph=graph)
    for c in range(3):
        assert torch.equal(early.x[c], late.x[c])
        assert torch.equal(early.v[c], late.v[c])
        assert torch.equal(early.a[c], late.a[c])
    assert early.split.placement() == late.split.placement()


def test_gather_box_matches_kick_box(dev):
    """ssim_tsc_gather_box: the accelerations of ssim_tsc_gather_corrected
    and of ssim_tsc_gather (no correction), and the box ssim_kick_drift_box
    gives for the state they leave, bit for bit."""
    print("gather box C-ABI", flush=True)
    pos, m, tpos = scenes.learned_scene((20, 20, 10), (20, 20, 20), seed=13,
                                        tracer_seed=113)
    n, t = len(m), len(tpos[0])
    nb = n + t
    rng = np.random.default_rng(14)
    vel = [rng.normal(0.0, 2.0, n) for _ in range(3)]
    tvel = [rng.normal(0.0, 2.0, t) for _ in range(3)]
    s = _system(dev, pos, m, tpos, vel, tvel)
    s.force(learned=True)                 # mesh fields, bins, network outputs
    sf = s.split
    lib = _lib.load()
    st = D.stream(dev)
    org = _lib.double_array(sf.origin)
    import ctypes
    corr = torch.from_numpy(rng.normal(0.0, 1.0, (nb, 3)).astype(np.float32)).to(dev)
    base = [D.to_dev(rng.normal(0.0, 1.0, nb), dev) for _ in range(3)]
    half, full = 1.25e-4, 2.5e-4
    gw = D.workspace(lib.ssim_tsc_gather_box_work_bytes(nb), dev)
    kw = D.workspace(lib.ssim_kick_box_work_bytes(nb), dev)
    for use_corr in (True, False):
        ref = [b.clone() for b in base]
        got = [b.clone() for b in base]
        if use_corr:
            _lib.call("ssim_tsc_gather_corrected", nb, _lib.ptr(sf.torder),
                      *_ptrs(s.x), sf.n, ctypes.addressof(org), sf.h, _lib.ptr(sf.f4),
                      *_ptrs(ref), 1, _lib.ptr(corr), 0.01, 1, *_ptrs(s.v), n, st)
        else:
            _lib.call("ssim_tsc_gather", nb, _lib.ptr(sf.torder), *_ptrs(s.x), sf.n,
                      ctypes.addressof(org), sf.h, P.F32, _lib.ptr(sf.f4), None, None,
                      *_ptrs(ref), 1, st)
        box = torch.empty(6, dtype=torch.float64, device=dev)
        _lib.call("ssim_tsc_gather_box", nb, _lib.ptr(sf.torder), *_ptrs(s.x), sf.n,
                  ctypes.addressof(org), sf.h, _lib.ptr(sf.f4), *_ptrs(got), 1,
                  _lib.ptr(corr) if use_corr else None, 0.01, 1, *_ptrs(s.v), n, half,
                  full, _lib.ptr(box), _lib.ptr(gw), gw.numel(), st)
        for c in range(3):
            assert torch.equal(got[c], ref[c]), (use_corr, c)
        xk, vk = [c.clone() for c in s.x], [c.clone() for c in s.v]
        kbox = torch.empty(6, dtype=torch.float64, device=dev)
        _lib.call("ssim_kick_drift_box", 3, nb, *_ptrs(xk), *_ptrs(vk), *_ptrs(got), half,
                  full, _lib.ptr(kbox), _lib.ptr(kw), kw.numel(), st)
        assert torch.equal(box, kbox), use_corr
