"""GPU parity of the raster integrator through the NBodyCuda drop-in:
against the serial oracle (same algorithm; FP64 state within 1e-12 of it,
identical sub-sample sets and contacts) and the reference's own tests
(test_kernel.py, test_subsamples.py, test_bewegter_stern.py)."""
import numpy as np
import pytest

import oracle
from oracle import nbody_np as nb
from solarsystemsimulator_amd.nbody import NBodyCuda
from tests import nbody_cases as C

pytestmark = pytest.mark.gpu
DAY = 1 / 365.25


@pytest.mark.parametrize("n_ast,k,raster", [(0, 2, 5 * DAY), (200, 1, 5 * DAY),
                                            (3000, 4, 0.5 * DAY),
                                            (20000, 3, 2 * DAY)])
def test_matches_oracle(dev, n_ast, k, raster):
    print(f"nbody n_ast={n_ast} k={k}", flush=True)
    st0 = list(nb.make_system(n_ast, seed=11))
    if n_ast:                                          # a few divers
        st0[0][10:30] = 0.3 * np.cos(np.arange(20.0))
        st0[1][10:30] = 0.3 * np.sin(np.arange(20.0))
        st0[2][10:30] = 0.0
        st0[3][10:30] = 0.0
    rr = np.full(len(st0[0]), 1e-5)
    rr[0] = 0.00465
    sim = NBodyCuda(0)
    st = sim.load_state(*st0, real_r=rr)
    out = sim.step_batch(st, raster, k)
    ref, ref_state, ref_sub, ref_hits = oracle.nbody_frames(*st0, rr, raster, k)
    got_state = sim.export_f64(st)
    scale = np.abs(ref_state).max()
    assert np.abs(got_state - ref_state).max() <= 1e-12 * scale
    assert np.abs(out - ref).max() <= 1e-6 * np.abs(ref).max()
    for (gi, gp), (ri, rp) in zip(st["sub"], ref_sub):
        assert np.array_equal(gi, ri)
        np.testing.assert_allclose(gp, rp, rtol=1e-6, atol=1e-7)
    assert np.array_equal(st["hits"][0], ref_hits[0])
    assert np.array_equal(st["hits"][2], ref_hits[2])
    np.testing.assert_allclose(st["hits"][1], ref_hits[1], rtol=1e-6, atol=1e-9)


def test_reference_physics_and_frozen_body(dev):
    """test_kernel.py:116-131."""
    sim = NBodyCuda(0)
    state = nb.make_system(200)
    n = len(state[0])
    st = sim.load_state(*state)
    out = sim.step(st, 5 * DAY)
    rx, ry, rvx, rvy = nb.reference_advance(5 * DAY, *state)
    assert np.abs(np.concatenate([out[:n] - rx, out[n:2 * n] - ry])).max() < 5e-6
    assert np.abs(np.concatenate([out[2 * n:3 * n] - rvx, out[3 * n:] - rvy])).max() < 5e-6
    assert abs(out[3] - state[0][3]) < 1e-5 and abs(out[n + 3] - state[1][3]) < 1e-5


def test_residency_and_delta_updates(dev):
    """test_kernel.py:133-155: 73 resident 5-day steps keep a planet on its
    orbit; point updates hit the right slots."""
    sim = NBodyCuda(0)
    st = sim.load_state(*nb.make_system(0))
    for _ in range(73):
        out = sim.step(st, 5 * DAY)
    n = 10
    # planet 3 is frozen (vis = 0) at 3 AU; the sun wanders only slightly
    assert abs(np.hypot(out[3] - out[0], out[n + 3] - out[n]) - 3.0) < 0.05
    state = nb.make_system(5)
    st = sim.load_state(*state)
    n = len(state[0])
    sim.apply_updates(st, np.array([12, 2]), np.array([[9.0, -9.0, 0.0, 0.0],
                                                       [4.0, 4.5, 0.0, 0.0]]))
    out = sim.step(st, 1e-9)
    assert abs(out[12] - 9.0) < 1e-4 and abs(out[n + 12] + 9.0) < 1e-4
    assert abs(out[2] - 4.0) < 1e-4 and abs(out[n + 2] - 4.5) < 1e-4
    sim.apply_deltas(st, np.array([12]), np.array([[1.0, 0.0, 0.0, 0.0]]))
    sim.deactivate_body(st, 13)
    x13 = float(sim.export_f64(st)[13])
    out = sim.step(st, 1 * DAY)
    assert abs(out[12] - 10.0) < 1e-2
    assert out[13] == np.float32(x13)


def test_batch_equals_singles(dev):
    """test_kernel.py:157-167."""
    sim = NBodyCuda(0)
    state = nb.make_system(500)
    st1 = sim.load_state(*state)
    st2 = sim.load_state(*state)
    batch = sim.step_batch(st1, 0.5 * DAY, 4)
    singles = [sim.step(st2, 0.5 * DAY) for _ in range(4)]
    assert max(np.abs(batch[i] - singles[i]).max() for i in range(4)) < 1e-6


def test_divers(dev):
    """test_kernel.py:190-219."""
    sim = NBodyCuda(0)
    x, y, vx, vy, mass, vis, is_ast = C.divers()
    st = sim.load_state(x, y, vx, vy, mass, vis, is_ast)
    out = sim.step(st, 8 * DAY)
    rx, ry, _, _ = nb.reference_advance(8 * DAY, x, y, vx, vy, mass, vis, is_ast)
    n1 = len(x)
    assert np.hypot(out[:n1] - rx, out[n1:2 * n1] - ry).max() < 2e-3
    assert np.hypot(out[1:n1] - out[0], out[n1 + 1:2 * n1] - out[n1]).max() < 0.5


def test_sub_samples(dev):
    """test_subsamples.py:46-97."""
    st0, diver = C.sub_sample_scene()
    sim = NBodyCuda(0, m_sub=8)
    st = sim.load_state(*st0)
    sim.step_batch(st, 0.5 * DAY, 1)
    idx, pos = st["sub"][0]
    assert 1 in idx.tolist() and 2 not in idx.tolist()
    col = idx.tolist().index(1)
    track = C.fine_orbit(*diver, 1.0, 0.5 * DAY, 200_000)
    for j in range(8):
        rx, ry = track[int(round((j + 1) / 8 * 200_000)) - 1]
        assert np.hypot(pos[j, 0, col] - rx, pos[j, 1, col] - ry) < 1e-4


@pytest.mark.parametrize("v_star", [0.0, 0.5, 2.0, 5.0])
def test_moving_star(dev, v_star):
    """test_bewegter_stern.py."""
    sim = NBodyCuda(0)
    st = sim.load_state(*C.moving_star(v_star))
    out = sim.step_batch(st, 0.5 * DAY, 16)
    rx, ry = C.moving_star_reference(v_star, 8 * DAY, 400_000)
    assert np.hypot(out[-1][1] - rx, out[-1][3] - ry) < 1e-3


def test_errors(dev):
    with pytest.raises(ValueError):
        NBodyCuda(0, m_sub=33)
    sim = NBodyCuda(0)
    n = 70
    with pytest.raises(ValueError):
        sim.load_state(np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n),
                       np.ones(n), np.ones(n, np.uint8), np.zeros(n, np.uint8))
    st = sim.load_state(*nb.make_system(3))
    with pytest.raises(ValueError):
        sim.step_batch(st, 0.01, 17)
    with pytest.raises(ValueError):
        sim.step_batch(None, 0.01, 1)
