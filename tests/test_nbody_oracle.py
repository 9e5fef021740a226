"""Pin the raster-integrator oracle (CPU only) with the reference's own
nbody tests: test_kernel.py (physics vs reference_advance, frozen body,
batch == single steps, divers), test_subsamples.py, test_bewegter_stern.py."""
import numpy as np
import pytest

import oracle
from oracle import nbody_np as nb
from tests import nbody_cases as C

DAY = 1 / 365.25


def test_frame_matches_reference_advance():
    """test_kernel.py:116-131: 200 asteroids, one 5-day frame, |dpos|,
    |dvel| < 5e-6 on the f32 output, the frozen planet does not move."""
    st = nb.make_system(200)
    n = len(st[0])
    out, _, _, _ = oracle.nbody_frames(*st, None, 5 * DAY, 1)
    rx, ry, rvx, rvy = nb.reference_advance(5 * DAY, *st)
    o = out[0]
    assert np.abs(np.concatenate([o[:n] - rx, o[n:2 * n] - ry])).max() < 5e-6
    assert np.abs(np.concatenate([o[2 * n:3 * n] - rvx, o[3 * n:] - rvy])).max() < 5e-6
    assert abs(o[3] - st[0][3]) < 1e-5 and abs(o[n + 3] - st[1][3]) < 1e-5


def test_batch_equals_single_steps():
    """test_kernel.py:157-167: K=4 in one call == 4 calls of one raster."""
    st = nb.make_system(500)
    raster = 0.5 * DAY
    batch, _, _, _ = oracle.nbody_frames(*st, None, raster, 4)
    cur = list(st)
    n = len(st[0])
    for k in range(4):
        out, state, _, _ = oracle.nbody_frames(*cur, None, raster, 1)
        assert np.abs(batch[k] - out[0]).max() < 1e-6
        cur[0:4] = [state[f * n:(f + 1) * n] for f in range(4)]


def test_divers_stay_on_track():
    """test_kernel.py:190-219: 20 bodies fall into the sun for 8 days:
    within 2e-3 AU of the all-fine reference, none catapulted beyond 0.5 AU."""
    st = C.divers()
    out, _, _, _ = oracle.nbody_frames(*st, None, 8 * DAY, 1)
    rx, ry, _, _ = nb.reference_advance(8 * DAY, *st)
    n = len(st[0])
    o = out[0]
    assert np.hypot(o[:n] - rx, o[n:2 * n] - ry).max() < 2e-3
    assert np.hypot(o[1:n] - o[0], o[n + 1:2 * n] - o[n]).max() < 0.5


def test_sub_samples_on_the_true_track():
    """test_subsamples.py: only the diver gets a complete sub-track, whose
    8 points lie within 1e-4 AU of a 200k-step Verlet reference."""
    st, diver = C.sub_sample_scene()
    raster = 0.5 * DAY
    _, _, sub, _ = oracle.nbody_frames(*st, None, raster, 1, m_sub=8)
    idx, pos = sub[0]
    assert 1 in idx.tolist() and 2 not in idx.tolist()
    assert pos.shape[:2] == (8, 2)
    col = idx.tolist().index(1)
    track = C.fine_orbit(*diver, 1.0, raster, 200_000)
    for j in range(8):
        rx, ry = track[int(round((j + 1) / 8 * 200_000)) - 1]
        assert np.hypot(pos[j, 0, col] - rx, pos[j, 1, col] - ry) < 1e-4


@pytest.mark.parametrize("v_star", [0.0, 0.5, 2.0, 5.0])
def test_moving_star_close_pass(v_star):
    """test_bewegter_stern.py: 16 rasters of 0.5 d past a moving star,
    within 1e-3 AU of a 400k-step reference."""
    st = C.moving_star(v_star)
    out, _, _, _ = oracle.nbody_frames(*st, None, 0.5 * DAY, 16)
    rx, ry = C.moving_star_reference(v_star, 8 * DAY, 400_000)
    assert np.hypot(out[-1][1] - rx, out[-1][3] - ry) < 1e-3


def test_contact_is_reported():
    """The fast diver of test_subsamples.py passes 0.004 AU from the sun's
    centre: with the solar radius as contact radius it reports a contact
    with partner 0 at the time it crosses r = 0.00465 AU (~0.15/160 yr),
    the far asteroid none."""
    st, _ = C.sub_sample_scene()
    rr = np.array([0.00465, 0.0, 0.0])
    _, _, _, hits = oracle.nbody_frames(*st, rr, 0.5 * DAY, 1)
    idx, t, partner = hits
    assert idx.tolist() == [1] and partner.tolist() == [0]
    assert abs(t[0] - (0.15 - 0.00465) / 160.0) < 2e-5
