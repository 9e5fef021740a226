"""GPU collision search: the ordered pair list and candidate count are
bit-exact against the oracle restatement of `_hits`, for whole scenes and
for stripes, through the Erkennungskarte drop-in."""
import numpy as np
import pytest
import torch

from oracle import collide_np as C
from solarsystemsimulator_amd.detection import Erkennungskarte, bounce_suche
from tests import collide_cases as S

pytestmark = pytest.mark.gpu


def fields(dev, *arrs):
    return tuple(torch.as_tensor(np.asarray(a, np.float32)).to(dev) for a in arrs)


def scenes():
    out = [(name, x, y, vx, vy, rr) for name, x, y, vx, vy, rr in S.stripe_scenes()]
    out.append(("cloud50k", *S.asteroid_cloud()))
    return out


@pytest.mark.parametrize("scene", scenes(), ids=lambda s: s[0])
def test_pair_list_bit_exact(dev, scene):
    name, x, y, vx, vy, rr = scene
    print(f"collide scene {name}", flush=True)
    n = len(x)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    karte = Erkennungskarte(0, ones != 0, -np.inf, np.inf)
    karte.stammdaten(ones, rr.astype(np.float32))
    got, cand = karte.bounce_hits(*fields(dev, x, y, vx, vy), S.DT_Y, h, halo)
    ref, rcand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    assert cand == rcand
    if ref is None:
        assert got is None
    else:
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    # stripes: each stripe's list equals the oracle's for that stripe
    cuts = np.percentile(np.float32(x), [50.0])
    for lo, hi in ((-np.inf, float(cuts[0])), (float(cuts[0]), np.inf)):
        karte.streifen(lo, hi)
        g, c = karte.bounce_hits(*fields(dev, x, y, vx, vy), S.DT_Y, h, halo)
        r, rc = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo, lo, hi)
        assert c == rc
        assert (g is None and r is None) or (np.array_equal(g[0], r[0]) and
                                             np.array_equal(g[1], r[1]))


def test_invisible_and_massive_bodies_excluded(dev):
    _, x, y, vx, vy, rr = S.stripe_scenes()[1]
    n = len(x)
    rng = np.random.default_rng(1)
    is_ast = rng.uniform(size=n) > 0.1
    vis = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    h, halo = S.cell_size(vx, vy, rr)
    karte = Erkennungskarte(0, is_ast, -np.inf, np.inf)
    karte.stammdaten(vis, rr.astype(np.float32))
    got, cand = karte.bounce_hits(*fields(dev, x, y, vx, vy), S.DT_Y, h, halo)
    ref, rcand = C.hits(x, y, vx, vy, rr, is_ast, vis, S.DT_Y, h, halo)
    assert cand == rcand
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_bounce_suche_single_stripe(dev):
    _, x, y, vx, vy, rr = S.stripe_scenes()[0]
    n = len(x)
    ones = np.ones(n, np.uint8)
    karte = Erkennungskarte(0, ones != 0, -np.inf, np.inf)
    karte.stammdaten(ones, rr.astype(np.float32))
    res, cand, h, shares = bounce_suche([karte], None,
                                        [fields(dev, x, y, vx, vy)], S.DT_Y,
                                        float(rr.max()))
    h_ref, halo = S.cell_size(vx, vy, rr)
    assert h == pytest.approx(h_ref, rel=1e-6)
    ref, rcand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    assert cand == rcand and shares == [1.0]
    assert np.array_equal(res[0], ref[0]) and np.array_equal(res[1], ref[1])


def test_pair_list_bit_exact_in_small_batches(dev):
    """The flattened candidate list cut into many batches (ssim_collide_batch,
    here 4096 candidates: ~320 batches for the cloud, hits of a run split
    across batch ends) gives the same ordered list as the oracle."""
    from solarsystemsimulator_amd import _lib
    x, y, vx, vy, rr = S.asteroid_cloud()
    n = len(x)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    karte = Erkennungskarte(0, ones != 0, -np.inf, np.inf)
    karte.stammdaten(ones, rr.astype(np.float32))
    lib = _lib.load()
    try:
        _lib.check(lib.ssim_collide_batch(4096))
        got, cand = karte.bounce_hits(*fields(dev, x, y, vx, vy), S.DT_Y, h, halo)
    finally:
        _lib.check(lib.ssim_collide_batch(1 << 25))
    ref, rcand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    assert cand == rcand
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    with pytest.raises(ValueError):
        _lib.check(lib.ssim_collide_batch(1000))


def test_dense_cells_hit_capacity_retry(dev):
    """A dense clump with more FP32 hits than the first workspace holds:
    the search reports the count, the drop-in retries with a larger one and
    the list equals the oracle's."""
    rng = np.random.default_rng(99)
    n = 6000
    x = 1.0 + rng.normal(0.0, 2e-3, n)
    y = rng.normal(0.0, 2e-3, n)
    vx, vy = rng.normal(0, 0.5, n), rng.normal(0, 0.5, n)
    rr = np.full(n, 2e-4)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    karte = Erkennungskarte(0, ones != 0, -np.inf, np.inf)
    karte.stammdaten(ones, rr.astype(np.float32))
    karte._cap = 64
    got, cand = karte.bounce_hits(*fields(dev, x, y, vx, vy), S.DT_Y, h, halo)
    ref, rcand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    assert cand == rcand and karte._cap > 64
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("n,frac_alive", [(50_000, 1.0), (20_000, 0.7), (2, 1.0), (9, 0.5)])
def test_speed_percentile_on_device(dev, n, frac_alive):
    """ssim_speed_percentile against NumPy's percentile of the FP32 speeds
    of the alive bodies (the cell size's v95, film_producer.py:357)."""
    from solarsystemsimulator_amd.detection import speed_percentile
    rng = np.random.default_rng(n)
    vx, vy = rng.normal(0, 3, n).astype(np.float32), rng.normal(0, 3, n).astype(np.float32)
    is_ast = np.ones(n, bool)
    vis = (rng.uniform(size=n) < frac_alive).astype(np.uint8)
    vis[:2] = 1
    karte = Erkennungskarte(0, is_ast, -np.inf, np.inf)
    karte.stammdaten(vis, np.full(n, 1e-5, np.float32))
    alive = np.flatnonzero(vis != 0)
    for q in (95.0, 50.0):
        got = speed_percentile(karte, *fields(dev, vx, vy), q)
        ref = np.percentile(np.hypot(vx[alive], vy[alive]), q)
        assert abs(got - float(ref)) <= 1.2e-7 * float(ref), (q, got, ref)
