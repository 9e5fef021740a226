"""Pin the collision-search oracle (CPU): its ordered list against an
independent O(n^2) brute force over the same reach rule and predicates, and
the stripe decomposition of test_erkennung_streifen.py (union of 2/3/5
stripes == one stripe, no pair twice)."""
import numpy as np
import pytest

from oracle import collide_np as C
from tests import collide_cases as S


@pytest.mark.parametrize("scene", S.stripe_scenes()[1:], ids=lambda s: s[0])
def test_hits_equal_brute_force(scene):
    _, x, y, vx, vy, rr = scene
    n = len(x)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    res, cand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    got = set() if res is None else set(zip(res[0].tolist(), res[1].tolist()))
    assert got == C.brute_force(x, y, vx, vy, rr, ones, ones, S.DT_Y, h)
    assert cand >= len(got)


@pytest.mark.parametrize("scene", S.stripe_scenes(), ids=lambda s: s[0])
def test_stripes_union_equals_single(scene):
    _, x, y, vx, vy, rr = scene
    n = len(x)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    res, _ = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    full = set() if res is None else set(zip(res[0].tolist(), res[1].tolist()))
    for k in (2, 3, 5):
        cuts = np.percentile(np.float32(x), 100.0 * np.cumsum([1.0 / k] * (k - 1)))
        edges = [-np.inf, *cuts, np.inf]
        union, twice = set(), []
        for lo, hi in zip(edges[:-1], edges[1:]):
            r, _ = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo, lo, hi)
            part = set() if r is None else set(zip(r[0].tolist(), r[1].tolist()))
            twice += list(union & part)
            union |= part
        assert not twice and union == full


def test_order_is_offset_then_row():
    """Hits come grouped by neighbour offset, rows ascending in each."""
    _, x, y, vx, vy, rr = S.stripe_scenes()[1]
    n = len(x)
    h, halo = S.cell_size(vx, vy, rr)
    ones = np.ones(n, np.uint8)
    (ci, cj), _ = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
    assert len(ci) > 10
    # within the (0,0) block every pair has i < j; the row index never
    # decreases inside one offset block, so it can only drop 4 times
    drops = int(np.sum(np.diff(ci) < 0))
    assert drops <= 4


def percentile_f32(values, q):
    """NumPy's 'linear' percentile of an FP32 array restated in FP32 (the
    rule ssim_speed_percentile runs on the device): virtual index
    (m - 1) * float32(q / 100), gamma its fraction, a + (b - a) gamma, or
    b - (b - a)(1 - gamma) when gamma >= 1/2."""
    v = np.sort(np.asarray(values, np.float32))
    m = len(v)
    qf = np.float32(q) / np.float32(100)
    vi = np.float32(m - 1) * qf
    lo = int(np.floor(vi))
    hi = min(lo + 1, m - 1)
    g = np.float32(vi - np.float32(lo))
    a, b = v[lo], v[hi]
    d = np.float32(b - a)
    return np.float32(b - d * np.float32(1 - g)) if g >= 0.5 else np.float32(a + d * g)


def test_percentile_restatement_matches_numpy():
    """The FP32 restatement equals np.percentile bit for bit (sizes where
    the virtual index has a fraction above, below and at 1/2, and m = 1)."""
    rng = np.random.default_rng(3)
    for m in (1, 2, 3, 7, 20, 21, 101, 1000, 50_000):
        x = rng.gamma(2.0, 1.5, m).astype(np.float32)
        for q in (95, 50, 5, 100, 0, 37.5):
            ref = np.percentile(x, q)
            assert ref.dtype == np.float32
            assert percentile_f32(x, q) == ref, (m, q)
