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
