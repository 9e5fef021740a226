"""What the trained network gains over the split force without it, on a
held-out 250k-mass + 500k-tracer cube (seeds 6 / 66; the committed weights
were trained on seeds 5 / 55), judged by the FP64 direct sum."""
import numpy as np
import pytest

from solarsystemsimulator_amd import scenes
from solarsystemsimulator_amd.learned_pm import (PMLearnedSystem, direct_system,
                                                 exact_accelerations, force_errors,
                                                 load_weights)

pytestmark = pytest.mark.gpu


def test_network_reduces_median_force_error(dev):
    print("learned gain on the held-out 250k cube", flush=True)
    pos, m, tpos = scenes.learned_scene(seed=6, tracer_seed=66)
    n, t = len(m), len(tpos[0])
    zn, zt = [np.zeros(n)] * 3, [np.zeros(t)] * 3
    sysm = PMLearnedSystem(dev, n, t, 1e-6, 1.0, net=load_weights(dev))
    sysm.set_bodies(pos, zn, m, tpos, zt)
    ref = [c.cpu().numpy() for c in
           exact_accelerations(direct_system(dev, n, t, 1e-6), sysm.x, n, sysm.m)]
    plain = force_errors(sysm.force(learned=False), ref)
    learned = force_errors(sysm.force(learned=True), ref)
    print(f"median plain {np.median(plain):.3e} learned {np.median(learned):.3e}")
    assert np.median(plain) < 1e-3 and np.median(learned) < 1e-3
    # at rest (v = 0) the masses get the whole correction too
    assert np.median(learned) < 0.97 * np.median(plain)
    assert np.percentile(learned, 90) < np.percentile(plain, 90)
