"""Energy drift of the 250k-mass cube over 1000 steps: FP64 direct sum vs
the split PM force with and without the learned correction.

    python -m solarsystemsimulator_amd.energy_drift [--steps 1000] [--every 100]
        [--small] [--weights PATH]

Every run starts from the same cold perturbed cube (250,000 masses + 500,000
tracers, eps = 1e-3, dt = 2.5e-4, leapfrog KDK, FP64 state).  The energy of
the masses is the exact FP64 sum (ssim_energy, all pairs) at the start and
every `--every` steps, for every method, so the diagnostic itself is the
same for all three.  Prints one JSON object: per method the relative drift
|E(t) - E(0)| / |E(0)| at the end and its maximum over the samples, and the
ratio of each method's drift to the direct run's.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib, scenes
from . import device as D
from .learned_pm import DEFAULT_WEIGHTS, DT, PMLearnedSystem, load_weights
from .selfgrav import SelfGravSystem


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _energy_of(judge: SelfGravSystem, x, v, n):
    for c in range(3):
        judge.x[c][:n].copy_(x[c][:n])
        judge.v[c][:n].copy_(v[c][:n])
    return float(judge.energy()[2])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    args = ap.parse_args(argv)
    dev = D.require_gpu(0)
    if args.small:
        pos, m, tpos = scenes.learned_scene((20, 25, 10), (20, 50, 10))
    else:
        pos, m, tpos = scenes.learned_scene()
    n, t = len(m), len(tpos[0])
    eps2, dt = 1e-6, DT
    zn, zt = [np.zeros(n)] * 3, [np.zeros(t)] * 3
    judge = SelfGravSystem(dev, 3, n, 0, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    judge.set_bodies(pos, zn, m)
    checkpoints = list(range(args.every, args.steps + 1, args.every))
    if not checkpoints or checkpoints[-1] != args.steps:
        checkpoints.append(args.steps)
    result = {"masses": n, "tracers": t, "dt": dt, "steps": args.steps}

    def track(name, run, state):
        e0 = _energy_of(judge, *state(), n)
        es, done, t0 = [e0], 0, time.time()
        for cp in checkpoints:
            run(cp - done)
            done = cp
            es.append(_energy_of(judge, *state(), n))
            log(f"{name}: step {cp}, E = {es[-1]:.12e}, "
                f"drift {abs(es[-1] - e0) / abs(e0):.3e} ({time.time() - t0:.0f} s)")
        d = np.abs(np.array(es) - e0) / abs(e0)
        result[name] = {"E0": e0, "E_end": es[-1], "drift_end": float(d[-1]),
                        "drift_max": float(d.max()), "energies": es,
                        "wall_s": time.time() - t0}

    direct = SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    direct.set_bodies(pos, zn, m, tpos, zt)
    track("direct_f64", lambda k: direct.run(dt, k), lambda: (direct.x, direct.v))
    del direct

    net = load_weights(dev, args.weights)
    for name, learned in (("pm_learned", True), ("pm_plain", False)):
        fast = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
        fast.set_bodies(pos, zn, m, tpos, zt)
        track(name, lambda k, f=fast, L=learned: f.run(dt, k, learned=L),
              lambda f=fast: (f.x, f.v))
        del fast
    ref = result["direct_f64"]
    for name in ("pm_learned", "pm_plain"):
        r = result[name]
        r["drift_end_vs_direct"] = r["drift_end"] / max(ref["drift_end"], 1e-300)
        r["drift_max_vs_direct"] = r["drift_max"] / max(ref["drift_max"], 1e-300)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
