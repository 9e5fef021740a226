"""Train the learned correction of the 250k-mass PM method on one MI355X.

    python -m solarsystemsimulator_amd.train_learned [--small] [--out PATH]

1. the 250,000-mass / 500,000-tracer perturbed cube (seeds 5 / 55),
   eps = 1e-3, dt = 2.5e-4, is integrated with the split PM force; every
   `--every` steps the FP64 all-pairs kernel computes the exact forces of
   all bodies (64 snapshots);
2. the network is trained on the residual exact - split (units G m / h^2)
   with the project's BF16-MFMA forward/backward kernels and Adam;
3. it is evaluated on a held-out run (seeds 6 / 66): median relative force
   error against the FP64 direct sum with and without the correction;
4. the weights and the evaluation go to solarsystemsimulator_amd/data/.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import device as D
from . import scenes
from .learned import LearnedCorrection
from .learned_pm import (DEFAULT_WEIGHTS, DT, PMLearnedSystem, direct_system,
                         exact_accelerations, force_errors, make_training_set,
                         save_weights)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def evaluate(dev, net, pos, m, tpos, eps2, dt, times=(0, 250, 500, 1000)):
    """Median / p90 relative force errors of the split force with and
    without the correction at several steps of a run of the method."""
    n, t = len(m), len(tpos[0])
    sysm = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
    sysm.set_bodies(pos, [np.zeros(n)] * 3, m, tpos, [np.zeros(t)] * 3)
    direct = direct_system(dev, n, t, eps2)
    rows, done = [], 0
    for step in times:
        if step > done:
            sysm.run(dt, step - done)
            done = step
        ref = [c.cpu().numpy() for c in exact_accelerations(direct, sysm.x, n, sysm.m)]
        plain = force_errors(sysm.force(learned=False), ref)
        learned = force_errors(sysm.force(learned=True), ref)
        rows.append({"step": step,
                     "median_plain": float(np.median(plain)),
                     "median_learned": float(np.median(learned)),
                     "p90_plain": float(np.percentile(plain, 90)),
                     "p90_learned": float(np.percentile(learned, 90))})
        log(json.dumps(rows[-1]))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--small", action="store_true", help="1/50 of the bodies")
    ap.add_argument("--snapshots", type=int, default=64)
    ap.add_argument("--every", type=int, default=15)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--dt", type=float, default=DT)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    args = ap.parse_args(argv)
    dev = D.require_gpu(0)
    if args.small:
        shapes = ((20, 25, 10), (20, 50, 10))
    else:
        shapes = ((100, 50, 50), (100, 100, 50))
    pos, m, tpos = scenes.learned_scene(*shapes, seed=5, tracer_seed=55)
    eps2, dt = 1e-6, args.dt
    t0 = time.time()
    net = LearnedCorrection(dev, seed=1)
    feat, tgt = make_training_set(dev, pos, m, tpos, eps2, dt, args.snapshots,
                                  args.every, net=net, log=log)
    t1 = time.time()
    log(f"training set: {feat.shape[0]} samples in {t1 - t0:.1f} s")
    hist = net.train(feat, tgt, epochs=args.epochs, batch=args.batch,
                     lr=args.lr, lr_decay=0.6, log=log)
    t2 = time.time()
    log(f"training: {t2 - t1:.1f} s")
    del feat, tgt
    hpos, hm, htpos = scenes.learned_scene(*shapes, seed=6, tracer_seed=66)
    rows = evaluate(dev, net, hpos, hm, htpos, eps2, dt)
    save_weights(net, args.out, loss=np.array(hist),
                 eval_json=json.dumps(rows), snapshots=args.snapshots,
                 every=args.every)
    print(json.dumps({"loss": hist, "heldout": rows,
                      "t_data_s": t1 - t0, "t_train_s": t2 - t1}))


if __name__ == "__main__":
    main()
