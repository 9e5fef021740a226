"""Benchmark: body-steps/s at 250k masses + 500k tracers on one MI355X.

Workload (G = 1, unit cube, cold start): 250,000 equal masses on a perturbed
100x50x50 lattice (seed 5) and 500,000 tracers on a 100x100x50 lattice
(seed 55), Plummer softening eps = 1e-3, dt = 1e-3, leapfrog KDK in FP64
state.  A step is one KDK step of all 750,000 bodies: half-kick + drift,
one force evaluation (all-pairs: 250k x 750k = 1.875e11 pair interactions),
half-kick and the f32 snapshot of the step, all on the GPU with inputs
resident in HBM.

Prints ONE JSON line.  Options: --gpus N (only 1 is supported: this build
runs on one GPU), --steps K timed steps, --warmup W untimed steps,
--small for a quick shape check with 1/50 of the bodies.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

FP64_PEAK_TFLOPS = 78.6          # MI355X vector / matrix FP64
FLOPS_PER_PAIR = 20              # 3-D pair interaction convention


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_kernel(fn, reps, dev):
    """Average duration of fn() over reps launches, HIP events on the
    current torch stream (the stream libssim launches on)."""
    import torch
    s = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def cpu_baseline(pos, m, tpos, eps2):
    """Plain FP64 all-pairs (the oracle's port, OpenMP on all host cores)
    on a bounded sample of targets; scaled to body-steps/s of the KDK
    workload (one force evaluation per step)."""
    import oracle
    nthreads = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(nthreads))
    n, t = len(m), len(tpos[0])
    k = max(1, (n + t) // 8)
    rng = np.random.default_rng(1)
    allpos = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    idx = np.sort(rng.choice(n + t, k, replace=False))
    sample = [c[idx] for c in allpos]
    t0 = time.perf_counter()
    oracle.allpairs(pos, m, sample, 1.0, eps2, plain=True)
    dt = time.perf_counter() - t0
    per_eval = dt * (n + t) / k
    return {"value": (n + t) / per_eval, "unit": "body-steps/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", nthreads)),
            "kind": "port",
            "sample": f"{k} of {n + t} targets x {n} sources, one FP64 force "
                      f"evaluation ({dt:.1f} s), extrapolated to all targets"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; "
                          "--gpus must be 1"}))
        sys.exit(2)
    import torch
    from solarsystemsimulator_amd import _lib, scenes
    from solarsystemsimulator_amd import device as D
    from solarsystemsimulator_amd.selfgrav import SelfGravSystem

    dev = D.require_gpu(0)
    if args.small:
        pos, m, tpos = scenes.learned_scene((20, 25, 10), (20, 50, 10))
    else:
        pos, m, tpos = scenes.learned_scene()
    n, t = len(m), len(tpos[0])
    eps2, dt = 1e-6, 1e-3
    log(f"bench: {n} masses + {t} tracers, direct FP64 KDK")

    sysm = SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    sysm.set_bodies(pos, [np.zeros(n)] * 3, m, tpos, [np.zeros(t)] * 3)
    snap = torch.empty((1, sysm.snap_width()), dtype=torch.float32, device=dev)
    sysm.accelerations()

    def step():
        sysm.run(dt, 1, snap)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    body_steps = (n + t) * args.steps / wall
    log(f"direct FP64: {ms_step:.2f} ms/step, {body_steps:.4g} body-steps/s")

    # dominant kernel: the all-pairs force launch, timed in isolation
    kt = time_kernel(sysm.accelerations, 3, dev)
    pairs = n * (n + t)
    achieved = FLOPS_PER_PAIR * pairs / kt / 1e12
    roof = {"bound": "valu", "kernel": "k_allpairs_f64<3,2>",
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None, "launch_ms": kt * 1e3,
            "work": f"{FLOPS_PER_PAIR} flop/pair x {pairs:.4g} pairs"}
    log(f"all-pairs kernel {kt * 1e3:.2f} ms, {achieved:.2f} TFLOP/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, m, tpos, eps2)
        log(f"cpu baseline: {cpu['value']:.4g} body-steps/s")

    out = {
        "metric": "body-steps/s at 250k masses: direct f64 vs PM+learned at "
                  "<=1e-3 force err",
        "value": body_steps, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"direct FP64 all-pairs KDK, {n} masses + {t} "
                               "tracers, eps=1e-3, dt=1e-3, 3-D unit cube",
                   "masses": n, "tracers": t, "method": "direct_f64"},
        "direct_f64_body_steps_per_s": body_steps,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
