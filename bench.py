"""Benchmark: body-steps/s at 250k masses + 500k tracers on one MI355X,
direct FP64 all-pairs vs PM + learned correction.

Workload (G = 1, unit cube, cold start): 250,000 equal masses on a perturbed
100x50x50 lattice (seed 5) and 500,000 tracers on a 100x100x50 lattice
(seed 55), Plummer softening eps = 1e-3, dt = 2.5e-4, leapfrog KDK with FP64
state.  A step is one KDK step of all 750,000 bodies: half-kick + drift,
one force evaluation, half-kick (+ the f32 snapshot for the direct run),
all on the GPU with inputs resident in HBM.

* direct FP64: all-pairs force, 250k x 750k = 1.875e11 pair interactions
  per step;
* PM + learned: isolated 128^3 mesh (FP32, zero-padded 256^3 FFTs) for the
  long range, near-field sum inside 5 rs, and the trained MLP correction
  (BF16 MFMA); its median relative force error against the FP64 direct sum
  is measured on the state the timed run ends in.

`value` is the PM + learned rate (the method the metric asks for); the
direct rate, the ratio and the force errors are in the same JSON line.
Options: --gpus N (1 only), --steps K (PM steps timed), --warmup W,
--direct-steps, --small (1/50 of the bodies), --no-cpu.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

FP64_PEAK_TFLOPS = 78.6          # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR = 20              # 3-D pair interaction convention


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--direct-steps", type=int, default=3)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--weights", default=None)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Timer:
    """HIP events on the current torch stream (the stream libssim uses)."""

    def __init__(self, dev):
        import torch
        self.torch, self.dev = torch, dev

    def __call__(self, fn, reps):
        torch = self.torch
        s = torch.cuda.current_stream(self.dev)
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
    """Plain FP64 all-pairs (the oracle's port of the reference's sum, OpenMP
    on all host cores) on a bounded sample of targets, scaled to
    body-steps/s of the direct KDK workload (one force evaluation/step)."""
    import oracle
    nthreads = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(nthreads))
    n, t = len(m), len(tpos[0])
    k = max(1, (n + t) // 8)
    rng = np.random.default_rng(1)
    allpos = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    idx = np.sort(rng.choice(n + t, k, replace=False))
    t0 = time.perf_counter()
    oracle.allpairs(pos, m, [c[idx] for c in allpos], 1.0, eps2, plain=True)
    dt = time.perf_counter() - t0
    return {"value": (n + t) / (dt * (n + t) / k), "unit": "body-steps/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", nthreads)),
            "kind": "port", "method": "direct FP64 all-pairs",
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
    from solarsystemsimulator_amd.learned_pm import (
        DEFAULT_WEIGHTS, DT, PMLearnedSystem, force_errors, load_weights)
    from solarsystemsimulator_amd.selfgrav import SelfGravSystem

    dev = D.require_gpu(0)
    timer = Timer(dev)
    if args.small:
        pos, m, tpos = scenes.learned_scene((20, 25, 10), (20, 50, 10))
    else:
        pos, m, tpos = scenes.learned_scene()
    n, t = len(m), len(tpos[0])
    nb = n + t
    eps2, dt = 1e-6, DT
    zeros_n, zeros_t = [np.zeros(n)] * 3, [np.zeros(t)] * 3

    # ---- direct FP64 ------------------------------------------------------
    direct = SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    direct.set_bodies(pos, zeros_n, m, tpos, zeros_t)
    snap = torch.empty((1, direct.snap_width()), dtype=torch.float32, device=dev)
    direct.run(dt, 1, snap)                         # warm-up (+ initial force)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.direct_steps):
        direct.run(dt, 1, snap)
    torch.cuda.synchronize(dev)
    wall_d = time.perf_counter() - t0
    direct_rate = nb * args.direct_steps / wall_d
    log(f"direct FP64: {wall_d / args.direct_steps * 1e3:.2f} ms/step, "
        f"{direct_rate:.4g} body-steps/s")
    kt = timer(direct.accelerations, 2)
    pairs = n * nb
    ach = FLOPS_PER_PAIR * pairs / kt / 1e12
    roof_direct = {"bound": "valu", "kernel": "k_allpairs_f64<3,2>",
                   "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                   "launch_ms": kt * 1e3,
                   "work": f"{FLOPS_PER_PAIR} flop/pair x {pairs:.4g} pairs"}

    # ---- PM + learned -----------------------------------------------------
    wpath = args.weights or DEFAULT_WEIGHTS
    net = load_weights(dev, wpath)
    fast = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
    fast.set_bodies(pos, zeros_n, m, tpos, zeros_t)
    fast.run(dt, max(args.warmup, 1))
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    fast.run(dt, args.steps)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    value = nb * args.steps / wall
    ms_step = wall / args.steps * 1e3
    log(f"PM + learned: {ms_step:.3f} ms/step, {value:.4g} body-steps/s, "
        f"{value / direct_rate:.1f}x direct")

    # force error of the method at the state its timed run ended in
    for c in range(3):
        direct.x[c][:nb].copy_(fast.x[c])
    ref = [c.cpu().numpy() for c in direct.accelerations()]
    err_l = force_errors(fast.force(learned=True), ref)
    err_p = force_errors(fast.force(learned=False), ref)
    fast.force(learned=True)
    log(f"median rel force error: learned {np.median(err_l):.3e}, "
        f"without MLP {np.median(err_p):.3e}")

    # dominant kernels of one PM + learned force evaluation, timed in isolation
    parts = {
        "mesh (deposit, FFTs, Green's function, gather)":
            timer(lambda: fast.split.mesh_accel(fast.x, n, fast.m, fast.a), 5),
        "near field (cell list + k_pp_short)":
            timer(lambda: fast.split.near_accel(fast.x, n, fast.m, fast.a), 5),
        "MLP (features + k_mlp_forward)":
            timer(lambda: fast.net.predict(fast.features(), add_to=fast.a,
                                           scale=fast.a_unit()), 5),
    }
    for k_, v_ in parts.items():
        log(f"  {k_}: {v_ * 1e3:.3f} ms")
    # algorithmic HBM bytes of the mesh part per evaluation (FP32 mesh):
    # deposit (n x 32 B + mesh 4 B/node), 1 forward + 3 inverse 3-D FFTs on
    # the padded mesh (3 passes each, read+write the half spectrum), 3
    # spectral products, 3 crops, gather (nb x (24 B in + 24 B out) + 3 x
    # mesh 4 B/node)
    P_ = 2 * fast.split.n
    spec = P_ * P_ * (P_ // 2 + 1) * 8
    mesh4 = fast.split.n ** 3 * 4
    mesh_bytes = (n * 32 + mesh4 + 4 * 3 * 2 * spec + 3 * 3 * spec
                  + 3 * 2 * mesh4 + nb * 48 + 3 * mesh4)
    mt = list(parts.values())[0]
    roof = {"bound": "hbm", "kernel": "PM mesh pipeline (k_fft_pass x12, "
            "k_cic_deposit, k_cmul x3, k_pm_pad, k_cic_gather)",
            "achieved": mesh_bytes / mt / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": mesh_bytes / mt / 1e9 / HBM_PEAK_GBS,
            "traffic": None, "launch_ms": mt * 1e3,
            "work": f"{mesh_bytes / 1e6:.1f} MB algorithmic per evaluation"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, m, tpos, eps2)
        log(f"cpu baseline: {cpu['value']:.4g} body-steps/s")

    out = {
        "metric": "body-steps/s at 250k masses: direct f64 vs PM+learned at "
                  "<=1e-3 force err",
        "value": value, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "precision": "f64 state; f32 mesh and near field; bf16 MLP with f32 "
                     "accumulation; direct run all f64",
        "config": {"workload": f"{n} masses + {t} tracers, perturbed unit cube, "
                               f"eps=1e-3, dt={dt:g}, leapfrog KDK, PM 128^3 "
                               "isolated + near field + learned correction",
                   "masses": n, "tracers": t, "method": "pm_learned"},
        "direct_f64_body_steps_per_s": direct_rate,
        "speedup_vs_direct_f64": value / direct_rate,
        "force_err_median": float(np.median(err_l)),
        "force_err_p90": float(np.percentile(err_l, 90)),
        "force_err_median_without_mlp": float(np.median(err_p)),
        "ms_per_eval_parts": {k_: v_ * 1e3 for k_, v_ in parts.items()},
        "roofline": roof,
        "roofline_direct": roof_direct,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
