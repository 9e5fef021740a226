"""Benchmark: body-steps/s at 250k masses + 500k tracers on one MI355X,
direct FP64 all-pairs vs PM + learned correction.

Workload (G = 1, unit cube, cold start): 250,000 equal masses on a perturbed
100x50x50 lattice (seed 5) and 500,000 tracers on a 100x100x50 lattice
(seed 55), Plummer softening eps = 1e-3, dt = 2.5e-4, leapfrog KDK with FP64
state.  A step is one KDK step of all 750,000 bodies: half-kick + drift,
one force evaluation, half-kick, all on the GPU with inputs resident in HBM.

* direct FP64: all-pairs force, 250k x 750k = 1.875e11 pair interactions
  per step (+ the f32 snapshot row, as the reference's selfgrav kernel);
* PM + learned: isolated 128^3 mesh (FP32, pruned 256^3 FFT convolution)
  for the long range, near-field sum inside rcut, and the trained MLP
  correction (BF16 MFMA); its median relative force error against the FP64
  direct sum is measured on the state the timed run ends in.

`value` is the PM + learned rate (the method the metric asks for); the
direct rate, the ratio and the force errors are in the same JSON line.
`gates` (skipped with --no-gates) checks the method's accuracy targets in
the same run: 1000 PM + learned steps and 1000 direct FP64 steps of the
masses from the cold start, the exact FP64 energy at both ends of each
(`energy_drift_ratio` = PM + learned drift / direct drift), and the median
force error of PM + learned against the FP64 direct sum every 250 steps.
Per-kernel durations come from libssim's launch timing (HIP events around
each launch on its stream) over `--profile-steps` extra steps after the
timed region; `roofline` is the kernel with the largest time per step.
Options: --gpus N (1 only), --steps K (PM steps timed), --warmup W,
--direct-steps, --small (1/50 of the bodies), --no-cpu, --no-gates,
--gate-steps, --probe (a short run for hardware-counter passes: a 256 MiB
copy, then three PM steps).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

FP64_PEAK_TFLOPS = 78.6          # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3         # MI355X vector FP32
BF16_PEAK_TFLOPS = 2500.0        # MI355X dense BF16 matrix
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR = 20              # 3-D pair interaction convention
TRAFFIC_FILE = Path(__file__).resolve().parent / "profiles" / "pmc_traffic.json"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--direct-steps", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--serial", action="store_true",
                    help="one stream: no overlap of the near field with the mesh")
    ap.add_argument("--no-gates", action="store_true",
                    help="skip the 1000-step energy-drift and force-error gates")
    ap.add_argument("--gate-steps", type=int, default=1000)
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel of a step instead of replaying "
                         "the captured step")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def quiet_gc():
    """Collect the set-up's garbage and freeze what survives before a timed
    region: a full collection of the interpreter's heap (~50 ms with torch
    loaded) would otherwise land at a random step of the timed loop."""
    import gc
    gc.collect()
    gc.freeze()


def kernel_times(lib, names):
    """{name: (total ms, launches)} of the launches recorded since
    ssim_timing(1)."""
    out = {}
    for k in names:
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        rc = lib.ssim_timing_report(k.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError(lib.ssim_last_error().decode())
        out[k] = (ms.value, cnt.value)
    return out


def cpu_model():
    """The host CPU's model name (lscpu "Model name", else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(pos, m, tpos, eps2):
    """Plain FP64 all-pairs (the oracle's port of the reference's sum, OpenMP
    on all host cores) on a bounded sample of targets, scaled to
    body-steps/s of the direct KDK workload (one force evaluation/step)."""
    import oracle
    nthreads = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(nthreads))
    n, t = len(m), len(tpos[0])
    k = max(1, (n + t) // 2)          # ~15 s on 16 host cores
    rng = np.random.default_rng(1)
    allpos = [np.concatenate([pos[c], tpos[c]]) for c in range(3)]
    idx = np.sort(rng.choice(n + t, k, replace=False))
    t0 = time.perf_counter()
    oracle.allpairs(pos, m, [c[idx] for c in allpos], 1.0, eps2, plain=True)
    dt = time.perf_counter() - t0
    return {"value": (n + t) / (dt * (n + t) / k), "unit": "body-steps/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", nthreads)),
            "kind": "port", "method": "direct FP64 all-pairs",
            "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
            "sample": f"{k} of {n + t} targets x {n} sources, one FP64 force "
                      f"evaluation ({dt:.1f} s), extrapolated to all targets"}


def conv_bytes(n, mesh_bytes=4):
    """Algorithmic HBM bytes of one isolated convolution of an n^3 FP32 mesh
    (padded P = 2n, half spectra of H = n + 1, complex 2 x mesh_bytes,
    kernel spectra real mesh_bytes): the 4 pruned FFT launches (R2C rows,
    y lines, the three components' inverse y lines, the three components'
    C2R rows written as one float4 field) and the fused x-axis kernel (x < n
    half in, the kx, ky <= n corners of three kernel spectra in, three x < n
    halves out)."""
    P, H = 2 * n, n + 1
    c = 2 * mesh_bytes
    passes = [n ** 3 * mesh_bytes + n * n * H * c,          # R2C rows from rho
              n * n * H * c + n * P * H * c,                # y lines, x < n
              3 * (n * P * H * c + n * n * H * c),          # inverse y lines
              3 * n * n * H * c + n ** 3 * 4 * mesh_bytes]  # C2R rows, float4
    # kernel spectra: odd / even in each axis, so the kx, ky <= n corner of
    # each table ((n + 1)^2 H values) holds all of it
    x3 = n * P * H * c + 3 * (n + 1) * (n + 1) * H * mesh_bytes + 3 * n * P * H * c
    return passes, x3


def algorithmic_bytes(n, t, n_mesh, mesh_bytes=4):
    """Algorithmic HBM bytes per launch of the PM + learned kernels (each
    input read once, each output written once): n masses, t tracers, an
    n_mesh^3 mesh.  The near field reads every target's position and
    writes its acceleration (48 B), reads the masses' packed float4
    sources (16 B) and the target order (4 B per target).  The step's
    gather (correction folded in, adding to the near field's sum, writing
    the next box) reads per body its slot in the cell order (4 B), its
    position (24 B), the sum it adds to (24 B), the network's outputs
    (12 B) and its velocity (24 B), writes the sum (24 B), and reads the
    three fields of every node (12 of the float4's 16 B)."""
    nb = n + t
    g3 = n_mesh ** 3
    passes, x3 = conv_bytes(n_mesh, mesh_bytes)
    return {
        "k_pp_near": nb * (24 + 24 + 4) + n * 16,
        # the filtered near field reads the masses' hi and lo float4 copies
        "k_pp_bits": nb * (24 + 24 + 4) + n * 32,
        "k_fft_pass": sum(passes) / len(passes),
        "k_conv_x3": x3,
        "k_tsc_deposit": n * (3 * 8 + 8) + g3 * mesh_bytes,
        "k_tsc_gather": nb * (4 + 24 + 24 + 12 + 24 + 24) + 3 * g3 * mesh_bytes,
        "k_cic_gather": nb * 3 * 8 + 3 * g3 * mesh_bytes + nb * 3 * 8,
        "k_mlp_forward": nb * (64 + 12),
        "k_mlp_features": nb * (24 + 64) + g3 * mesh_bytes,
    }


# rocprofv3 kernel names of the bench's timed kernels (PMC passes)
PMC_KERNELS = {"k_pp_bits": "k_pp_bits<false", "k_pp_near": "k_pp_near",
               "k_fft_pass": "k_line256",
               "k_conv_x3": "k_conv_x3", "k_tsc_deposit": "k_tsc_deposit_tiles",
               "k_tsc_gather": "k_tsc_gather4<true", "k_mlp_forward": "k_mlp_infer",
               "k_mlp_features": "k_mlp_features"}


def pmc_summary(fetch_dir, write_dir, out):
    """HBM bytes per launch of the PM kernels from two counter passes over
    `bench.py --probe` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
    each scaled by what the counter reads for the probe's 256 MiB copy
    kernel (torch.mul: 256 MiB in, 256 MiB out)."""
    import csv
    rows = {k: list(csv.DictReader(open(Path(d) / "run_counter_collection.csv")))
            for k, d in (("fetch", fetch_dir), ("write", write_dir))}
    # the probe's copy is its first elementwise kernel after the fills
    copy = {k: [r for r in v if "elementwise" in r["Kernel_Name"]
                and "Fill" not in r["Kernel_Name"]][0] for k, v in rows.items()}
    copy_kib = 256 * 1024
    factor = {k: copy_kib / float(copy[k]["Counter_Value"]) for k in rows}
    alg = algorithmic_bytes(250_000, 500_000, 128)
    rec = {"workload": "250k_cube", "copy_bytes": copy_kib * 1024,
           "copy_fetch_kib": float(copy["fetch"]["Counter_Value"]),
           "copy_write_kib": float(copy["write"]["Counter_Value"]),
           "factor_fetch": factor["fetch"], "factor_write": factor["write"],
           "kernels": {},
           "commands": [
               f"rocprofv3 --pmc FETCH_SIZE -d {Path(fetch_dir).name} -o run -f csv -- "
               "python -u bench.py --probe",
               f"rocprofv3 --pmc WRITE_SIZE -d {Path(write_dir).name} -o run -f csv -- "
               "python -u bench.py --probe",
               f"python bench.py --pmc-summary {Path(fetch_dir).name} "
               f"{Path(write_dir).name} {Path(out).name}"]}
    for name, pat in PMC_KERNELS.items():
        kib = {}
        for k, v in rows.items():
            vals = [float(r["Counter_Value"]) for r in v if pat in r["Kernel_Name"]]
            if not vals:
                break
            last = vals[-(len(vals) // 4 or 1):]      # the last of the 4 steps
            kib[k] = sum(last) / len(last)
        if len(kib) < 2:
            continue
        traffic = (kib["fetch"] * factor["fetch"] + kib["write"] * factor["write"]) * 1024
        rec["kernels"][name] = {"rocprof_name": pat, "fetch_kib_per_launch": kib["fetch"],
                                "write_kib_per_launch": kib["write"],
                                "traffic_bytes_per_launch": traffic,
                                "algorithmic_bytes_per_launch": alg.get(name),
                                "traffic_over_algorithmic":
                                    traffic / alg[name] if alg.get(name) else None}
    Path(out).write_text(json.dumps(rec, indent=1))
    print(json.dumps(rec, indent=1))


def near_hilo_cost(lib, fast, dt, args, direct, nb, ms_step, kernels):
    """The same run with the near field's precise form (ssim_pp_near_bits
    with FP64 positions (hi + lo) and pair terms: p99 error of the near
    field ~2e-7 against ~2e-6):
    its step time in the timed run's launch mode, its kernel's duration
    beside the plain kernel's, and the method's force error with it."""
    from solarsystemsimulator_amd.learned_pm import force_errors
    sf = fast.split
    graph = not args.eager
    sf.near_hilo = True
    try:
        fast.run(dt, 2, graph=graph)
        import torch
        torch.cuda.synchronize(fast.dev)
        quiet_gc()
        t0 = time.perf_counter()
        fast.run(dt, args.steps, graph=graph)
        torch.cuda.synchronize(fast.dev)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        lib.ssim_timing(1)
        fast.run(dt, args.profile_steps, graph=False)
        tot, cnt = kernel_times(lib, ["k_pp_bits_precise"])["k_pp_bits_precise"]
        lib.ssim_timing(0)
        for c in range(3):
            direct.x[c][:nb].copy_(fast.x[c])
        ref = [c.cpu().numpy() for c in direct.accelerations()]
        err = force_errors(fast.force(learned=True), ref)
    finally:
        sf.near_hilo = False
    fast.force(learned=True)
    plain = kernels.get("k_pp_bits", {}).get("avg_us")
    us = tot / max(cnt, 1) * 1e3
    log(f"near field hi/lo: {ms:.3f} ms/step ({ms / ms_step:.2f}x), kernel "
        f"{us:.1f} us against {plain or 0:.1f} us, median force error {np.median(err):.3e}")
    return {"ms_per_step": ms, "step_ratio": ms / ms_step, "kernel": "k_pp_bits_precise",
            "kernel_us": us, "kernel_us_plain": plain,
            "force_err_median": float(np.median(err)),
            "near_field_p99_err": "<= 1e-6 (tests/test_p3m_gpu.py)"}


def gates(dev, pos, m, tpos, net, eps2, dt, steps, every=250):
    """The method's accuracy gates over `steps` steps from the cold start:
    the exact FP64 energy of the masses (ssim_energy, all pairs) at the
    start and the end of a PM + learned run and of a direct FP64 run, and
    the median relative force error of PM + learned (and of the split force
    without the network) against the FP64 direct sum every `every` steps.
    The direct run integrates the masses alone: the tracers are massless,
    so the masses' trajectories are the same with or without them."""
    from solarsystemsimulator_amd import _lib
    from solarsystemsimulator_amd.learned_pm import PMLearnedSystem, force_errors
    from solarsystemsimulator_amd.selfgrav import SelfGravSystem
    n, t = len(m), len(tpos[0])
    nb = n + t
    zn, zt = [np.zeros(n)] * 3, [np.zeros(t)] * 3
    judge = SelfGravSystem(dev, 3, n, 0, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    judge.set_bodies(pos, zn, m)

    def energy(x, v):
        for c in range(3):
            judge.x[c][:n].copy_(x[c][:n])
            judge.v[c][:n].copy_(v[c][:n])
        return float(judge.energy()[2])

    ref = SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    ref.set_bodies(pos, zn, m, tpos, zt)
    fast = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
    fast.set_bodies(pos, zn, m, tpos, zt)
    e0 = energy(fast.x, fast.v)
    errs, done, t0 = [], 0, time.perf_counter()
    for cp in list(range(0, steps + 1, every)) + ([steps] if steps % every else []):
        if cp > done:
            fast.run(dt, cp - done)
            done = cp
        for c in range(3):
            ref.x[c][:nb].copy_(fast.x[c])
        a_ref = [c.cpu().numpy() for c in ref.accelerations()]
        el = force_errors(fast.force(learned=True), a_ref)
        ep = force_errors(fast.force(learned=False), a_ref)
        fast.force(learned=True)
        errs.append({"step": cp, "median": float(np.median(el)),
                     "p90": float(np.percentile(el, 90)),
                     "median_without_mlp": float(np.median(ep))})
        log(f"gate step {cp}: median force error {errs[-1]['median']:.3e} "
            f"(without MLP {errs[-1]['median_without_mlp']:.3e})")
    e_pm = energy(fast.x, fast.v)
    pm_s = time.perf_counter() - t0
    fast.close()
    del fast, ref
    direct = SelfGravSystem(dev, 3, n, 0, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    direct.set_bodies(pos, zn, m)
    t0 = time.perf_counter()
    left = steps
    while left > 0:                           # chunks keep progress visible
        k = min(100, left)
        direct.run(dt, k)
        left -= k
        log(f"gate: direct FP64 masses, {steps - left}/{steps} steps "
            f"({time.perf_counter() - t0:.0f} s)")
    e_d = energy(direct.x, direct.v)
    d_pm = abs(e_pm - e0) / abs(e0)
    d_d = abs(e_d - e0) / abs(e0)
    return {"steps": steps, "dt": dt, "E0": e0,
            "energy_drift_pm_learned": d_pm, "energy_drift_direct_f64": d_d,
            "energy_drift_ratio": d_pm / max(d_d, 1e-300),
            "force_err": errs,
            "force_err_median_max": max(e["median"] for e in errs),
            "pm_run_s": pm_s, "direct_run_s": time.perf_counter() - t0,
            "energy": "exact FP64 sum over the masses (ssim_energy)"}


def main():
    if len(sys.argv) == 5 and sys.argv[1] == "--pmc-summary":
        pmc_summary(*sys.argv[2:5])
        return
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; "
                          "--gpus must be 1"}))
        sys.exit(2)
    import torch
    from solarsystemsimulator_amd import _lib, scenes
    from solarsystemsimulator_amd import device as D
    from solarsystemsimulator_amd.learned_pm import (
        DEFAULT_WEIGHTS, DT, SPLIT, PMLearnedSystem, force_errors, load_weights)
    from solarsystemsimulator_amd.selfgrav import SelfGravSystem

    dev = D.require_gpu(0)
    lib = _lib.load()
    if args.small:
        pos, m, tpos = scenes.learned_scene((20, 25, 10), (20, 50, 10))
    else:
        pos, m, tpos = scenes.learned_scene()
    n, t = len(m), len(tpos[0])
    nb = n + t
    eps2, dt = 1e-6, DT
    zeros_n, zeros_t = [np.zeros(n)] * 3, [np.zeros(t)] * 3
    net = load_weights(dev, args.weights or DEFAULT_WEIGHTS)

    if args.probe:
        # hardware-counter passes: one 256 MiB copy (the calibration kernel),
        # then three PM + learned steps
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize(dev)
        torch.mul(a, 1.0, out=b)          # reads 256 MiB, writes 256 MiB
        fast = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
        fast.set_bodies(pos, zeros_n, m, tpos, zeros_t)
        fast.split.overlap = False        # one kernel at a time on the GPU
        fast.run(dt, 3, graph=False)
        torch.cuda.synchronize(dev)
        print(json.dumps({"probe": "done", "copy_bytes": a.numel() * 4}))
        return

    # ---- direct FP64 ------------------------------------------------------
    direct = SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, 1.0, eps2)
    direct.set_bodies(pos, zeros_n, m, tpos, zeros_t)
    snap = torch.empty((1, direct.snap_width()), dtype=torch.float32, device=dev)
    direct.run(dt, 1, snap)                         # warm-up (+ initial force)
    torch.cuda.synchronize(dev)
    quiet_gc()
    t0 = time.perf_counter()
    for _ in range(args.direct_steps):
        direct.run(dt, 1, snap)
    torch.cuda.synchronize(dev)
    wall_d = time.perf_counter() - t0
    direct_rate = nb * args.direct_steps / wall_d
    log(f"direct FP64: {wall_d / args.direct_steps * 1e3:.2f} ms/step, "
        f"{direct_rate:.4g} body-steps/s")
    lib.ssim_timing(1)
    direct.accelerations()
    direct.accelerations()
    ms, cnt = kernel_times(lib, ["k_allpairs_f64"])["k_allpairs_f64"]
    lib.ssim_timing(0)
    kt = ms / cnt / 1e3
    pairs = n * nb
    ach = FLOPS_PER_PAIR * pairs / kt / 1e12
    roof_direct = {"bound": "valu", "kernel": "k_allpairs_f64<3,2>",
                   "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                   "launch_us": kt * 1e6,
                   "work": f"{FLOPS_PER_PAIR} flop/pair x {pairs:.4g} pairs"}

    # ---- PM + learned -----------------------------------------------------
    fast = PMLearnedSystem(dev, n, t, eps2, 1.0, net=net)
    fast.set_bodies(pos, zeros_n, m, tpos, zeros_t)
    fast.split.overlap = not args.serial
    graph = not args.eager
    fast.run(dt, max(args.warmup, 1), graph=graph)
    torch.cuda.synchronize(dev)
    quiet_gc()
    t0 = time.perf_counter()
    fast.run(dt, args.steps, graph=graph)
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

    # ---- per-kernel durations (after the timed region) ---------------------
    sf = fast.split
    sf.bin_bodies(fast.x, n, fast.m)
    tested, inside = sf.near_pair_counts(fast.x)
    names = ["k_pp_bits", "k_pp_near", "k_pp_cells", "k_fft_pass", "k_conv_x3", "k_tsc_deposit",
             "k_tsc_gather", "k_cic_deposit", "k_cic_gather", "k_mlp_forward",
             "k_mlp_features"]
    # HIP events per launch over eager steps in the timed run's stream
    # configuration (events cannot sit inside the replayed graph), so each
    # duration is the kernel's own under the same sharing of the GPU
    lib.ssim_timing(1)
    fast.run(dt, args.profile_steps, graph=False)
    times = kernel_times(lib, names)
    lib.ssim_timing(0)
    steps_p = args.profile_steps
    n_mesh = sf.n
    pass_bytes, x3_bytes = conv_bytes(n_mesh)
    mlp_flops = 2 * nb * (64 * 32 + 2 * 64 * 64 + 16 * 64)
    alg = algorithmic_bytes(n, t, n_mesh)
    work = {
        "k_pp_bits": ("valu", FLOPS_PER_PAIR * inside, FP32_PEAK_TFLOPS,
                      "TFLOP/s", f"{FLOPS_PER_PAIR} flop x {inside:.4g} "
                      f"pairs inside rcut ({tested:.4g} pairs through the FP16 "
                      f"MFMA filter)"),
        "k_pp_near": ("valu", FLOPS_PER_PAIR * inside, FP32_PEAK_TFLOPS,
                      "TFLOP/s", f"{FLOPS_PER_PAIR} flop x {inside:.4g} "
                      f"pairs inside rcut ({tested:.4g} pairs tested)"),
        "k_pp_cells": ("valu", FLOPS_PER_PAIR * inside, FP32_PEAK_TFLOPS,
                       "TFLOP/s", f"{FLOPS_PER_PAIR} flop x {inside:.4g} "
                       f"pairs inside rcut ({tested:.4g} pairs tested)"),
        "k_fft_pass": ("hbm", sum(pass_bytes) / len(pass_bytes), HBM_PEAK_GBS,
                       "GB/s", f"{sum(pass_bytes) / 1e6:.1f} MB over the 4 pruned "
                       f"FFT passes of one {2 * n_mesh}^3 isolated convolution"),
        "k_conv_x3": ("hbm", x3_bytes, HBM_PEAK_GBS, "GB/s",
                      f"{x3_bytes / 1e6:.1f} MB: x < {n_mesh} half in, 3 kernel "
                      f"spectrum corners in, 3 x < {n_mesh} halves out"),
        "k_tsc_deposit": ("hbm", alg["k_tsc_deposit"], HBM_PEAK_GBS, "GB/s",
                          "N (3 x 8 B position + 8 B mass) + n^3 x 4 B mesh"),
        "k_tsc_gather": ("hbm", alg["k_tsc_gather"], HBM_PEAK_GBS, "GB/s",
                         "N (4 B slot + 24 B position + 24 B sum in + 12 B network "
                         "outputs + 24 B velocity + 24 B sum out) + 3 n^3 x 4 B meshes"),
        "k_cic_deposit": ("hbm", alg["k_tsc_deposit"], HBM_PEAK_GBS, "GB/s",
                          "N (3 x 8 B position + 8 B mass) + n^3 x 4 B mesh"),
        "k_cic_gather": ("hbm", alg["k_cic_gather"], HBM_PEAK_GBS, "GB/s",
                         "N x 3 x 8 B position + 3 n^3 x 4 B meshes + N x 3 x 8 B "
                         "acceleration"),
        "k_mlp_forward": ("mfma", mlp_flops, BF16_PEAK_TFLOPS, "TFLOP/s",
                          "2 x 11,264 MAC per body, BF16"),
        "k_mlp_features": ("hbm", alg["k_mlp_features"], HBM_PEAK_GBS, "GB/s",
                           "24 B position in + 64 B feature row out per body + "
                           "n^3 x 4 B mesh"),
    }
    kernels = {}
    for k, (tot_ms, cnt) in times.items():
        if cnt == 0:
            continue
        bound, w, peak, unit, desc = work[k]
        avg_s = tot_ms / cnt / 1e3
        scale = 1e9 if unit == "GB/s" else 1e12
        ach = w / avg_s / scale
        kernels[k] = {"bound": bound, "avg_us": avg_s * 1e6,
                      "launches_per_step": cnt / steps_p,
                      "us_per_step": tot_ms / steps_p * 1e3,
                      "achieved": ach, "peak": peak, "unit": unit,
                      "frac": ach / peak, "work": desc}
        log(f"  {k}: {cnt / steps_p:.0f} x {avg_s * 1e6:.1f} us/step, "
            f"{ach:.1f} {unit} ({ach / peak:.1%} of peak)")
    dom = max(kernels, key=lambda k: kernels[k]["us_per_step"])
    d = kernels[dom]
    traffic = None
    if TRAFFIC_FILE.exists():
        rec = json.loads(TRAFFIC_FILE.read_text())
        if rec.get("workload") == "250k_cube" and not args.small:
            traffic = rec.get("kernels", {}).get(dom, {}).get("traffic_bytes_per_launch")
    roof = {"bound": d["bound"], "kernel": dom, "achieved": d["achieved"],
            "peak": d["peak"], "unit": d["unit"], "frac": d["frac"],
            "traffic": traffic, "launch_us": d["avg_us"], "work": d["work"]}

    # ---- the hi/lo near field (FP64 pair arithmetic): its cost -------------
    hilo = near_hilo_cost(lib, fast, dt, args, direct, nb, ms_step, kernels)

    gate = None
    if not args.no_gates and not args.small:
        fast.close()
        gate = gates(dev, pos, m, tpos, net, eps2, dt, args.gate_steps)
        log(f"gates: energy drift PM + learned {gate['energy_drift_pm_learned']:.3e}, "
            f"direct {gate['energy_drift_direct_f64']:.3e} "
            f"(ratio {gate['energy_drift_ratio']:.2f}), worst median force error "
            f"{gate['force_err_median_max']:.3e}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, m, tpos, eps2)
        log(f"cpu baseline: {cpu['value']:.4g} body-steps/s")

    split = {k: v for k, v in SPLIT.items() if k != "prec"}
    out = {
        "metric": "body-steps/s at 250k masses: direct f64 vs PM+learned at "
                  "<=1e-3 force err",
        "value": value, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "precision": "force arithmetic f32 (mesh, FFT, near-field pair terms); "
                     "f64 state, kicks, drifts and energy; bf16 MLP with f32 "
                     "accumulation; the direct run is f64 throughout",
        "config": {"workload": f"{n} masses + {t} tracers, perturbed unit cube, "
                               f"eps=1e-3, dt={dt:g}, leapfrog KDK, PM "
                               f"{n_mesh}^3 isolated + near field + learned "
                               "correction",
                   "masses": n, "tracers": t, "method": "pm_learned",
                   "split": split,
                   "launch": "eager" if args.eager else "hip_graph",
                   "streams": 1 if args.serial else 2},
        "direct_f64_body_steps_per_s": direct_rate,
        "speedup_vs_direct_f64": value / direct_rate,
        "force_err_median": float(np.median(err_l)),
        "force_err_p90": float(np.percentile(err_l, 90)),
        "force_err_median_without_mlp": float(np.median(err_p)),
        "gates": gate,
        "energy_drift_ratio": None if gate is None else gate["energy_drift_ratio"],
        "kernels": kernels,
        "near_hilo": hilo,
        "roofline": roof,
        "roofline_direct": roof_direct,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    D.shutdown()


if __name__ == "__main__":
    main()
