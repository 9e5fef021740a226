"""Throughput of the other benchmark configurations on one MI355X.

    python -m solarsystemsimulator_amd.bench_configs [--quick]

Prints one JSON object with a record per configuration (the headline
250k-mass comparison is bench.py's):

* solar_system — star + 8 planets + 10,000 test particles (2-D), raster
  integrator (frame kernel, Yoshida-4 sub-steps of <= 0.5 day), dt = 1 day,
  1000 frames: frames/s and body-steps/s;
* solar_system_3d — the same scene in space with inclinations, G = 1
  units, Plummer eps = 1e-6 AU (the 3-D frame kernel);
* film_belt_200k / film_clumps_200k — the film producer's batch loop on
  the film benchmark's belt and interpenetrating-clumps scenes with
  200,000 asteroids (200,001 bodies), 0.5-day raster: simulated days per
  second with ring slots, collision search, merges and bounces;
* allpairs_65k — 65,536 equal masses in the perturbed unit cube, Yoshida-4,
  dt = 1e-3, eps = 1e-3: ms/step in FP64 and in FP32 with compensated
  sums, and the FP32 path's force error against FP64 (median / p99);
* periodic_pm_262k — 262,144 bodies in the periodic cube, 128^3 mesh, CIC
  deposit, R2C Poisson, central-difference gradient, CIC gather: ms per
  force evaluation and per leapfrog step, FP32 and FP64 mesh, and µs per
  CIC deposit alone, tiled (reproducible) and with FP atomics;
* near_field_250k — the near field of the bench.py workload (250k masses +
  500k tracers) after one binning: µs per call of the filtered sum (FP32 /
  FP64 terms) and of the cell walk (FP32 / hi-lo), and their per-body
  differences from the FP64 filtered sum;
* collisions_50k — star + 4 planets + 50,000 asteroids in a cloud at
  2.5 AU: raster frames/s of the frame kernel, the uniform-grid collision
  search of one sample (candidates, contact pairs, ms) and the merge
  response of 40 contacts + 8 runaway bodies on the device (one call).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib, scenes
from . import device as D


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed(fn, reps=1):
    import gc
    import torch
    torch.cuda.synchronize()
    # the set-up's garbage collected and frozen first: a full collection
    # (~50 ms with torch loaded) must not land inside the timed loop
    gc.collect()
    gc.freeze()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def solar_system(frames):
    from .nbody import NBodyCuda
    from .constants import K_MAX
    x, y, vx, vy, m, vis, ast = scenes.solar_system(10_000, seed=1)
    sim = NBodyCuda(0)
    st = sim.load_state(x, y, vx, vy, m, vis, ast)
    day = 1.0 / 365.25
    sim.step_batch(st, day, K_MAX)                        # warm-up
    batches = max(1, frames // K_MAX)
    t, _ = timed(lambda: [sim.step_batch(st, day, K_MAX) for _ in range(batches)])
    nf = batches * K_MAX
    return {"bodies": len(x), "frames": nf, "dt": "1 day", "s": t,
            "frames_per_s": nf / t, "body_frames_per_s": nf * len(x) / t}


def solar_system_3d(frames):
    """The solar system in space in its own units (G = 1, AU, solar masses,
    one day = 2 pi / 365.25), inclinations, Plummer eps = 1e-6 AU."""
    from .nbody import NBodyCuda3D
    from .constants import K_MAX
    sc = scenes.solar_system_3d(10_000, seed=1)
    day = 2 * np.pi / 365.25
    sim = NBodyCuda3D(0, G=1.0, softening=1e-12, max_sub_dt=0.5 * day)
    st = sim.load_state(*sc)
    sim.step_batch(st, day, K_MAX)                        # warm-up
    batches = max(1, frames // K_MAX)
    t, _ = timed(lambda: [sim.step_batch(st, day, K_MAX) for _ in range(batches)])
    nf = batches * K_MAX
    return {"bodies": len(sc[0]), "frames": nf, "dt": "1 day", "s": t,
            "frames_per_s": nf / t, "body_frames_per_s": nf * len(sc[0]) / t}


def film(scene, n_ast, batches):
    """The film producer's loop (FilmProducer) on the film benchmark's seeded
    scene: 0.5-day raster, K = 8 samples per batch, 8 sub-samples, the
    collision search and bounces on, a ring of 1024 slots."""
    from .nbody import NBodyCuda
    from .producer import FilmProducer, sub_max_fuer
    s = scene(n_ast)
    n = len(s["x"])
    sim = NBodyCuda(0, m_sub=8)
    st = sim.load_state(s["x"], s["y"], s["vx"], s["vy"], s["mass"], s["visible"],
                        s["is_ast"], s["real_r"])
    sub_max = sub_max_fuer(n)
    sb = int(_lib.load().ssim_ring_slot_bytes(n, 8, sub_max))
    ring = np.zeros((1024, sb), np.uint8)
    prod = FilmProducer(sim, st, ring, 0.5, mass=s["mass"], real_r=s["real_r"],
                        visible=s["visible"], is_ast=s["is_ast"],
                        is_star_bh=s["is_star_bh"], sub_max=sub_max, K=8)
    prod.run(2)                                            # warm-up
    k0, c0 = prod.k, prod.candidates
    times0 = dict(prod.times)
    import gc
    gc.collect()
    gc.freeze()
    t0 = time.perf_counter()
    prod.run(batches)
    t = time.perf_counter() - t0
    samples = prod.k - k0
    share = {key: (prod.times[key] - times0[key]) / t for key in prod.times}
    out = {"bodies": n, "raster_days": 0.5, "samples": samples, "s": t,
           "sim_days_per_s": samples * 0.5 / t,
           "candidates_per_sample": (prod.candidates - c0) / samples,
           "events": prod.n_events, "collisions": prod.collisions,
           "time_share": {k: round(v, 3) for k, v in share.items()}}
    prod.close()
    return out


def ring_slots_50k(batches):
    """Raster batches of the 50,005-body asteroid cloud (m_sub 8, 16 samples
    of 0.5 day per batch) delivered as film ring slots: step_batch (frame,
    D2H of the rows, host scatter to original order and track collection;
    the per-sample slot writing of the producer comes on top and is not
    timed) against step_batch_slots (slots written on the device, one D2H)."""
    from .constants import K_MAX
    from .nbody import NBodyCuda
    x, y, vx, vy, m, rr, ast = scenes.asteroid_cloud(50_000, seed=4)
    vis = np.ones(len(x), np.uint8)
    raster = 0.5 / 365.25
    sub_max = min(len(x), max(2048, len(x) // 32))       # film_producer.py:75-76
    res = {"bodies": len(x), "samples_per_batch": K_MAX, "sub_max": sub_max}
    for name in ("step_batch", "step_batch_slots"):
        sim = NBodyCuda(0, m_sub=8)
        st = sim.load_state(x, y, vx, vy, m, vis, ast, real_r=rr)
        if name == "step_batch":
            run = lambda: sim.step_batch(st, raster, K_MAX)
        else:
            run = lambda: sim.step_batch_slots(st, raster, K_MAX, sub_max)
        run()                                                # warm-up
        t, _ = timed(lambda: [run() for _ in range(batches)])
        res[f"{name}_ms_per_batch"] = 1e3 * t / batches
    res["slot_bytes"] = 8 * len(x) + 4 + sub_max * (4 + 64)
    return res


def allpairs_65k(steps):
    from .selfgrav import SelfGravSystem
    pos, m = scenes.selfgrav_cube()
    n = len(m)
    zeros = [np.zeros(n)] * 3
    dev = D.require_gpu(0)
    out = {"bodies": n, "steps": steps, "scheme": "yoshida4", "dt": 1e-3}
    acc = {}
    for name, prec in (("f64", _lib.PREC_F64), ("f32c", _lib.PREC_F32C)):
        s = SelfGravSystem(dev, 3, n, 0, prec, _lib.SCHEME_YOSHIDA4, 1.0, 1e-6)
        s.set_bodies(pos, zeros, m)
        acc[name] = np.array([c.cpu().numpy() for c in s.accelerations()])
        s.run(1e-3, 1)
        t, _ = timed(lambda: s.run(1e-3, steps))
        out[f"ms_per_step_{name}"] = t / steps * 1e3
    err = (np.linalg.norm(acc["f32c"] - acc["f64"], axis=0) /
           np.linalg.norm(acc["f64"], axis=0))
    out["f32c_force_err_median"] = float(np.median(err))
    out["f32c_force_err_p99"] = float(np.percentile(err, 99))
    return out


def periodic_pm(steps):
    from . import pm as P
    dev = D.require_gpu(0)
    pos = scenes.perturbed_cube(64, 64, 64, seed=3, periodic=True)
    n = pos[0].size
    m = np.full(n, 1.0 / n)
    zeros = [np.zeros(n)] * 3
    out = {"bodies": n, "mesh": 128, "steps": steps}
    for name, prec in (("f64", P.F64), ("f32", P.F32)):
        for sname, scheme in (("kdk", _lib.SCHEME_KDK), ("yoshida4", _lib.SCHEME_YOSHIDA4)):
            s = P.PeriodicPMSystem(dev, 128, n, prec=prec, scheme=scheme)
            s.set_bodies(pos, zeros, m)
            s.accelerations()
            if sname == "kdk":
                t_eval, _ = timed(s.accelerations, 5)
                out[f"ms_per_eval_{name}"] = t_eval * 1e3
            s.run(1e-3, 1)                                  # warm-up
            t_step, _ = timed(lambda: s.run(1e-3, steps))
            out[f"ms_per_step_{sname}_{name}"] = t_step / steps * 1e3
        # the CIC deposit alone on the same bodies: the reproducible tiled
        # one (each node written once) and the FP atomic one (+ its fill)
        x = [D.to_dev(c, dev) for c in pos]
        md = D.to_dev(m, dev)
        for kind, tiled in (("tiled", True), ("atomic", False)):
            dep = lambda: P.deposit(x, md, [128] * 3, [0.0] * 3, 1.0 / 128, True,  # noqa: E731
                                    prec, tiled=tiled)
            dep()
            t_dep, _ = timed(dep, 20)
            out[f"us_per_deposit_{kind}_{name}"] = t_dep * 1e6
    return out


def near_field_250k(reps):
    """The near field of the 250k-mass + 500k-tracer cube (the bench.py
    workload) after one binning of all bodies: µs per call of the filtered
    sum (ssim_pp_near_bits, FP32 and FP64 terms) and of the cell walk
    (ssim_pp_near and its hi/lo form) on the same lists, and each one's
    per-body difference from the FP64 filtered sum (median / p99)."""
    import torch
    from . import pm as P
    from .p3m import SplitForce
    dev = D.require_gpu(0)
    pos, m, tpos = scenes.learned_scene()
    n, t = len(m), len(tpos[0])
    sf = SplitForce(dev, mesh_n=128, eps2=1e-6, rs_cells=1.0, rcut_rs=5.0, prec=P.F32,
                    assign="tsc", margin=4.0, freeze=True)
    x = [D.to_dev(np.concatenate([pos[c], tpos[c]]), dev) for c in range(3)]
    md = D.to_dev(m, dev)
    sf.set_mass_bound(md)
    sf.place(x)
    sf.bin_bodies(x, n, md)
    a = [torch.zeros(n + t, dtype=torch.float64, device=dev) for _ in range(3)]
    res, acc = {"bodies": n + t, "masses": n}, {}
    for name, bits, hilo in (("bits_f32", True, False), ("bits_f64", True, True),
                             ("cells_f32", False, False), ("cells_hilo", False, True)):
        sf.near_bits, sf.near_hilo = bits, hilo
        sf.near_sum(x, a, accumulate=False)
        acc[name] = np.array([c.cpu().numpy() for c in a])
        t_call, _ = timed(lambda: sf.near_sum(x, a, accumulate=False), reps)
        res[f"us_{name}"] = t_call * 1e6
    ref = acc["bits_f64"]
    nrm = np.linalg.norm(ref, axis=0)
    for name, v in acc.items():
        if name != "bits_f64":
            e = np.linalg.norm(v - ref, axis=0) / nrm
            res[f"diff_{name}"] = [float(np.median(e)), float(np.percentile(e, 99))]
    return res


def collisions_50k(frames):
    import torch
    from .constants import K_MAX
    from .detection import Erkennungskarte, bounce_suche
    from .nbody import NBodyCuda
    x, y, vx, vy, m, rr, ast = scenes.asteroid_cloud(50_000, seed=4)
    n = len(x)
    vis = np.ones(n, np.uint8)
    sim = NBodyCuda(0)
    st = sim.load_state(x, y, vx, vy, m, vis, ast, real_r=rr)
    day = 1.0 / 365.25
    sim.step_batch(st, day, K_MAX)
    batches = max(1, frames // K_MAX)
    t, rows = timed(lambda: [sim.step_batch(st, day, K_MAX) for _ in range(batches)])
    out = {"bodies": n, "frames": batches * K_MAX, "frames_per_s": batches * K_MAX / t}
    dev = D.require_gpu(0)
    det = Erkennungskarte(0, ast, -np.inf, np.inf)
    det.stammdaten(vis, rr)
    sample = rows[-1][-1]
    g = det.hochladen(sample)
    row = (g[0:n], g[n:2 * n], g[2 * n:3 * n], g[3 * n:4 * n])
    bounce_suche([det], None, [row], day, float(rr.max()))
    t, res = timed(lambda: bounce_suche([det], None, [row], day, float(rr.max())), 5)
    hits, cand, h, _ = res
    out.update({"search_ms": t * 1e3, "candidates": int(cand),
                "contact_pairs": 0 if hits is None else int(len(hits[0])),
                "cell_h_au": h})
    # the merge response on the device: 40 massive-body contacts and 8
    # runaway bodies applied to the state (one call, events read back)
    from .detection import apply_merges
    rng = np.random.default_rng(5)
    massive = np.flatnonzero(np.asarray(ast) == 0)
    pairs = [(int(massive[k % len(massive)]), int(j), 0.0)
             for k, j in enumerate(rng.choice(np.flatnonzero(ast), 40, replace=False))]
    runaway = [int(j) for j in rng.choice(np.flatnonzero(ast), 8, replace=False)]
    mirrors = {"mass": D.to_dev(np.asarray(m, np.float64), dev),
               "vis": D.to_dev(vis, dev, torch.uint8),
               "real_r": D.to_dev(np.asarray(rr, np.float64), dev),
               "is_ast": D.to_dev(np.asarray(ast, np.uint8), dev, torch.uint8),
               "is_star_bh": torch.zeros(n, dtype=torch.uint8, device=dev)}
    contact = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    smp = D.to_dev(np.asarray(sample, np.float32), dev, torch.float32)
    t, (ev, merges, _) = timed(lambda: apply_merges(st, smp, pairs, runaway, 0, mirrors,
                                                    contact, True, 2.959e-4, 0.0, 1.0))
    out.update({"apply_merges_ms": t * 1e3, "merge_pairs": len(pairs),
                "merges": merges, "event_records": len(ev) // 32})
    del dev
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--only", action="append", default=None,
                    help="run only this configuration (repeatable)")
    args = ap.parse_args(argv)
    q = args.quick
    res = {}
    for name, fn, arg in (("solar_system", solar_system, 160 if q else 1000),
                          ("solar_system_3d", solar_system_3d, 160 if q else 1000),
                          ("film_belt_200k", lambda b: film(scenes.film_belt, 200_000, b),
                           4 if q else 20),
                          ("film_clumps_200k",
                           lambda b: film(scenes.film_clumps, 200_000, b), 4 if q else 20),
                          ("allpairs_65k", allpairs_65k, 5 if q else 20),
                          ("periodic_pm_262k", periodic_pm, 5 if q else 50),
                          ("near_field_250k", near_field_250k, 5 if q else 20),
                          ("collisions_50k", collisions_50k, 32 if q else 160),
                          ("ring_slots_50k", ring_slots_50k, 3 if q else 10)):
        if args.only and name not in args.only:
            continue
        log(f"== {name}")
        res[name] = fn(arg)
        log(json.dumps(res[name]))
    print(json.dumps(res), flush=True)
    D.shutdown()


if __name__ == "__main__":
    main()
