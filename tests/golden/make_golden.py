"""Writes the golden fixtures of this directory: small seeded inputs and
their expected outputs, so that a change of NumPy's RNG streams or of the
oracle cannot move the parity targets unnoticed.

The reference ships no golden vectors of its own, and its modules cannot be
imported here (they need CuPy and an NVIDIA device), so every expectation
is computed by an independent route:

* ``allpairs_wolke.npz`` -- the 2-D cloud of test_selfgrav.py:43-51 (seed
  11, 256 bodies); accelerations of selfgrav_kernel.py:222-260 as the
  exactly rounded sum (``math.fsum``) of the FP64 pair terms;
* ``nbody_seed7.npz`` -- the system of test_kernel.py:84-108 (seed 7, 200
  tracers), two 5-day frames of the serial C restatement of
  nbody_kernel.py:97-560 (outputs, FP64 state, sub-samples, contacts);
* ``cic_seed42.npz`` -- CIC cells (the floor / clip rule of
  pm_kernel.py:63-90, restated here as a closed form) and the deposited
  mesh, 2-D isolated and 3-D periodic;
* ``collide_4711.npz`` -- the clump and on-the-border scenes of
  test_erkennung_streifen.py:76-111 (seed 4711), the ordered pair list and
  candidate count of film_producer.py:223-317 (the oracle's order; its pair
  set is asserted equal to the O(n^2) brute force over the same reach), and
  the bounce response of film_producer.py:916-1019;
* ``mlp_forward.npz`` -- the learned correction's network on 64 BF16
  feature rows, from the NumPy forward pass in float64.

Run ``python -m tests.golden.make_golden`` from the repository root.
"""
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def allpairs_wolke():
    from oracle import selfgrav_np as S
    x, y, vx, vy, m = S.wolke(256, 0.3, seed=11)
    G, eps2 = S.G_AU, 1e-4 ** 2
    ax, ay = np.empty(len(x)), np.empty(len(x))
    for i in range(len(x)):
        dx, dy = x - x[i], y - y[i]
        r2 = dx * dx + dy * dy + eps2
        w = G * m / (r2 * np.sqrt(r2))
        ax[i] = math.fsum(w * dx)
        ay[i] = math.fsum(w * dy)
    return dict(x=x, y=y, m=m, G=G, eps2=eps2, ax=ax, ay=ay)


def nbody_seed7():
    import oracle
    from oracle import nbody_np as nb
    st0 = nb.make_system(200, seed=7)
    rr = np.full(len(st0[0]), 1e-5)
    rr[0] = 0.00465
    raster, k = 5 / 365.25, 2
    out, state, sub, hits = oracle.nbody_frames(*st0, rr, raster, k)
    names = ("x", "y", "vx", "vy", "mass", "visible", "is_ast")
    d = {n: np.asarray(v) for n, v in zip(names, st0)}
    d.update(real_r=rr, raster=raster, k=k, out=out, state=state,
             sub_counts=np.array([len(i) for i, _ in sub]),
             sub_idx=np.concatenate([i for i, _ in sub]),
             sub_pos=np.concatenate([p for _, p in sub], axis=-1),
             hit_idx=hits[0], hit_t=hits[1], hit_partner=hits[2])
    return d


def cic_closed_form(pos, origin, h, n, periodic):
    """Cell index per axis: floor((x - x0) / h), clipped to [0, n - 2] on an
    isolated mesh, wrapped modulo n on a periodic one."""
    cells = []
    for x, o, k in zip(pos, origin, n):
        c = np.floor((np.asarray(x) - o) / h).astype(np.int64)
        cells.append(np.mod(c, k) if periodic else np.clip(c, 0, k - 2))
    return np.array(cells, np.int32)


def cic_seed42():
    from oracle import pm_np
    rng = np.random.default_rng(42)
    d = {}
    for dim, periodic in ((2, False), (3, True)):
        n = [16, 8, 32][:dim]
        origin = [-0.3, 0.1, 0.7][:dim]
        h = 1.0 / 37.0
        pos = [rng.uniform(o - 2 * h, o + (k + 1) * h, 2000) for o, k in zip(origin, n)]
        m = rng.uniform(0.5, 2.0, 2000)
        tag = f"{dim}d"
        d.update({f"{tag}_pos": np.array(pos), f"{tag}_m": m,
                  f"{tag}_n": np.array(n), f"{tag}_origin": np.array(origin),
                  f"{tag}_h": h, f"{tag}_periodic": periodic,
                  f"{tag}_cells": cic_closed_form(pos, origin, h, n, periodic),
                  f"{tag}_mesh": pm_np.cic_deposit(pos, m, n, origin, h, periodic)})
    return d


def collide_4711():
    from oracle import collide_np as C
    from oracle import merge_np as M
    from tests import collide_cases as S
    scenes = S.stripe_scenes()
    d = {}
    for name, x, y, vx, vy, rr in (scenes[1], scenes[3]):
        x, y, vx, vy, rr = (np.asarray(a, np.float32) for a in (x, y, vx, vy, rr))
        h, halo = S.cell_size(vx, vy, rr)
        ones = np.ones(len(x), np.uint8)
        pairs, cand = C.hits(x, y, vx, vy, rr, ones, ones, S.DT_Y, h, halo)
        brute = C.brute_force(x, y, vx, vy, rr, ones, ones, S.DT_Y, h)
        assert set(zip(pairs[0].tolist(), pairs[1].tolist())) == brute
        d.update({f"{name}_in": np.array([x, y, vx, vy, rr]), f"{name}_h": h,
                  f"{name}_halo": halo, f"{name}_cand": cand,
                  f"{name}_i": pairs[0].astype(np.int32),
                  f"{name}_j": pairs[1].astype(np.int32)})
        if name == "klumpen":
            sample = np.concatenate([x, y, vx, vy]).astype(np.float32)
            mass = np.random.default_rng(2).uniform(1e-12, 3e-12, len(x))
            # the producer's radii are FP64 host arrays (film_producer.py:916-1019)
            hi, hj, hit, dl = M.bounce_deltas(sample, pairs[0], pairs[1],
                                              rr.astype(np.float64), mass, S.DT_Y)
            d.update(bounce_mass=mass, bounce_hi=hi.astype(np.int32),
                     bounce_hj=hj.astype(np.int32),
                     bounce_hit=hit.astype(np.int32), bounce_deltas=dl)
    return d


def mlp_forward():
    from oracle import mlp_np as M
    rng = np.random.default_rng(3)
    p = M.init_params(3)
    dd = M.unpack(p.copy())
    for b in ("b1", "b2", "b3"):
        dd[b] = rng.normal(0, 0.1, dd[b].shape).astype(np.float32)
    p = M.pack(dd)
    x = np.zeros((64, 32), np.float32)
    x[:, :30] = rng.normal(0, 1, (64, 30))
    x = M.bf16(x)
    exact, _ = M.forward(M.bf16(p), x, rounding=False)
    return dict(params=p, x=x, out=exact)


FIXTURES = {"allpairs_wolke": allpairs_wolke, "nbody_seed7": nbody_seed7,
            "cic_seed42": cic_seed42, "collide_4711": collide_4711,
            "mlp_forward": mlp_forward}


def load(name):
    with np.load(HERE / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        np.savez_compressed(HERE / f"{name}.npz", **fn())
        print(name, (HERE / f"{name}.npz").stat().st_size, "bytes")
