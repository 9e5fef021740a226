"""The film producer's batch loop (FilmProducer: ring slots written on the
device, detection from device rows on a second stream one batch behind)
against the reference's host route, restated here from film_producer.py:
1152-1295 and its closures: `step_batch` to host rows, `schreibe_slot` on
the host, the batch uploaded to the detection device, the analysis of
batch b applied before step b+1, events packed as "<dIIfIff".

Both routes call the same analysis kernels (pinned against the NumPy
restatements in test_merge_gpu.py and test_collide_gpu.py), so what is
compared here is the loop: slot bytes, event records, counters and the
final state must be identical.  Scenes: the film benchmark's belt and
clumps (bench_film.py:45-86) at reduced size; the clumps get a planet and
a few fast bodies so that merges (kind 0), bounces (kind 1) and runaway
removals (kind 2) all occur."""
import struct

import numpy as np
import pytest
import torch

from oracle import ring_np as R
from solarsystemsimulator_amd import _lib, scenes
from solarsystemsimulator_amd import device as D
from solarsystemsimulator_amd.constants import BOUNCE_E, EV_BYTES, EV_FORMAT, G_AU
from solarsystemsimulator_amd.detection import (Erkennungskarte, apply_merges,
                                                bounce_deltas, bounce_suche,
                                                merge_scan)
from solarsystemsimulator_amd.nbody import NBodyCuda
from solarsystemsimulator_amd.producer import FilmProducer, sub_max_fuer

pytestmark = pytest.mark.gpu
RASTER = 0.5
M_SUB = 8


def _scene(kind, n_ast):
    if kind == "belt":
        return scenes.film_belt(n_ast)
    s = scenes.film_clumps(n_ast)
    # a planet in the first cloud (merges) and five bodies far beyond the
    # escape speed (runaway removals)
    s["x"] = np.concatenate([s["x"][:1], [2.0], s["x"][1:]])
    s["y"] = np.concatenate([s["y"][:1], [0.0], s["y"][1:]])
    s["vx"] = np.concatenate([s["vx"][:1], [0.0], s["vx"][1:]])
    s["vy"] = np.concatenate([s["vy"][:1], [np.sqrt(G_AU / 2.0)], s["vy"][1:]])
    for key, val in (("mass", 1e-3), ("real_r", 0.01), ("visible", 1), ("is_ast", 0),
                     ("is_star_bh", 0)):
        s[key] = np.concatenate([s[key][:1], np.array([val], s[key].dtype), s[key][1:]])
    s["vx"][-5:] = 40.0
    return s


def _load(s):
    sim = NBodyCuda(0, m_sub=M_SUB)
    st = sim.load_state(s["x"], s["y"], s["vx"], s["vy"], s["mass"], s["visible"],
                        s["is_ast"], s["real_r"])
    return sim, st


def host_route(s, batches, K, capacity, sub_max):
    """The reference's producer loop on host rows (film_producer.py:
    1152-1295), without throttling and diagnostics."""
    sim, st = _load(s)
    dev = sim._dev
    n = len(s["x"])
    dt_years = RASTER / 365.25
    sb = R.slot_bytes(n, M_SUB, sub_max)
    ring = np.zeros((capacity, sb), np.uint8)
    buf = memoryview(ring.reshape(-1))           # a byte buffer, like the shm ring
    mass = np.array(s["mass"], np.float64)
    real_r = np.array(s["real_r"], np.float64)
    vis = np.array(s["visible"], np.uint8)
    is_ast = np.asarray(s["is_ast"]) != 0
    mirrors = {"mass": D.to_dev(mass, dev), "vis": D.to_dev(vis, dev, torch.uint8),
               "real_r": D.to_dev(real_r, dev),
               "is_ast": D.to_dev(is_ast.astype(np.uint8), dev, torch.uint8),
               "is_star_bh": D.to_dev(np.asarray(s["is_star_bh"], np.uint8), dev,
                                      torch.uint8)}
    contact = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    karte = Erkennungskarte(0, is_ast, -np.inf, np.inf)
    state = {"prev": None, "dirty": False, "anteile": [1.0], "events": bytearray(),
             "coll": 0, "over": 0}

    def stammdaten():
        karte.stammdaten(vis, real_r.astype(np.float32))
        return float(real_r[is_ast].max())
    rr_max = [stammdaten()]

    def analyze_sample(out_np, k_ev, g):
        gx, gy, gvx, gvy = g[0:n], g[n:2 * n], g[2 * n:3 * n], g[3 * n:4 * n]
        px, py = (gx, gy) if state["prev"] is None else state["prev"]
        m_all = st["m_idx_h"]
        m_alive = m_all[(vis[m_all] != 0) & (mass[m_all] > 0)]
        pairs, runaway = [], None
        if len(m_alive):
            mw = mass[m_alive]
            msum = float(mw.sum())
            bary = None
            if msum > 0:
                bary = (float((out_np[0:n][m_alive] * mw).sum() / msum),
                        float((out_np[n:2 * n][m_alive] * mw).sum() / msum))
            hit, runaway = merge_scan(dev, gx, gy, gvx, gvy, px, py, karte.g_rr,
                                      karte.g_vis, karte.g_ast, m_alive, dt_years,
                                      barycentre=bary, two_gm=2.0 * G_AU * msum)
            pairs = [(int(m_alive[r]), int(j), float(fr)) for r, j, fr in hit]
        state["prev"] = (gx.clone(), gy.clone())
        hits, _, _, state["anteile"] = bounce_suche([karte], None, [(gx, gy, gvx, gvy)],
                                                    dt_years, rr_max[0], state["anteile"])
        bounce = None if hits is None else bounce_deltas(dev, out_np, hits, real_r, mass,
                                                         dt_years, BOUNCE_E)
        return {"k": k_ev, "sample": out_np, "pairs": pairs, "runaways": runaway,
                "bounce": bounce}

    def analyze_batch(outs, k0):
        g_all = karte.hochladen(outs)
        if state["dirty"]:
            rr_max[0] = stammdaten()
            state["dirty"] = False
        return [analyze_sample(outs[i], k0 + i, g_all[i]) for i in range(len(outs))]

    def apply(res):
        run = [] if res["runaways"] is None else res["runaways"]
        ev, merges, shatter = apply_merges(st, res["sample"], res["pairs"], run, res["k"],
                                           mirrors, contact, False, G_AU, 0.0, RASTER)
        assert shatter is None
        state["events"] += ev
        state["coll"] += merges
        if ev:
            mass[:] = mirrors["mass"].cpu().numpy()
            vis[:] = mirrors["vis"].cpu().numpy()
            real_r[:] = mirrors["real_r"].cpu().numpy()
            state["dirty"] = True
        if res["bounce"] is not None:
            hi, hj, aff, deltas = res["bounce"]
            sim.apply_deltas(st, aff, deltas)
            sample = res["sample"]
            for a_i, b_i in zip(hi.tolist(), hj.tolist()):
                heavy, light = (a_i, b_i) if mass[a_i] >= mass[b_i] else (b_i, a_i)
                state["events"] += struct.pack(
                    EV_FORMAT, 0.0 + (res["k"] + 1 + 0.0) * RASTER, heavy & 0xFFFFFFFF,
                    light, 0.0, 1, float(sample[heavy]), float(sample[n + heavy]))
            state["coll"] += len(hi)

    k, pending = 0, None
    for _ in range(batches):
        if pending is not None:
            for res in pending:
                apply(res)
            pending = None
        outs = sim.step_batch(st, dt_years, K)
        idx, t_rel, _ = st["hits"]
        for a_idx, t in zip(idx.tolist(), t_rel.tolist()):
            contact[a_idx] = k + float(t) / dt_years
        pending = analyze_batch(outs, k)
        subs = st["sub"]
        for i in range(K):
            if R.schreibe_slot(buf, ((k + i) % capacity) * sb, n, sub_max, outs[i],
                               subs[i] if subs else None):
                state["over"] += 1
        k += K
    for res in pending or []:
        apply(res)
    return ring, state, sim.export_f64(st)


def _defined_equal(got, exp, n, sub_max):
    head = 8 * n + 4
    assert got[:head].tobytes() == exp[:head].tobytes()
    nh = struct.unpack_from("<I", exp, 8 * n)[0]
    assert got[head:head + 4 * nh].tobytes() == exp[head:head + 4 * nh].tobytes()
    t0 = head + 4 * sub_max
    assert got[t0:t0 + 8 * M_SUB * nh].tobytes() == exp[t0:t0 + 8 * M_SUB * nh].tobytes()
    return nh


@pytest.mark.parametrize("kind,n_ast,batches,capacity", [
    ("belt", 20_000, 6, 48),        # no wrap
    ("clumps", 8_000, 6, 20),       # the ring wraps twice
])
def test_producer_matches_host_route(dev, kind, n_ast, batches, capacity):
    print(f"producer {kind} n_ast={n_ast} batches={batches} capacity={capacity}",
          flush=True)
    K = 8
    s = _scene(kind, n_ast)
    n = len(s["x"])
    sub_max = sub_max_fuer(n)
    exp_ring, exp, exp_state = host_route(s, batches, K, capacity, sub_max)

    sim, st = _load(s)
    ring = np.zeros((capacity, R.slot_bytes(n, M_SUB, sub_max)), np.uint8)
    prod = FilmProducer(sim, st, ring, RASTER, mass=s["mass"], real_r=s["real_r"],
                        visible=s["visible"], is_ast=s["is_ast"],
                        is_star_bh=s["is_star_bh"], sub_max=sub_max, K=K)
    rate = prod.run(batches)
    print(f"  {rate:.1f} sim-days/s, events {prod.n_events}, "
          f"collisions {prod.collisions}", flush=True)
    assert prod.k == prod.head == batches * K
    tracks = sum(_defined_equal(ring[c], exp_ring[c], n, sub_max)
                 for c in range(min(capacity, batches * K)))
    assert bytes(prod.events) == bytes(exp["events"])
    assert prod.collisions == exp["coll"] and prod.sub_overflow == exp["over"]
    assert np.array_equal(sim.export_f64(st), exp_state)
    kinds = [struct.unpack_from(EV_FORMAT, prod.events, EV_BYTES * q)[4]
             for q in range(prod.n_events)]
    if kind == "clumps":
        assert {0, 1, 2} <= set(kinds)       # merges, bounces and runaways occurred
        assert tracks > 0                    # hot asteroids near the planet
    prod.close()


def test_producer_errors(dev):
    s = scenes.film_belt(100)
    sim, st = _load(s)
    n = len(s["x"])
    sb = R.slot_bytes(n, M_SUB, 16)
    kw = dict(mass=s["mass"], real_r=s["real_r"], visible=s["visible"], is_ast=s["is_ast"],
              sub_max=16)
    with pytest.raises(ValueError):
        FilmProducer(sim, st, np.zeros((8, sb + 4), np.uint8), RASTER, **kw)
    with pytest.raises(ValueError):
        FilmProducer(sim, st, np.zeros((4, sb), np.uint8), RASTER, K=8, **kw)
    with pytest.raises(ValueError):
        FilmProducer(sim, st, np.zeros((32, sb), np.uint8), RASTER, K=17, **kw)
