"""The golden fixtures of tests/golden: the oracle reproduces them (CPU), and
the HIP path matches them through the drop-in API (GPU).  The fixtures'
expectations come from independent routes (exactly rounded sums, closed-form
cell indices, the brute-force pair set); tests/golden/make_golden.py writes
them."""
import numpy as np
import pytest

import oracle
from oracle import collide_np as C
from oracle import merge_np as M
from oracle import mlp_np as MLP
from oracle import pm_np
from tests.golden.make_golden import cic_closed_form, load

DT_Y = 0.5 / 365.25


# --- CPU: the oracle against the fixtures ----------------------------------

def test_oracle_allpairs_matches_exact_sums():
    g = load("allpairs_wolke")
    ax, ay = oracle.allpairs([g["x"], g["y"]], g["m"], [g["x"], g["y"]],
                             float(g["G"]), float(g["eps2"]))
    scale = np.abs(np.concatenate([g["ax"], g["ay"]])).max()
    assert np.abs(ax - g["ax"]).max() <= 1e-13 * scale
    assert np.abs(ay - g["ay"]).max() <= 1e-13 * scale


def _nbody_inputs(g):
    return [g[k] for k in ("x", "y", "vx", "vy", "mass", "visible", "is_ast")]


def test_oracle_nbody_reproduces_fixture():
    g = load("nbody_seed7")
    out, state, sub, hits = oracle.nbody_frames(*_nbody_inputs(g), g["real_r"],
                                                float(g["raster"]), int(g["k"]))
    assert np.array_equal(out, g["out"]) and np.array_equal(state, g["state"])
    assert np.array_equal([len(i) for i, _ in sub], g["sub_counts"])
    assert np.array_equal(np.concatenate([i for i, _ in sub]), g["sub_idx"])
    assert np.array_equal(hits[0], g["hit_idx"]) and np.array_equal(hits[2], g["hit_partner"])


@pytest.mark.parametrize("tag", ["2d", "3d"])
def test_oracle_cic_cells_and_mesh(tag):
    g = load("cic_seed42")
    pos, m = list(g[f"{tag}_pos"]), g[f"{tag}_m"]
    n, origin, h = list(g[f"{tag}_n"]), list(g[f"{tag}_origin"]), float(g[f"{tag}_h"])
    periodic = bool(g[f"{tag}_periodic"])
    cells, _ = pm_np.cic_cells(pos, origin, h, n, periodic)
    assert np.array_equal(cells, g[f"{tag}_cells"])
    assert np.array_equal(cic_closed_form(pos, origin, h, n, periodic), g[f"{tag}_cells"])
    mesh = pm_np.cic_deposit(pos, m, n, origin, h, periodic)
    np.testing.assert_allclose(mesh, g[f"{tag}_mesh"], rtol=1e-14, atol=1e-15 * m.sum())
    assert abs(mesh.sum() - m.sum()) <= 1e-12 * m.sum()          # mass conserved


@pytest.mark.parametrize("scene", ["klumpen", "auf-der-grenze"])
def test_oracle_pair_list_and_brute_force(scene):
    g = load("collide_4711")
    x, y, vx, vy, rr = g[f"{scene}_in"]
    ones = np.ones(len(x), np.uint8)
    h, halo = float(g[f"{scene}_h"]), float(g[f"{scene}_halo"])
    (i, j), cand = C.hits(x, y, vx, vy, rr, ones, ones, DT_Y, h, halo)
    assert cand == int(g[f"{scene}_cand"])
    assert np.array_equal(i, g[f"{scene}_i"]) and np.array_equal(j, g[f"{scene}_j"])
    brute = C.brute_force(x, y, vx, vy, rr, ones, ones, DT_Y, h)
    assert set(zip(i.tolist(), j.tolist())) == brute


def test_oracle_bounce_reproduces_fixture():
    g = load("collide_4711")
    x, y, vx, vy, rr = g["klumpen_in"]
    sample = np.concatenate([x, y, vx, vy]).astype(np.float32)
    hi, hj, hit, dl = M.bounce_deltas(sample, g["klumpen_i"].astype(np.int64),
                                      g["klumpen_j"].astype(np.int64),
                                      rr.astype(np.float64), g["bounce_mass"], DT_Y)
    assert np.array_equal(hi, g["bounce_hi"]) and np.array_equal(hj, g["bounce_hj"])
    assert np.array_equal(hit, g["bounce_hit"])
    assert np.array_equal(dl, g["bounce_deltas"])


def test_oracle_mlp_forward_reproduces_fixture():
    g = load("mlp_forward")
    out, _ = MLP.forward(MLP.bf16(g["params"]), g["x"], rounding=False)
    np.testing.assert_allclose(out, g["out"], rtol=1e-12, atol=1e-12)


# --- GPU: the HIP path against the fixtures --------------------------------

@pytest.mark.gpu
def test_gpu_allpairs_fixture(dev):
    from solarsystemsimulator_amd import _lib
    from tests.test_selfgrav_gpu import gpu_allpairs
    g = load("allpairs_wolke")
    print("golden allpairs_wolke", flush=True)
    ax, ay = gpu_allpairs(dev, [g["x"], g["y"]], g["m"], [g["x"], g["y"]],
                          float(g["G"]), float(g["eps2"]), _lib.PREC_F64)
    scale = np.abs(np.concatenate([g["ax"], g["ay"]])).max()
    assert np.abs(ax - g["ax"]).max() <= 1e-12 * scale
    assert np.abs(ay - g["ay"]).max() <= 1e-12 * scale


@pytest.mark.gpu
def test_gpu_nbody_fixture(dev):
    from solarsystemsimulator_amd.nbody import NBodyCuda
    g = load("nbody_seed7")
    print("golden nbody_seed7", flush=True)
    sim = NBodyCuda(0)
    st = sim.load_state(*_nbody_inputs(g), real_r=g["real_r"])
    out = sim.step_batch(st, float(g["raster"]), int(g["k"]))
    state = sim.export_f64(st)
    assert np.abs(state - g["state"]).max() <= 1e-12 * np.abs(g["state"]).max()
    assert np.abs(out - g["out"]).max() <= 1e-6 * np.abs(g["out"]).max()
    assert np.array_equal([len(i) for i, _ in st["sub"]], g["sub_counts"])
    assert np.array_equal(np.concatenate([i for i, _ in st["sub"]]), g["sub_idx"])
    np.testing.assert_allclose(np.concatenate([p for _, p in st["sub"]], axis=-1),
                               g["sub_pos"], rtol=1e-6, atol=1e-7)
    hits = st["hits"]
    assert np.array_equal(hits[0], g["hit_idx"]) and np.array_equal(hits[2], g["hit_partner"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["2d", "3d"])
def test_gpu_cic_fixture(dev, tag):
    import torch
    from solarsystemsimulator_amd import pm as P
    g = load("cic_seed42")
    print(f"golden cic {tag}", flush=True)
    pos, m = g[f"{tag}_pos"], g[f"{tag}_m"]
    n, origin, h = list(g[f"{tag}_n"]), list(g[f"{tag}_origin"]), float(g[f"{tag}_h"])
    dim, npart = len(n), len(m)
    cells = torch.empty(dim * npart, dtype=torch.int32, device=dev)
    tp = [torch.as_tensor(np.ascontiguousarray(c)).to(dev) for c in pos]
    mesh = P.deposit(tp, torch.as_tensor(m).to(dev), n, origin, h,
                     bool(g[f"{tag}_periodic"]), cells=cells)
    assert np.array_equal(cells.cpu().numpy().reshape(npart, dim).T, g[f"{tag}_cells"])
    np.testing.assert_allclose(mesh.cpu().numpy(), g[f"{tag}_mesh"], rtol=1e-12,
                               atol=1e-14 * m.sum())


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["klumpen", "auf-der-grenze"])
def test_gpu_pair_list_fixture(dev, scene):
    import torch
    from solarsystemsimulator_amd.detection import Erkennungskarte
    g = load("collide_4711")
    print(f"golden collide {scene}", flush=True)
    x, y, vx, vy, rr = g[f"{scene}_in"]
    ones = np.ones(len(x), np.uint8)
    karte = Erkennungskarte(0, ones != 0, -np.inf, np.inf)
    karte.stammdaten(ones, rr)
    fields = [torch.as_tensor(np.ascontiguousarray(a)).to(dev) for a in (x, y, vx, vy)]
    got, cand = karte.bounce_hits(*fields, DT_Y, float(g[f"{scene}_h"]),
                                  float(g[f"{scene}_halo"]))
    assert cand == int(g[f"{scene}_cand"])
    assert np.array_equal(got[0], g[f"{scene}_i"]) and np.array_equal(got[1], g[f"{scene}_j"])


@pytest.mark.gpu
def test_gpu_bounce_fixture(dev):
    from solarsystemsimulator_amd.detection import bounce_deltas
    g = load("collide_4711")
    print("golden bounce", flush=True)
    x, y, vx, vy, rr = g["klumpen_in"]
    sample = np.concatenate([x, y, vx, vy]).astype(np.float32)
    got = bounce_deltas(dev, sample, (g["klumpen_i"].astype(np.int64),
                                      g["klumpen_j"].astype(np.int64)),
                        rr.astype(np.float64), g["bounce_mass"], DT_Y)
    assert np.array_equal(got[0], g["bounce_hi"]) and np.array_equal(got[1], g["bounce_hj"])
    assert np.array_equal(got[2], g["bounce_hit"])
    np.testing.assert_allclose(got[3], g["bounce_deltas"], rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_gpu_mlp_fixture(dev):
    import torch
    from solarsystemsimulator_amd.learned import LearnedCorrection
    from tests.test_mlp_gpu import to_bf16_tensor
    g = load("mlp_forward")
    print("golden mlp_forward", flush=True)
    net = LearnedCorrection(dev, g["params"])
    out = torch.empty((len(g["x"]), 3), dtype=torch.float32, device=dev)
    net.predict(to_bf16_tensor(g["x"], dev), out=out)
    # BF16 activations between the layers: the same bound as the network's
    # own forward-pass test (relative L2 error of the BF16-rounded network)
    ref = g["out"]
    assert np.linalg.norm(out.cpu().numpy() - ref) <= 2e-2 * np.linalg.norm(ref)
