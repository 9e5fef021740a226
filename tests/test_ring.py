"""Film ring slots (film_producer.py:76-123): the oracle against hand-packed
slots (CPU), and the device-written slots of NBodyCuda.step_batch_slots
against the reference's host route, step_batch + schreibe_slot per sample,
bit-exact in every defined byte (GPU)."""
import struct

import numpy as np
import pytest

from oracle import ring_np as R


def test_slot_bytes_matches_abi():
    from solarsystemsimulator_amd import _lib
    lib = _lib.load()
    for n, ms, sm in ((0, 0, 0), (10, 8, 2048), (50_005, 32, 2048), (7, 0, 3)):
        assert lib.ssim_ring_slot_bytes(n, ms, sm) == R.slot_bytes(n, ms, sm)
    assert R.slot_bytes(10, 8, 2048) == 80 + 4 + 2048 * 68


def test_schreibe_slot_hand_packed():
    n, ms, sm = 3, 2, 2
    row = np.arange(12, dtype=np.float32) * 0.5            # x|y|vx|vy
    idx = np.array([2, 0, 1])
    pos = np.arange(12, dtype=np.float32).reshape(ms, 2, 3) + 100
    buf = bytearray(R.slot_bytes(n, ms, sm))
    assert R.schreibe_slot(buf, 0, n, sm, row, (idx, pos)) is True
    exp = struct.pack("<6f", *row[:6]) + struct.pack("<III", 2, 0, 1)
    for b in (0, 1):                                         # lowest indices kept
        col = int(np.flatnonzero(idx == b)[0])
        for s in range(ms):
            exp += struct.pack("<2f", pos[s, 0, col], pos[s, 1, col])
    assert bytes(buf) == exp
    buf2 = bytearray(b"\xff" * len(buf))
    assert R.schreibe_slot(buf2, 0, n, sm, row, None) is False
    assert buf2[:28] == struct.pack("<6fI", *row[:6], 0) and buf2[28:] == b"\xff" * (len(buf) - 28)
    assert R.schreibe_slot(buf2, 0, n, 0, row, (idx, pos)) is False


def _defined_equal(got, exp, n, ms, sm):
    head = 8 * n + 4
    assert got[:head].tobytes() == exp[:head].tobytes()
    nh = struct.unpack_from("<I", exp, 8 * n)[0]
    assert got[head:head + 4 * nh].tobytes() == exp[head:head + 4 * nh].tobytes()
    t0 = head + 4 * sm
    assert got[t0:t0 + 8 * ms * nh].tobytes() == exp[t0:t0 + 8 * ms * nh].tobytes()
    return nh


def _scene(n_ast, seed=11):
    from oracle import nbody_np as nb
    st0 = list(nb.make_system(n_ast, seed=seed))
    if n_ast >= 40:                                          # divers: complete fine tracks
        st0[0][10:40] = 0.3 * np.cos(np.arange(30.0))
        st0[1][10:40] = 0.3 * np.sin(np.arange(30.0))
        st0[2][10:40] = 0.0
        st0[3][10:40] = 0.0
    rr = np.full(len(st0[0]), 1e-5)
    rr[0] = 0.00465
    return st0, rr


@pytest.mark.gpu
@pytest.mark.parametrize("n_ast,m_sub,sub_max,k,massive", [
    (3000, 8, 5, 4, True), (3000, 8, 2048, 3, True), (200, 0, 16, 2, True),
    (0, 8, 16, 2, True), (20000, 32, 7, 16, True),
    # sub_max 0 with tracks due: count field 0, no overflow (schreibe_slot's
    # `not sub_max` branch)
    (500, 8, 0, 2, True),
    # no massive bodies at all: the massive rows and index map are empty
    (400, 8, 16, 2, False)])
def test_gpu_slots_match_host_route(dev, n_ast, m_sub, sub_max, k, massive):
    from solarsystemsimulator_amd.nbody import NBodyCuda
    print(f"ring slots n_ast={n_ast} m_sub={m_sub} sub_max={sub_max} k={k} "
          f"massive={massive}", flush=True)
    st0, rr = _scene(n_ast)
    if not massive:                     # every body an asteroid
        st0[6] = np.ones_like(st0[6])
    # a 5-day raster: above the 0.5-day largest sub-step, so the divers get
    # complete fine tracks (a 0.5-day raster would need no sub-steps)
    raster = 5 / 365.25
    host_sim, dev_sim = NBodyCuda(0, m_sub=m_sub), NBodyCuda(0, m_sub=m_sub)
    sa = host_sim.load_state(*st0, real_r=rr)
    sb_ = dev_sim.load_state(*st0, real_r=rr)
    n = len(st0[0])
    size = R.slot_bytes(n, m_sub, sub_max)
    ring = np.zeros((2 * k + 1, size), np.uint8)             # a ring with one spare slot
    seen = 0
    for batch in range(2):
        out = host_sim.step_batch(sa, raster, k)
        exp = np.zeros((k, size), np.uint8)
        ovf = []
        for f in range(k):                                   # the ring is a byte buffer
            slot = bytearray(size)
            ovf.append(R.schreibe_slot(slot, 0, n, sub_max, out[f],
                                       None if sa["sub"] is None else sa["sub"][f]))
            exp[f] = np.frombuffer(slot, np.uint8)
        dest = ring[1 + batch * k:1 + (batch + 1) * k]
        got, govf = dev_sim.step_batch_slots(sb_, raster, k, sub_max, out=dest)
        assert got is dest
        nhs = [_defined_equal(got[f], exp[f], n, m_sub, sub_max) for f in range(k)]
        assert list(govf) == ovf
        for a, b in zip(sa["hits"], sb_["hits"]):
            assert np.array_equal(a, b)
        seen += sum(nhs)
    if n_ast >= 40 and m_sub and sub_max and massive:
        assert seen > 0                                      # the case has tracks
    if not sub_max:
        assert seen == 0
    assert not ring[0].any()                                 # nothing outside its slots
    assert np.array_equal(host_sim.export_f64(sa), dev_sim.export_f64(sb_))


@pytest.mark.gpu
def test_gpu_slots_errors(dev):
    from solarsystemsimulator_amd.nbody import NBodyCuda
    st0, rr = _scene(100)
    sim = NBodyCuda(0)
    st = sim.load_state(*st0, real_r=rr)
    with pytest.raises(ValueError):
        sim.step_batch_slots(None, 0.01, 1, 16)
    with pytest.raises(ValueError):
        sim.step_batch_slots(st, 0.01, 0, 16)
    with pytest.raises(ValueError):
        sim.step_batch_slots(st, 0.01, 2, -1)
    with pytest.raises(ValueError):
        sim.step_batch_slots(st, 0.01, 2, 16, out=np.zeros((2, 5), np.uint8))
