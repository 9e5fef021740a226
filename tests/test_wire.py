"""The live wire path (server.py:15-23, 82-97, 1093-1115, 1152-1256): byte
layouts against hand-packed messages, the session's transitions and error
replies with a recording simulator stand-in (CPU), and FULL / STEP / DELTA
frames through the GPU simulator against direct NBodyCuda calls and the
oracle (GPU)."""
import asyncio
import json
import struct

import numpy as np
import pytest

from solarsystemsimulator_amd import wire as W


def scene(n=7, seed=3):
    rng = np.random.default_rng(seed)
    cols = [rng.normal(size=n) for _ in range(5)]
    return cols + [rng.integers(0, 2, n).astype(np.uint8),
                   (np.arange(n) > 1).astype(np.uint8)]


def test_full_layout_and_round_trip():
    x, y, vx, vy, m, vis, ast = scene()
    n = len(x)
    buf = W.build_full(0.25, x, y, vx, vy, m, vis, ast)
    hand = struct.pack("<IId", 0, n, 0.25)
    for c in (x, y, vx, vy, m):
        hand += struct.pack(f"<{n}d", *c)
    hand += bytes(vis.tolist()) + bytes(ast.tolist())
    assert buf == hand and len(buf) == 16 + 42 * n
    dt, cols, v, a = W.parse_full(buf)
    assert dt == 0.25
    for got, exp in zip(cols, (x, y, vx, vy, m)):
        assert np.array_equal(got, exp)
    assert np.array_equal(v, vis) and np.array_equal(a, ast)
    with pytest.raises(ValueError, match="Protokollfehler"):
        W.parse_full(buf + b"\0")
    with pytest.raises(ValueError, match="Protokollfehler"):
        W.parse_full(buf[:-1])


def test_delta_step_response_layouts():
    idx, vals = np.array([4, 0]), np.arange(8.0).reshape(2, 4)
    buf = W.build_delta(1e-3, idx, vals)
    hand = struct.pack("<IId", 2, 2, 1e-3)
    for i, v in zip(idx, vals):
        hand += struct.pack("<II4d", i, 0, *v)
    assert buf == hand and W.DELTA_REC.itemsize == 40
    dt, gi, gv = W.parse_delta(buf)
    assert dt == 1e-3 and np.array_equal(gi, idx) and np.array_equal(gv, vals)
    with pytest.raises(ValueError):
        W.parse_delta(buf[:-8])
    assert W.build_step(0.5) == struct.pack("<IId", 1, 0, 0.5)
    rows = np.arange(12, dtype=np.float32)
    assert W.build_response(3, rows) == struct.pack("<II12f", 0, 3, *rows)
    assert W.build_error("kaputt") == struct.pack("<II", 1, 0) + b"kaputt"
    got, err = W.parse_response(W.build_response(3, rows))
    assert err is None and np.array_equal(got, rows)
    assert W.parse_response(W.build_error("x")) == (None, "x")


class RecordingSim:
    """Test double of NBodyCuda's live interface: records calls and steps a
    trivial state so that the session's transitions can be checked on the
    CPU (the GPU tests below run the real simulator)."""

    def __init__(self):
        self.calls = []

    def load_state(self, x, y, vx, vy, mass, visible, is_ast):
        self.calls.append("load")
        return {"N": len(x), "rows": np.concatenate([x, y, vx, vy]).astype(np.float32)}

    def apply_updates(self, st, idx, vals):
        self.calls.append(("upd", idx.tolist()))
        n = st["N"]
        for i, v in zip(idx, vals):
            st["rows"][[i, n + i, 2 * n + i, 3 * n + i]] = v

    def step(self, st, dt):
        if st is None:
            raise ValueError("kein Zustand geladen")
        self.calls.append(("step", dt))
        return st["rows"].copy()


def test_session_transitions_and_errors():
    sim = RecordingSim()
    made = []
    s = W.LiveSession(lambda: made.append(1) or sim, device_name="MI355X")
    assert json.loads(s.ping()) == {"backend": "cuda", "device": "MI355X", "mMax": W.M_MAX}
    assert s.handle("ping") == s.ping() and not made           # no simulator for a ping
    assert W.parse_response(s.handle(W.build_step(1.0)))[1] == "kein Zustand geladen"
    assert "DELTA ohne Zustand" in W.parse_response(s.handle(W.build_delta(1.0, [0], [[0] * 4])))[1]
    x, y, vx, vy, m, vis, ast = scene()
    rows, err = W.parse_response(s.handle(W.build_full(0.5, x, y, vx, vy, m, vis, ast)))
    assert err is None and np.array_equal(rows, np.concatenate([x, y, vx, vy]).astype(np.float32))
    rows, _ = W.parse_response(s.handle(W.build_delta(0.25, [2], [[9, 8, 7, 6]])))
    assert rows[[2, 9, 16, 23]].tolist() == [9, 8, 7, 6]
    assert "DELTA-Index 7" in W.parse_response(s.handle(W.build_delta(0.25, [7], [[0] * 4])))[1]
    assert "unbekannter Nachrichtentyp 9" in W.parse_response(
        s.handle(struct.pack("<IId", 9, 0, 0.0)))[1]
    assert "Filmmodus" in W.parse_response(s.handle(struct.pack("<IId", 3, 0, 0.0)))[1]
    assert "Filmmodus" in W.parse_response(s.handle(struct.pack("<IId", 7, 0, 0.0)))[1]
    # FILM_STOP with no film running is ignored without a reply
    # (server.py:1195-1209), so it cannot be taken for the next STEP's reply
    assert s.handle(struct.pack("<IId", W.MSG_FILM_STOP, 0, 0.0)) is None
    assert "Protokollfehler" in W.parse_response(
        s.handle(W.build_full(0.5, x, y, vx, vy, m, vis, ast)[:-1]))[1]
    assert sim.calls == ["load", ("step", 0.5), ("upd", [2]), ("step", 0.25)]
    assert (s.frames, s.fulls, len(made)) == (2, 1, 1)


def test_serve_loop():
    class FakeWs:
        def __init__(self, msgs):
            self.msgs, self.sent = list(msgs), []

        def __aiter__(self):
            return self

        async def __anext__(self):
            if not self.msgs:
                raise StopAsyncIteration
            return self.msgs.pop(0)

        async def send(self, b):
            self.sent.append(b)

    x, y, vx, vy, m, vis, ast = scene()
    ws = FakeWs(["hallo", W.build_full(0.5, x, y, vx, vy, m, vis, ast), W.build_step(0.5)])
    asyncio.run(W.serve(ws, W.LiveSession(RecordingSim)))
    assert len(ws.sent) == 3 and isinstance(ws.sent[0], str)
    assert all(W.parse_response(b)[1] is None for b in ws.sent[1:])


@pytest.mark.gpu
def test_gpu_live_frames_match_direct_calls_and_oracle(dev):
    import oracle
    from oracle import nbody_np as nb
    from solarsystemsimulator_amd.nbody import NBodyCuda
    print("wire live frames, make_system(300, seed 7)", flush=True)
    st0 = nb.make_system(300, seed=7)
    day = 1 / 365.25
    sim = NBodyCuda(0)
    s = W.LiveSession(lambda: sim, device_name=sim.name())
    assert W.parse_response(s.handle(W.build_step(day)))[1]          # no state yet
    got, err = W.parse_response(s.handle(W.build_full(day, *st0)))
    assert err is None
    ref, _, _, _ = oracle.nbody_frames(*st0, None, day, 1)
    assert np.abs(got - ref[0]).max() <= 1e-6 * np.abs(ref[0]).max()
    direct = NBodyCuda(0)
    st = direct.load_state(*st0)
    assert np.array_equal(got, direct.step(st, day))
    got2, _ = W.parse_response(s.handle(W.build_step(day)))
    assert np.array_equal(got2, direct.step(st, day))
    idx = np.array([0, 5, 120])
    vals = np.array([[0.0, 0.0, 0.0, 0.0], [1.1, -0.2, 0.3, 6.0], [2.5, 0.1, -0.4, 3.9]])
    got3, err = W.parse_response(s.handle(W.build_delta(0.5 * day, idx, vals)))
    assert err is None
    direct.apply_updates(st, idx, vals)
    assert np.array_equal(got3, direct.step(st, 0.5 * day))
    assert "DELTA-Index" in W.parse_response(
        s.handle(W.build_delta(day, [len(st0[0])], [[0] * 4])))[1]
    got4, _ = W.parse_response(s.handle(W.build_step(day)))          # still usable
    assert np.array_equal(got4, direct.step(st, day))
