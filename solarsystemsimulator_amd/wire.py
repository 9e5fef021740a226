"""Live wire path of the simulator's WebSocket server on the GPU backend.

The browser drives the live (non-film) physics with three binary messages
(server.py:15-23, little-endian):

  FULL:     u32 type=0 | u32 N | f64 dtYears |
            x[N] f64 | y[N] | vx[N] | vy[N] | mass[N] | visible[N] u8 | isAst[N] u8
  STEP:     u32 type=1 | u32 pad | f64 dtYears
  DELTA:    u32 type=2 | u32 count | f64 dtYears |
            count x (u32 idx | u32 pad | f64 x | y | vx | vy)
  Response: u32 status=0 | u32 N | x[N] f32 | y[N] | vx[N] | vy[N]
            (status 1: u32 0, then the UTF-8 error text)

FULL uploads a state and steps it; STEP steps the resident state; DELTA
overwrites x/y/vx/vy of some bodies, then steps.  Every frame answers with
the f32 render rows of one `NBodyCuda.step`.  A text message is the client's
auto-detection ping and is answered with the backend's JSON description.

`LiveSession` is the per-connection state machine of server.py:1152-1256
(live branch), independent of the transport; `serve` runs it over any
object that yields messages asynchronously and has an async `send` (a
`websockets` connection is one).  The state stays resident in HBM between
frames; the only host traffic per STEP frame is 16 bytes in and 16 N bytes
out.  Film messages (types 3, 4, 7) belong to the film producer, which runs
in its own process; without a film handler they are answered with an error.
"""
from __future__ import annotations

import asyncio
import json
import struct
import threading

import numpy as np

from .constants import M_MAX

HEADER = struct.Struct("<IId")          # type, N / pad / count, dtYears (server.py:82)
MSG_FULL = 0
MSG_STEP = 1
MSG_DELTA = 2
MSG_FILM_START = 3
MSG_FILM_STOP = 4
MSG_FILM_SUB = 7
FILM_TYPES = (MSG_FILM_START, MSG_FILM_STOP, MSG_FILM_SUB)
# one DELTA record: u32 idx | u32 pad | f64 x, y, vx, vy (server.py:96-97)
DELTA_REC = np.dtype([("idx", "<u4"), ("pad", "<u4"), ("v", "<f8", (4,))])


def parse_full(buf: bytes):
    """(dtYears, [x, y, vx, vy, mass], visible, isAst) of a FULL message,
    views into `buf`; ValueError unless its length is exactly the header
    plus 42 N bytes (server.py:1093-1107)."""
    _typ, n, dt_years = HEADER.unpack_from(buf, 0)
    need = HEADER.size + 42 * n
    if len(buf) != need:
        raise ValueError(f"Protokollfehler: {len(buf)} Bytes, erwartet {need}")
    off = HEADER.size
    cols = np.frombuffer(buf, "<f8", 5 * n, off).reshape(5, n)
    flags = np.frombuffer(buf, np.uint8, 2 * n, off + 40 * n).reshape(2, n)
    return dt_years, list(cols), flags[0], flags[1]


def build_full(dt_years, x, y, vx, vy, mass, visible, is_ast) -> bytes:
    """The client side of FULL (index.html's encoder), for tests and tools."""
    n = len(x)
    cols = np.stack([np.asarray(c, "<f8") for c in (x, y, vx, vy, mass)])
    flags = np.stack([np.asarray(visible, np.uint8), np.asarray(is_ast, np.uint8)])
    return HEADER.pack(MSG_FULL, n, dt_years) + cols.tobytes() + flags.tobytes()


def build_step(dt_years) -> bytes:
    return HEADER.pack(MSG_STEP, 0, dt_years)


def build_delta(dt_years, idx, vals) -> bytes:
    recs = np.zeros(len(idx), DELTA_REC)
    recs["idx"] = idx
    recs["v"] = np.asarray(vals, np.float64).reshape(len(idx), 4)
    return HEADER.pack(MSG_DELTA, len(idx), dt_years) + recs.tobytes()


def parse_delta(buf: bytes):
    """(dtYears, idx int64, vals f64 (count, 4)); ValueError when the
    message is shorter than its count of records."""
    _typ, count, dt_years = HEADER.unpack_from(buf, 0)
    recs = np.frombuffer(buf, DELTA_REC, count, HEADER.size)
    return dt_years, recs["idx"].astype(np.int64), recs["v"]


def build_response(n: int, f32_state: np.ndarray) -> bytes:
    """status 0, N, then the (4N,) f32 rows x | y | vx | vy (server.py:1110-1111)."""
    return struct.pack("<II", 0, n) + np.ascontiguousarray(f32_state, "<f4").tobytes()


def build_error(msg: str) -> bytes:
    """status 1, 0, then the UTF-8 text (server.py:1114-1115)."""
    return struct.pack("<II", 1, 0) + msg.encode()


def parse_response(buf: bytes):
    """(f32 (4N,) rows, None) for status 0, (None, text) otherwise."""
    status, n = struct.unpack_from("<II", buf, 0)
    if status != 0:
        return None, buf[8:].decode()
    return np.frombuffer(buf, "<f4", 4 * n, 8), None


class LiveSession:
    """One client connection of the live path: the resident state and the
    FULL / STEP / DELTA transitions of server.py:1175-1251.

    `sim_factory` returns the shared `NBodyCuda`; it is called lazily on the
    first frame that needs it (the server builds the simulator only when a
    client uses the live path, server.py:1118-1137) and at most once per
    session.  `film` optionally handles film messages; it receives the
    message and returns the bytes to send or None."""

    def __init__(self, sim_factory, device_name: str = "", film=None):
        self._factory = sim_factory
        self._sim = None
        self._lock = threading.Lock()
        self.device_name = device_name
        self.film = film
        self.state = None
        self.frames = 0
        self.fulls = 0

    def sim(self):
        with self._lock:
            if self._sim is None:
                self._sim = self._factory()
            return self._sim

    def ping(self) -> str:
        """The answer to the auto-detection ping: backend, device and the
        massive-body limit the client checks before uploading
        (server.py:1164-1173; the client reads `mMax`, index.html:5425-5428)."""
        return json.dumps({"backend": "cuda", "device": self.device_name,
                           "mMax": M_MAX}, separators=(",", ":"))

    def handle(self, message):
        """The reply to one message (str for a ping, bytes otherwise, None
        when nothing is sent).  Errors of a frame become an error response,
        as in the reference; the session stays usable."""
        if isinstance(message, str):
            return self.ping()
        try:
            typ, _n, dt_years = HEADER.unpack_from(message, 0)
            if typ in FILM_TYPES:
                if self.film is None:
                    # FILM_STOP without an active film is ignored, with no
                    # reply (server.py:1195-1209); FILM_SUB / FILM_START
                    # report that no film runs
                    if typ == MSG_FILM_STOP:
                        return None
                    raise ValueError(f"Filmmodus nicht verfuegbar (Nachrichtentyp {typ})")
                return self.film(message)
            sim = self.sim()
            if typ == MSG_FULL:
                dt_years, (x, y, vx, vy, mass), visible, is_ast = parse_full(message)
                self.state = sim.load_state(x, y, vx, vy, mass, visible, is_ast)
                self.fulls += 1
            elif typ == MSG_DELTA:
                if self.state is None:
                    raise ValueError("DELTA ohne Zustand (erst FULL senden)")
                _, idx, vals = parse_delta(message)
                if len(idx) and (idx.max() >= self.state["N"]):
                    raise ValueError(f"DELTA-Index {int(idx.max())} >= N={self.state['N']}")
                sim.apply_updates(self.state, idx, vals)
            elif typ != MSG_STEP:
                raise ValueError(f"unbekannter Nachrichtentyp {typ}")
            out = sim.step(self.state, dt_years)
            self.frames += 1
            return build_response(len(out) // 4, out)
        except Exception as e:                              # reported to the client
            return build_error(str(e))


async def serve(ws, session: LiveSession) -> None:
    """The connection loop of server.py:1152-1256 over `ws` (async iterable
    of messages with an async `send`); GPU work runs in a worker thread so
    that the event loop keeps serving other connections."""
    async for message in ws:
        reply = await asyncio.to_thread(session.handle, message)
        if reply is not None:
            await ws.send(reply)
