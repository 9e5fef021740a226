"""CPU oracle of the GPU physics backend — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, as the checker or as the timed CPU baseline.  The
product package never imports it.

* ``liboracle.so`` (built from ``ssim_oracle.c`` by ``make -C oracle``):
  long-double all-pairs force, energy and the KDK / Yoshida-4 self-gravity
  integrator; see the header of ssim_oracle.c for the reference lines each
  function follows and how they are pinned.
* NumPy restatements in the sibling modules (one module per reference area).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", os.fspath(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(os.fspath(LIB))
        d = ctypes.c_double
        p = ctypes.c_void_p
        i = ctypes.c_int
        _lib.oracle_allpairs.argtypes = [i, i, p, p, p, p, i, p, p, p, p, p, p, d, d]
        _lib.oracle_allpairs_f64_plain.argtypes = _lib.oracle_allpairs.argtypes
        _lib.oracle_energy.argtypes = [i, i, p, p, p, p, p, p, p, d, d, p]
        _lib.oracle_selfgrav_run.argtypes = [i, i, i, i, p, p, p, p, p, p, p,
                                             p, p, p, d, d, d, i, i]
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data


def allpairs(src_pos, src_m, tgt_pos, G: float, eps2: float, plain=False):
    """a(targets) from sources; positions are lists of dim arrays."""
    dim = len(src_pos)
    sp = [_f64(c) for c in src_pos] + [None] * (3 - dim)
    tp = [_f64(c) for c in tgt_pos] + [None] * (3 - dim)
    m = _f64(src_m)
    nt = len(tp[0])
    out = [np.zeros(nt) for _ in range(dim)] + [None] * (3 - dim)
    fn = lib().oracle_allpairs_f64_plain if plain else lib().oracle_allpairs
    fn(dim, len(m), _p(sp[0]), _p(sp[1]), _p(sp[2]), _p(m), nt, _p(tp[0]),
       _p(tp[1]), _p(tp[2]), _p(out[0]), _p(out[1]), _p(out[2]), G, eps2)
    return out[:dim]


def energy(pos, vel, m, G: float, eps2: float) -> np.ndarray:
    dim = len(pos)
    p = [_f64(c) for c in pos] + [None] * (3 - dim)
    v = [_f64(c) for c in vel] + [None] * (3 - dim)
    m = _f64(m)
    out = np.zeros(3)
    lib().oracle_energy(dim, len(m), _p(p[0]), _p(p[1]), _p(p[2]), _p(v[0]),
                        _p(v[1]), _p(v[2]), _p(m), G, eps2, _p(out))
    return out


def selfgrav_run(pos, vel, m, tpos, tvel, G, eps2, dt, steps, scheme=0,
                 acc=None):
    """Integrate n masses (+ t tracers); returns (pos, vel, acc) lists of
    dim arrays of length n + t.  `acc` (if given) are the current
    accelerations, otherwise they are computed first."""
    dim = len(pos)
    n = len(m)
    if tpos is None:
        tpos = [np.zeros(0)] * dim
        tvel = [np.zeros(0)] * dim
    t = len(tpos[0])
    x = [_f64(np.concatenate([pos[c], tpos[c]])) for c in range(dim)]
    v = [_f64(np.concatenate([vel[c], tvel[c]])) for c in range(dim)]
    a = ([np.zeros(n + t) for _ in range(dim)] if acc is None
         else [_f64(c).copy() for c in acc])
    xs = x + [None] * (3 - dim)
    vs = v + [None] * (3 - dim)
    as_ = a + [None] * (3 - dim)
    mm = _f64(m)
    lib().oracle_selfgrav_run(dim, scheme, n, t, _p(xs[0]), _p(xs[1]),
                              _p(xs[2]), _p(vs[0]), _p(vs[1]), _p(vs[2]),
                              _p(as_[0]), _p(as_[1]), _p(as_[2]), _p(mm), G,
                              eps2, dt, steps, 1 if acc is None else 0)
    return x, v, a


def nbody_frames(x, y, vx, vy, mass, visible, is_ast, real_r, dt_raster, k,
                 m_sub=8, G=4 * np.pi ** 2, soft=1e-6,
                 max_sub_dt=0.5 / 365.25, max_sub_steps=100_000):
    """Serial restatement of one `step_batch` (nbody_kernel.py:783-859):
    returns (out f32 (k, 4n) in original order, state f64 (4n), sub,
    hits) with sub/hits shaped like NBodyCuda's st["sub"] / st["hits"].
    The inputs are not modified."""
    return _frames(2, [x, y], [vx, vy], mass, visible, is_ast, real_r,
                   dt_raster, k, m_sub, G, soft, max_sub_dt, max_sub_steps)


def nbody_frames3d(x, y, z, vx, vy, vz, mass, visible, is_ast, real_r,
                   dt_raster, k, m_sub=8, G=4 * np.pi ** 2, soft=1e-6,
                   max_sub_dt=0.5 / 365.25, max_sub_steps=100_000):
    """The same batch in space: out f32 (k, 6n) [x|y|z|vx|vy|vz], state f64
    (6n), sub pos (m_sub, 3, nh)."""
    return _frames(3, [x, y, z], [vx, vy, vz], mass, visible, is_ast, real_r,
                   dt_raster, k, m_sub, G, soft, max_sub_dt, max_sub_steps)


def _frames(dim, pos, vel, mass, visible, is_ast, real_r, dt_raster, k,
            m_sub, G, soft, max_sub_dt, max_sub_steps):
    c = ctypes
    F = 2 * dim
    ast = np.asarray(is_ast) != 0
    m_idx, a_idx = np.flatnonzero(~ast), np.flatnonzero(ast)
    na, m = len(a_idx), len(m_idx)
    n = len(pos[0])
    cols = [np.asarray(v, np.float64) for v in list(pos) + list(vel) + [mass]]
    a = [_f64(v[a_idx]).copy() for v in cols]
    M = [_f64(v[m_idx]).copy() for v in cols]
    vis = np.asarray(visible).astype(np.uint8)
    rr = np.zeros(n) if real_r is None else np.asarray(real_r, np.float64)
    avis, mvis = np.ascontiguousarray(vis[a_idx]), np.ascontiguousarray(vis[m_idx])
    arad, mrad = _f64(rr[a_idx]).copy(), _f64(rr[m_idx]).copy()
    aacc = [np.zeros(max(na, 1)) for _ in range(dim)]
    macc = [np.zeros(max(m, 1)) for _ in range(dim)]
    ms = max(m_sub, 1)
    snap = np.zeros(k * F * max(na, 1), np.float32)
    snapm = np.zeros(k * F * max(m, 1), np.float32)
    subpos = np.zeros(k * ms * dim * max(na, 1), np.float32)
    subn = np.zeros(k * max(na, 1), np.uint8)
    hitt = np.full(max(na, 1), -1.0, np.float32)
    hitm = np.zeros(max(na, 1), np.int32)
    hot = np.zeros(max(na, 1), np.uint8)
    w1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
    w0 = -(2.0 ** (1.0 / 3.0)) / (2.0 - 2.0 ** (1.0 / 3.0))
    fn = lib().oracle_nbody_frame if dim == 2 else lib().oracle_nbody_frame3d
    P = c.c_void_p
    nv = 3 * dim + 3                     # pos, vel, acc, mass, vis, radius
    fn.argtypes = ([c.c_int] + [P] * nv + [c.c_int] + [P] * nv +
                   [c.c_double, c.c_double, c.c_double, c.c_int, c.c_double,
                    c.c_int, c.c_double, c.c_double, c.c_int] + [P] * 7)
    fn(na, *[_p(v) for v in a[:F]], *[_p(v) for v in aacc], _p(a[F]),
       _p(avis), _p(arad), m, *[_p(v) for v in M[:F]], *[_p(v) for v in macc],
       _p(M[F]), _p(mvis), _p(mrad), G, soft, dt_raster, k,
       max_sub_dt, max_sub_steps, w1, w0, m_sub, _p(snap), _p(snapm),
       _p(subpos), _p(subn), _p(hitt), _p(hitm), _p(hot))
    out = np.empty((k, F * n), np.float32)
    sn = snap[:k * F * na].reshape(k, F, na)
    smn = snapm[:k * F * m].reshape(k, F, m)
    for f in range(F):
        out[:, f * n + a_idx] = sn[:, f, :]
        out[:, f * n + m_idx] = smn[:, f, :]
    state = np.empty(F * n)
    for f in range(F):
        state[f * n + a_idx] = a[f][:na]
        state[f * n + m_idx] = M[f][:m]
    sel = np.flatnonzero(hitt[:na] >= 0)
    hits = (a_idx[sel], hitt[sel], m_idx[hitm[sel]])
    sub = None
    if m_sub > 0:
        sub = []
        subn_k = subn[:k * na].reshape(k, na)
        pos_ = subpos[:k * ms * dim * na].reshape(k, ms, dim, na)
        for ks in range(k):
            s = np.flatnonzero(subn_k[ks] == m_sub)
            sub.append((a_idx[s], pos_[ks][:, :, s]))
    return out, state, sub, hits
