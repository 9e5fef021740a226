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
