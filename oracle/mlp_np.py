"""NumPy oracle of the learned correction MLP — TEST INFRASTRUCTURE ONLY.

The network has no reference counterpart, so this restatement is pinned by
mathematics only: its f64 backward pass is checked against central finite
differences (tests/test_mlp_oracle.py), and its BF16 mode rounds exactly
where the GPU kernels round (inputs, weights, activations, output
gradients and layer deltas to BF16, round to nearest even) so the GPU's
forward, gradients and Adam step can be compared to it.  The inference
kernel rounds the activations scaled by -log2(e) (`act_scale`); its result
is also checked against the pure-f64 network, within the error of the
BF16-rounded network itself.

Parameter block layout (float32, NPARAM values):
  W1[64][32] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[16][64] b4[16]
rows 3..15 of W4 and b4 are zero.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT, NOUT_USED = 32, 64, 16, 3
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k]).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], np.float32).ravel() for n, _ in SIZES])


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform fan-in initialisation, output rows 3..15 zero."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[1]) * (0.5 if name == "W4" else 1.0)
            d[name] = rng.uniform(-lim, lim, shape).astype(np.float32)
        else:
            d[name] = np.zeros(shape, np.float32)
    d["W4"][NOUT_USED:] = 0
    return pack(d)


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


# activation scale of the inference kernel: it keeps the BF16 rounding of
# -log2(e) silu(z) (its hidden layers carry t = -log2(e) z), the training
# kernels the rounding of silu(z) itself
INFER_ACT_SCALE = -1.4426950408889634


def forward(p, x, rounding=True, act_scale=1.0):
    """Returns (out [n,3], cache).  rounding=False is the pure-f64 network.

    With rounding, each hidden activation h = silu(z) enters the next layer
    as bf16(act_scale h) / act_scale (exact in f64 for the scale): 1 for the
    training kernels, INFER_ACT_SCALE for the inference kernel."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    P = unpack(np.asarray(p, np.float64) if not rounding else p)
    W = {k: (r(v) if k.startswith("W") else np.asarray(v, np.float64))
         for k, v in P.items()}
    h0 = r(x).astype(np.float64)
    zs, hs = [], [h0]
    h = h0
    for l in (1, 2, 3):
        z = h @ W[f"W{l}"].T.astype(np.float64) + W[f"b{l}"]
        if rounding:
            z = z.astype(np.float32).astype(np.float64)
        if rounding:
            h = bf16((act_scale * silu(z)).astype(np.float32)).astype(np.float64) / act_scale
        else:
            h = silu(z)
        zs.append(z)
        hs.append(h)
    out = h @ W["W4"].T.astype(np.float64) + W["b4"]
    return out[:, :NOUT_USED], (W, zs, hs)


def backward(cache, gout, rounding=True):
    """Parameter gradient (flat, NPARAM) for dL/dout = gout [n,3]."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W, zs, hs = cache
    n = gout.shape[0]
    g = np.zeros((n, NOUT))
    g[:, :NOUT_USED] = gout
    d = r(g.astype(np.float32) if rounding else g).astype(np.float64)
    grads = {"W4": d.T @ hs[3], "b4": d.sum(0)}
    for l in (3, 2, 1):
        back = d @ W[f"W{l + 1}"].astype(np.float64)
        dz = back * silu_grad(zs[l - 1])
        d = r(dz.astype(np.float32) if rounding else dz).astype(np.float64)
        grads[f"W{l}"] = d.T @ hs[l - 1]
        grads[f"b{l}"] = d.sum(0)
    grads["W4"][NOUT_USED:] = 0
    grads["b4"][NOUT_USED:] = 0
    return pack({k: grads[k] for k, _ in SIZES}).astype(np.float64)


def mse(out, y):
    return float(np.mean(np.sum((out - y) ** 2, axis=1)))


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def features(pos, rho, origin, h, mass_unit, n_mass=0):
    """27 stencil node masses around the nearest node + 2x offsets + the
    mass flag (1 for the first n_mass bodies) (n, 32)."""
    n = rho.shape[0]
    f = np.stack([(np.asarray(pos[a]) - origin[a]) / h for a in range(3)])
    c = np.floor(f + 0.5)
    off = 2.0 * (f - c)
    c = np.clip(c.astype(np.int64), 0, n - 1)
    out = np.zeros((f.shape[1], 32))
    q = 0
    for da in (-1, 0, 1):
        for db in (-1, 0, 1):
            for dc in (-1, 0, 1):
                i, j, k = c[0] + da, c[1] + db, c[2] + dc
                ok = (i >= 0) & (i < n) & (j >= 0) & (j < n) & (k >= 0) & (k < n)
                v = np.zeros(f.shape[1])
                v[ok] = rho[i[ok], j[ok], k[ok]]
                out[:, q] = v / mass_unit
                q += 1
    out[:, 27:30] = off.T
    out[:n_mass, 30] = 1.0
    return out
