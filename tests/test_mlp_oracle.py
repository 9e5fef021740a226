"""Pin the MLP oracle (CPU only): finite differences, BF16 rounding against
torch's own conversion, Adam's first step, feature stencil indexing."""
import numpy as np
import torch

from oracle import mlp_np as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(3).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)          # non-zero biases too
    p[M.NPARAM - 16 + 3:] = 0
    x = rng.normal(0, 1, (7, 32))
    y = rng.normal(0, 1, (7, 3))

    def loss(q):
        out, _ = M.forward(q, x, rounding=False)
        return M.mse(out, y)

    out, cache = M.forward(p, x, rounding=False)
    g = M.backward(cache, 2 * (out - y) / len(x), rounding=False)
    mask = M.unpack(np.ones(M.NPARAM))
    mask["W4"][3:] = 0          # padding rows of the output layer
    mask["b4"][3:] = 0
    live = np.flatnonzero(M.pack(mask))
    idx = rng.choice(live, 60, replace=False)
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = 1e-6
        fd = (loss(p + e) - loss(p - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.normal(0, 10, 5000).astype(np.float32),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5,
                                    1e-30, 3.4e38])])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(a), ref)


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g), rtol=1e-12)


def test_features_pick_the_nearest_node_stencil():
    n = 8
    rho = np.arange(n ** 3, dtype=np.float64).reshape(n, n, n)
    h, origin = 0.5, [1.0, -1.0, 0.0]
    # body 0.3 cells past node (3, 4, 5) -> nearest node (3, 4, 5)
    pos = [np.array([origin[0] + 3.3 * h]), np.array([origin[1] + 4.3 * h]),
           np.array([origin[2] + 4.8 * h])]
    f = M.features(pos, rho, origin, h, 2.0)
    assert f[0, 13] == rho[3, 4, 5] / 2.0           # centre of the stencil
    assert f[0, 0] == rho[2, 3, 4] / 2.0
    assert f[0, 26] == rho[4, 5, 6] / 2.0
    np.testing.assert_allclose(f[0, 27:30], [0.6, 0.6, -0.4], atol=1e-12)
    assert f[0, 30] == 0.0 and M.features(pos, rho, origin, h, 2.0, 1)[0, 30] == 1.0


def test_scaled_activation_rounding():
    """act_scale only moves the BF16 rounding points: a power of two commutes
    with the rounding (identical result), and the inference kernel's scale
    -log2(e) stays as close to the f64 network as the unscaled rounding."""
    rng = np.random.default_rng(21)
    p = M.init_params(3)
    d = M.unpack(p.copy())
    for k in ("b1", "b2", "b3"):
        d[k][:] = rng.normal(0, 0.3, d[k].shape)
    p = M.pack(d)
    x = M.bf16(rng.normal(0, 1, (500, M.NIN)).astype(np.float32))
    base, _ = M.forward(p, x)
    two, _ = M.forward(p, x, act_scale=2.0)
    assert np.array_equal(base, two)
    scaled, _ = M.forward(p, x, act_scale=M.INFER_ACT_SCALE)
    assert not np.array_equal(base, scaled)
    exact, _ = M.forward(M.bf16(p), x, rounding=False)
    e_base = np.linalg.norm(base - exact)
    e_scaled = np.linalg.norm(scaled - exact)
    assert 0.5 * e_base < e_scaled < 1.5 * e_base
