"""Pins the CPU oracle of the scheduler network before the GPU kernels are
checked against it: analytic backward vs central finite differences, forward
vs a per-row loop, Adam vs a hand-unrolled step."""
import numpy as np

import mlp_ref as M


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, M.IN))
    ym = rng.integers(0, M.MODES, n)
    yc = rng.integers(0, M.CHUNKS, n)
    p = rng.normal(0, 0.3, M.NPARAM)
    return p, x, ym, yc


def test_backward_matches_finite_differences():
    p, x, ym, yc = _data(17, 0)
    _, g = M.loss_grad(p, x, ym, yc)
    idx = np.random.default_rng(1).choice(M.NPARAM, 300, replace=False)
    # every block of the layout is sampled
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = np.unique(np.concatenate([idx, offs[:-1], offs[1:] - 1]))
    ng = M.numeric_grad(p, x, ym, yc, idx)
    assert np.allclose(g[idx], ng, rtol=1e-5, atol=1e-8)


def test_forward_matches_row_loop():
    p, x, _, _ = _data(5, 2)
    P = M.unpack(p)
    o, _ = M.forward(p, x)
    for r in range(len(x)):
        h1 = np.maximum(P["W1"] @ x[r] + P["b1"], 0)
        h2 = np.maximum(P["W2"] @ h1 + P["b2"], 0)
        assert np.allclose(o[r], P["W3"] @ h2 + P["b3"])


def test_adam_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = M.adam_step(p, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first step of Adam moves every coordinate by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
