"""The scheduler network's GPU kernels against the CPU oracle (oracle/mlp_ref.py),
and what the trained network does inside the solver."""
import os

import numpy as np
import pytest

import mlp_ref as M
import sssp_amd as S
from conftest import ROOT, bits

WEIGHTS = os.path.join(ROOT, "parallel-task1_amd", "models", "scheduler_mlp.f32")


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, M.IN)).astype(np.float32)
    ym = rng.integers(0, M.MODES, n).astype(np.int32)
    yc = rng.integers(0, M.CHUNKS, n).astype(np.int32)
    p = rng.normal(0, 0.3, M.NPARAM).astype(np.float32)
    return p, x, ym, yc


@pytest.mark.gpu
def test_forward_matches_oracle(gpu):
    p, x, _, _ = _case(37, 0)
    m = S.MLP(0)
    m.set_params(p)
    assert np.array_equal(m.params(), p)
    out = m.forward(x)
    ref, _ = M.forward(p.astype(np.float64), x.astype(np.float64))
    # fp32 dot products of length <= 64 against float64: relative 1e-5
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
def test_gradient_matches_oracle(gpu):
    p, x, ym, yc = _case(64, 1)
    m = S.MLP(0)
    m.set_params(p)
    loss, g = m.loss_grad(x, ym, yc)
    rl, rg = M.loss_grad(p.astype(np.float64), x.astype(np.float64), ym, yc)
    assert abs(loss - rl) <= 1e-5 * max(1.0, abs(rl))
    scale = np.abs(rg).max()
    assert np.allclose(g, rg, rtol=1e-4, atol=1e-5 * scale)


@pytest.mark.gpu
def test_adam_step_matches_oracle(gpu):
    p, x, ym, yc = _case(32, 2)
    m = S.MLP(0)
    m.set_params(p)
    m.train(x, ym, yc, epochs=1, batch=32, lr=1e-2, seed=0)  # one full-batch step
    _, rg = M.loss_grad(p.astype(np.float64), x.astype(np.float64), ym, yc)
    ref, _, _ = M.adam_step(p.astype(np.float64), rg, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, 1e-2)
    assert np.allclose(m.params(), ref, rtol=1e-4, atol=1e-6)


@pytest.mark.gpu
def test_training_lowers_loss_on_a_learnable_rule(gpu):
    rng = np.random.default_rng(3)
    x = rng.random((512, M.IN)).astype(np.float32)
    ym = (x[:, 1] > 0.5).astype(np.int32) + (x[:, 2] > 0.7).astype(np.int32)
    yc = (x[:, 0] * 4).astype(np.int32).clip(0, 3)
    m = S.MLP(0, seed=1)
    l0, _ = m.loss_grad(x, ym, yc)
    losses = m.train(x, ym, yc, epochs=150, batch=64, lr=5e-3, seed=1)
    assert losses[-1] < 0.5 * l0
    acc = (m.forward(x)[:, :3].argmax(1) == ym).mean()
    assert acc > 0.85


@pytest.mark.gpu
def test_any_network_keeps_the_fixpoint(gpu, oracle):
    """Whatever the network picks, distances are the Bellman-Ford fixpoint:
    random and constant-output networks give bit-identical results."""
    src, dst, w = S.rmat(13, 16, seed=12, threads=8)
    V = 1 << 13
    ref = oracle.dijkstra(src, dst, w, V, 0)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for seed in range(3):
            m = S.MLP(0, seed=seed)
            g.set_mlp(m)
            d, st = g.solve(0, "learned")
            assert st["mlp_rounds"] > 0
            assert np.array_equal(bits(d), bits(ref))
        for mode in range(3):  # network that always answers `mode`
            p = np.zeros(S.MLP_NPARAM, np.float32)
            p[-S.MLP_OUT + mode] = 5.0
            m = S.MLP(0)
            m.set_params(p)
            g.set_mlp(m)
            d, st = g.solve(0, "learned")
            assert np.array_equal(bits(d), bits(ref)), mode
        g.set_mlp(None)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_trained_network_against_heuristic(gpu):
    """What the trained network gains over the hand-written rule on an R-MAT
    graph no training run saw (scale 24, edge factor 16, seed 99; training
    used seeds 7-12 and 57-60 and uniform graphs): 8 sources, the two
    schedules interleaved, best of three timings each. The learned schedule
    must not be slower than the rule, and its distances are identical."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights (run parallel-task1_amd/train_scheduler.py)")
    g = S.DeviceGraph.rmat(24, 16, seed=99, threads=64)
    try:
        deg = g.out_degrees()
        srcs = S.pick_sources(np.nonzero(deg)[0].astype(np.int32), g.V, 99, 8)
        m = S.MLP(0)
        m.set_params(np.fromfile(WEIGHTS, np.float32))
        g.set_mlp(m)
        th, tl = [], []
        for s in srcs:
            dh, _ = g.solve(int(s), "heuristic")
            dl, sl = g.solve(int(s), "learned")
            assert np.array_equal(bits(dh), bits(dl))
            assert sl["mlp_rounds"] > 0
            bh = bl = float("inf")
            for _ in range(3):
                _, sh = g.solve(int(s), "heuristic", download=False)
                _, sl = g.solve(int(s), "learned", download=False)
                bh = min(bh, sh["t_solve_ms"])
                bl = min(bl, sl["t_solve_ms"])
            th.append(bh)
            tl.append(bl)
        gain = sum(th) / sum(tl)
        print(f"learned vs heuristic on held-out RMAT-24: {gain:.3f}x (heuristic {sum(th):.2f} ms, "
              f"learned {sum(tl):.2f} ms over {len(srcs)} sources)")
        assert gain >= 1.0, f"learned schedule {gain:.3f}x the rule"
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_trained_bfs_network_against_heuristic(gpu):
    """The same for the BFS network, on R-MAT scale 24 seed 3 (training used
    seed 7 and uniform graphs): 4 sources, interleaved, best of three; the
    levels are identical and the learned BFS is not slower than the rule."""
    bfs_weights = os.path.join(os.path.dirname(WEIGHTS), "scheduler_bfs_mlp.f32")
    if not os.path.exists(bfs_weights):
        pytest.skip("no trained BFS weights")
    g = S.DeviceGraph.rmat(24, 16, seed=3, threads=64)
    try:
        srcs = S.pick_sources(np.nonzero(g.out_degrees())[0].astype(np.int32), g.V, 3, 4)
        m = S.MLP(0)
        m.set_params(np.fromfile(bfs_weights, np.float32))
        g.set_bfs_mlp(m)
        th, tl = [], []
        for s in srcs:
            _, lh, _ = g.bfs(int(s), "heuristic")
            _, ll, sl = g.bfs(int(s), "learned")
            assert np.array_equal(lh, ll)
            assert sl["mlp_rounds"] > 0
            bh = bl = float("inf")
            for _ in range(3):
                bh = min(bh, g.bfs(int(s), "heuristic", levels=False)[2]["t_solve_ms"])
                bl = min(bl, g.bfs(int(s), "learned", levels=False)[2]["t_solve_ms"])
            th.append(bh)
            tl.append(bl)
        gain = sum(th) / sum(tl)
        print(f"learned vs heuristic BFS on held-out RMAT-24: {gain:.3f}x (heuristic {sum(th):.2f} ms, "
              f"learned {sum(tl):.2f} ms over {len(srcs)} sources)")
        assert gain >= 1.0, f"learned BFS {gain:.3f}x the rule"
    finally:
        g.close()


@pytest.mark.gpu
def test_bfs_network_slot(gpu, oracle):
    """BFS has its own network slot: without one a learned BFS runs the rule
    (the SSSP network is not used); with any network the levels stay the
    hop distances."""
    src, dst, w = S.rmat(13, 16, seed=21, threads=8)
    V = 1 << 13
    _, lev_ref, _ = oracle.bfs(src, dst, V, 0)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_mlp(S.MLP(0, seed=1))          # an SSSP network: ignored by BFS
        _, lev, st = g.bfs(0, "learned")
        assert np.array_equal(lev, lev_ref) and st["mlp_rounds"] == 0
        for seed in range(2):
            g.set_bfs_mlp(S.MLP(0, seed=seed + 5))
            _, lev, st = g.bfs(0, "learned")
            assert np.array_equal(lev, lev_ref) and st["mlp_rounds"] > 0
        # BFS timing collection (rollouts) leaves a certified BFS result
        g.set_param("collect_kind", 1)
        feats, times = g.collect_timings(0)
        assert len(feats) > 0 and np.all(times[:, :8] > 0)
        ok, counts, _ = g.certify(0)
        assert ok, counts
        g.set_mlp(None)
        g.set_bfs_mlp(None)
