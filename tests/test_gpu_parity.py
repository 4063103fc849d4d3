"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: bit-exact distances for f32 and u32 (the fixpoint of d[v] <= d[u] (+) w
is unique for w >= 0, so every schedule must land on the same bits as the
oracle's restatement of cpu_bellman_ford, sssp.hpp:3-71), plus the
shortest-path certificate (all edges satisfied, tight parent tree).
"""
import numpy as np
import pytest

import sssp_amd as S
from conftest import bits, small_graphs

SCHEDS = ["plain", "heuristic", "learned", "push", "all", "pull"]


def _check_solution(g, oracle, src, dst, w, V, source, d, sched):
    ref = oracle.dijkstra(src, dst, w, V, source)
    assert np.array_equal(bits(d), bits(ref)), f"{sched}: distances differ from the oracle"
    ok, counts, par = g.certify(source)
    assert ok, f"{sched}: GPU certificate failed: {counts}"
    rc, oc = oracle.certify(src, dst, w, V, source, d, par)
    assert rc == 0, f"{sched}: oracle rejects the GPU parents/distances: {oc}"


@pytest.mark.gpu
def test_device_is_gfx950(gpu):
    info = S.device_info(0)
    print(info)
    assert info["arch"].startswith("gfx950")


@pytest.mark.gpu
@pytest.mark.parametrize("name,V,src,dst,w", small_graphs(), ids=[g[0] for g in small_graphs()])
def test_small_graphs_every_schedule(gpu, oracle, name, V, src, dst, w):
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for sched in SCHEDS:
            d, st = g.solve(0, sched)
            _check_solution(g, oracle, src, dst, w, V, 0, d, sched)
        g.set_param("plain_kernel", 2)  # the plain pass over the blocked list too
        d, st = g.solve(0, "plain")
        _check_solution(g, oracle, src, dst, w, V, 0, d, "plain (blocked)")
    if name == "chain":
        assert np.array_equal(d, (np.arange(10) * np.float32(0.5)).astype(np.float32))
    if name == "disconnected":
        assert np.all(d[3:] == S.F32_INF)


@pytest.mark.gpu
def test_device_ingest_from_bin(gpu, oracle, tmp_path):
    """A graph streamed from the reference's binary file format equals the one
    built from the same arrays: layouts, and distances under every schedule."""
    src, dst, w = S.rmat(14, 16, seed=12, threads=8)
    V = 1 << 14
    p = tmp_path / "g.bin"
    S.save_bin(p, src, dst, w, V)
    with S.DeviceGraph.from_bin(p) as g1, S.DeviceGraph(src, dst, w, V) as g2:
        g1.set_param("plain_below_edges", 0)  # exercise the round engine
        g2.set_param("plain_below_edges", 0)  # exercise the round engine
        assert (g1.V, g1.E) == (V, len(src))
        L1, L2 = g1.layout(), g2.layout()
        for k in L1:
            assert np.array_equal(L1[k].view(np.uint8), L2[k].view(np.uint8)), k
        ref = oracle.dijkstra(src, dst, w, V, 3)
        for sched in SCHEDS:
            d, _ = g1.solve(3, sched)
            assert np.array_equal(bits(d), bits(ref)), sched
    su, tu, wu = S.er(1 << 12, 8, seed=5, weight_mode=1)
    pu = tmp_path / "u.bin"
    S.save_bin(pu, su, tu, wu, 1 << 12)
    with S.DeviceGraph.from_bin(pu, weight_dtype=np.uint32) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        d, _ = g.solve(0, "heuristic")
        assert np.array_equal(d, oracle.dijkstra(su, tu, wu, 1 << 12, 0))
    # several chunks (4 Mi records each) through the two pinned buffers
    s2, t2, w2 = S.rmat(20, 16, seed=13, threads=64)
    p2 = tmp_path / "big.bin"
    S.save_bin(p2, s2, t2, w2, 1 << 20)
    with S.DeviceGraph.from_bin(p2, csc=False) as g1, S.DeviceGraph(s2, t2, w2, 1 << 20, csc=False) as g2:
        g1.set_param("plain_below_edges", 0)  # exercise the round engine
        g2.set_param("plain_below_edges", 0)  # exercise the round engine
        a, _ = g1.solve(5, "plain")
        b, _ = g2.solve(5, "plain")
        assert np.array_equal(bits(a), bits(b))
        L1, L2 = g1.layout(csc=False), g2.layout(csc=False)
        assert np.array_equal(L1["el_dst"], L2["el_dst"])


@pytest.mark.gpu
@pytest.mark.parametrize("relabel", [True, False])
def test_layout_is_stable_sort(gpu, relabel):
    """The layouts are stable sorts of the edge list renamed into slots; slots
    rank the vertices by descending in- plus out-degree, then descending
    out-degree, then vertex id (the identity with relabel=False)."""
    src, dst, w = S.rmat(11, 16, seed=3, threads=4)
    V = 1 << 11
    with S.DeviceGraph(src, dst, w, V, relabel=relabel) as g:
        L = g.layout()
        order = g.vertex_order()
    if relabel:
        out = np.bincount(src, minlength=V)
        tot = out + np.bincount(dst, minlength=V)
        exp = np.empty(V, np.int32)
        exp[np.lexsort((np.arange(V), -out, -tot))] = np.arange(V, dtype=np.int32)
        assert np.array_equal(order, exp)
        assert not np.array_equal(order, np.arange(V))
    else:
        assert np.array_equal(order, np.arange(V))
    src, dst = order[src], order[dst]
    p = np.argsort(src, kind="stable")
    assert np.array_equal(L["el_src"], src[p])
    assert np.array_equal(L["el_dst"], dst[p])
    assert np.array_equal(bits(L["el_w"]), bits(w[p]))
    rp = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=V))])
    assert np.array_equal(L["row_ptr"], rp)
    q = np.argsort(dst, kind="stable")
    assert np.array_equal(L["csc_src"], src[q])
    assert np.array_equal(bits(L["csc_w"]), bits(w[q]))
    cp = np.concatenate([[0], np.cumsum(np.bincount(dst, minlength=V))])
    assert np.array_equal(L["col_ptr"], cp)


@pytest.mark.gpu
@pytest.mark.parametrize("relabel", [None, True], ids=["ids", "slots"])
@pytest.mark.parametrize("scale,ef,seed,threads", [(10, 16, 1, 64), (14, 16, 9, 64), (16, 16, 0, 1)])
def test_rmat_every_schedule_bit_exact(gpu, oracle, scale, ef, seed, threads, relabel):
    src, dst, w = S.rmat(scale, ef, seed=seed, threads=threads)
    V = 1 << scale
    sources = [0] + list(S.pick_sources(src, V, seed + 100, 3))
    with S.DeviceGraph(src, dst, w, V, relabel=relabel) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for s in sources:
            ref, _ = oracle.bf(src, dst, w, V, s)
            for sched in SCHEDS:
                d, st = g.solve(s, sched)
                assert np.array_equal(bits(d), bits(ref)), (s, sched, st)
                assert st["reached"] == int((d != S.F32_INF).sum()), (s, sched, st)
            ok, counts, par = g.certify(s)
            assert ok, counts
            rc, oc = oracle.certify(src, dst, w, V, s, d, par)
            assert rc == 0, oc


@pytest.mark.gpu
@pytest.mark.parametrize("logV,deg", [(12, 16), (15, 8)])
def test_er_u32_bit_exact(gpu, oracle, logV, deg):
    V = 1 << logV
    src, dst, w = S.er(V, deg, seed=3, weight_mode=1)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        ref, _ = oracle.bf(src, dst, w, V, 0)
        dj = oracle.dijkstra(src, dst, w, V, 0)
        assert np.array_equal(ref, dj)
        for sched in SCHEDS:
            d, st = g.solve(0, sched)
            assert d.dtype == np.uint32
            assert np.array_equal(d, ref), sched
            assert st["reached"] == int((d != np.uint32(0xFFFFFFFF)).sum()), (sched, st)
        ok, counts, par = g.certify(0)
        assert ok, counts
        rc, oc = oracle.certify(src, dst, w, V, 0, d, par)
        assert rc == 0, oc


@pytest.mark.gpu
def test_edge_order_does_not_change_distances(gpu):
    src, dst, w = S.rmat(13, 16, seed=4, threads=8)
    V = 1 << 13
    perm = np.random.default_rng(0).permutation(len(src))
    with S.DeviceGraph(src, dst, w, V) as g1, S.DeviceGraph(src[perm], dst[perm], w[perm], V) as g2:
        g1.set_param("plain_below_edges", 0)  # exercise the round engine
        g2.set_param("plain_below_edges", 0)  # exercise the round engine
        for sched in ("plain", "heuristic"):
            a, _ = g1.solve(5, sched)
            b, _ = g2.solve(5, sched)
            assert np.array_equal(bits(a), bits(b))


@pytest.mark.gpu
def test_repeat_solves_reuse_graph(gpu, oracle):
    """Stamps and ticks carry across solves on one graph; results must not."""
    src, dst, w = S.rmat(12, 8, seed=7, threads=8)
    V = 1 << 12
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for s in [0, 1, 2, 0, 3]:
            for sched in ("push", "all", "pull", "heuristic"):
                d, _ = g.solve(s, sched)
                ref = oracle.dijkstra(src, dst, w, V, s)
                assert np.array_equal(bits(d), bits(ref))


@pytest.mark.gpu
def test_verify_pass_and_trace(gpu):
    src, dst, w = S.rmat(12, 16, seed=2, threads=8)
    with S.DeviceGraph(src, dst, w, 1 << 12) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("trace_events", 1)
        d, st = g.solve(0, "heuristic")
        tr = g.trace()
        assert st["verify_passes"] >= 1
        assert tr["mode"][-1] == S.MODE_VERIFY
        assert len(tr["mode"]) == st["iterations"]
        assert st["t_solve_ms"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("tail_edges,group", [(1 << 14, 4), (1 << 14, 1), (64, 4), (64, 3)])
def test_tail_rounds_bit_exact(gpu, oracle, tail_edges, group):
    """Tail rounds (push rounds queued without a decision or host round trip
    between them, entry lists built in one pass) land on the oracle's
    fixpoint: f32 R-MAT (several sources, heuristic and learned), u32 ER and
    BFS levels. A bound of 64 out-edges makes the frontier cross it again and
    again (tail runs that stop and resume); group 1 runs under trace_events
    (one round per host check, each timed)."""
    src, dst, w = S.rmat(16, 16, seed=13, threads=8)
    V = 1 << 16
    sources = [0] + list(S.pick_sources(src, V, 131, 3))
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("tail_edges", tail_edges)
        g.set_param("tail_group", group)
        g.set_param("trace_events", 1 if group == 1 else 0)
        tails = 0
        for s in sources:
            for sched in ("heuristic", "learned"):
                print(f"tail rounds: rmat16 f32 source {s} {sched} tail_edges {tail_edges} "
                      f"group {group}", flush=True)
                d, st = g.solve(int(s), sched)
                assert np.array_equal(bits(d), bits(oracle.dijkstra(src, dst, w, V, int(s)))), \
                    (s, sched)
                tr = g.trace()
                assert len(tr["mode"]) == st["iterations"]
                assert tr["mode"][-1] == S.MODE_VERIFY
                k = tr["mode"] == S.MODE_TAIL
                tails += int(k.sum())
                if group == 1 and k.any():
                    assert (tr["frontier"][k] > 0).all() and (tr["ms"][k] > 0).all()
                    assert (tr["frontier_edges"][k] <= tail_edges).all()
        assert tails > 0, "no tail rounds ran"
        ok, counts, par = g.certify(int(sources[-1]))
        assert ok, counts
        print(f"tail rounds: rmat16 BFS tail_edges {tail_edges} group {group}", flush=True)
        vis, lev, st = g.bfs(int(sources[1]), "heuristic")
        rvis, rlev, _ = oracle.bfs(src, dst, V, int(sources[1]))
        assert np.array_equal(lev, rlev) and np.array_equal(vis.view(np.uint8), rvis)
    V = 1 << 15
    src, dst, w = S.er(V, 8, seed=17, weight_mode=1)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)
        g.set_param("tail_edges", tail_edges)
        g.set_param("tail_group", group)
        print(f"tail rounds: er15 u32 tail_edges {tail_edges} group {group}", flush=True)
        d, st = g.solve(5, "heuristic")
        assert np.array_equal(d, oracle.dijkstra(src, dst, w, V, 5))


@pytest.mark.gpu
def test_push_rounds_with_frontier_bitmap(gpu, oracle):
    """Push rounds that mark their lowered vertices in a bitmap and queue them
    in one pass (push_bits_edges) land on the oracle's fixpoint: f32 R-MAT with
    the push-only, heuristic and learned schedules (tail rounds on and off),
    u32 ER, BFS levels."""
    src, dst, w = S.rmat(16, 16, seed=31, threads=8)
    V = 1 << 16
    sources = [0] + list(S.pick_sources(src, V, 311, 3))
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("push_bits_edges", 1)    # every push round
        for tail in (0, 1 << 14):
            g.set_param("tail_edges", tail)
            for s in sources:
                for sched in ("push", "heuristic", "learned"):
                    print(f"bitmap push rounds: rmat16 source {s} {sched} tail {tail}", flush=True)
                    d, st = g.solve(int(s), sched)
                    assert np.array_equal(bits(d), bits(oracle.dijkstra(src, dst, w, V, int(s)))), \
                        (s, sched, tail)
        ok, counts, par = g.certify(int(sources[-1]))
        assert ok, counts
        print("bitmap push rounds: rmat16 BFS", flush=True)
        vis, lev, _ = g.bfs(int(sources[1]), "push")
        rvis, rlev, _ = oracle.bfs(src, dst, V, int(sources[1]))
        assert np.array_equal(lev, rlev)
    V = 1 << 15
    src, dst, w = S.er(V, 8, seed=37, weight_mode=1)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)
        g.set_param("push_bits_edges", 1)
        print("bitmap push rounds: er15 u32", flush=True)
        for sched in ("push", "heuristic"):
            d, st = g.solve(9, sched)
            assert np.array_equal(d, oracle.dijkstra(src, dst, w, V, 9)), sched


@pytest.mark.gpu
@pytest.mark.parametrize("scale,ef,seed,threads", [(12, 16, 3, 64), (16, 16, 0, 1), (18, 8, 11, 7)])
def test_device_rmat_generator_matches_host(gpu, scale, ef, seed, threads):
    s, d, w = S.rmat(scale, ef, seed=seed, threads=threads)
    g, (ds, dd, dw) = S.DeviceGraph.rmat(scale, ef, seed=seed, threads=threads, download_edges=True)
    try:
        assert np.array_equal(ds, s) and np.array_equal(dd, d)
        assert np.array_equal(bits(dw), bits(w))
        L = g.layout()
    finally:
        g.close()
    with S.DeviceGraph(s, d, w, 1 << scale) as h:
        h.set_param("plain_below_edges", 0)  # exercise the round engine
        M = h.layout()
    for k in ("row_ptr", "el_src", "el_dst", "col_ptr", "csc_src"):
        assert np.array_equal(L[k], M[k]), k


# ---------------------------------------------------------------- batches
@pytest.mark.gpu
@pytest.mark.parametrize("policy", [1, 2, 3])
def test_batch_sources_match_single_solves(gpu, oracle, policy):
    src, dst, w = S.rmat(13, 16, seed=21, threads=16)
    V = 1 << 13
    sources = S.pick_sources(src, V, 21, 48)
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        D, st = g.solve_batch(sources, policy=policy)
        assert D.shape == (48, V) and st["t_solve_ms"] > 0
        for i in (0, 7, 47):
            ref, _ = oracle.bf(src, dst, w, V, int(sources[i]))
            assert np.array_equal(bits(D[i]), bits(ref))
        for i in range(48):
            d, _ = g.solve(int(sources[i]), "push")
            assert np.array_equal(bits(D[i]), bits(d)), i
        # replay of the captured launch with other sources
        s2 = sources[::-1].copy()
        D2, _ = g.solve_batch(s2, policy=policy)
        assert np.array_equal(bits(D2[0]), bits(D[-1]))


@pytest.mark.gpu
def test_batch_captured_launch_replays(gpu, oracle):
    """Policy 3 runs each chunk of sources as captured hipGraph launches (change
    flags and pass base cleared by a kernel, init, runs of 32 passes reading
    their pass number from the device, transposed output). Three replays with
    other sources, eager solves between them: each bit-equal to the passes
    queued from the host and to policy 1 (a workgroup per source)."""
    src, dst, w = S.rmat(14, 16, seed=41, threads=16)
    V = 1 << 14
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g:
        for k in range(3):
            sources = S.pick_sources(src, V, 41 + k, 64)
            print(f"captured batch replay {k}: 64 sources, RMAT-14", flush=True)
            g.set_param("batch_capture", 1)
            D, _ = g.solve_batch(sources, policy=3)
            assert g.get_param("multi_passes") >= 1
            d, _ = g.solve(int(sources[k]), "heuristic")
            assert np.array_equal(bits(D[k]), bits(d))
            g.set_param("batch_capture", 0)
            D0, _ = g.solve_batch(sources, policy=3)
            assert np.array_equal(bits(D), bits(D0))
            D1, _ = g.solve_batch(sources, policy=1)
            assert np.array_equal(bits(D), bits(D1))
        ref, _ = oracle.bf(src, dst, w, V, int(sources[5]))
        assert np.array_equal(bits(D[5]), bits(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 2], ids=["edge_list", "blocked"])
def test_plain_solve_queued_and_traced(gpu, oracle, kernel):
    """The plain schedule queues its passes four at a time (change flags tested
    on the device); with trace_events each pass is launched alone between
    events. Same distances; the traced solve records one row per pass, the
    last one a pass that changed nothing; a pass cap below the fixpoint gives
    up and leaves the graph usable. Both plain pass kernels."""
    src, dst, w = S.rmat(15, 16, seed=42, threads=16)
    V = 1 << 15
    ref = oracle.dijkstra(src, dst, w, V, 7)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_kernel", kernel)
        assert g.get_param("plain_blocked") == (kernel == 2)
        d, st = g.solve(7, "plain")
        assert np.array_equal(bits(d), bits(ref))
        assert 2 <= st["iterations"] <= V + 1 and st["verify_passes"] == 1
        g.set_param("trace_events", 1)
        d2, st2 = g.solve(7, "plain")
        g.set_param("trace_events", 0)
        assert np.array_equal(bits(d2), bits(ref))
        tr = g.trace()
        assert len(tr["ms"]) == st2["iterations"] and np.all(tr["ms"] > 0)
        # a plain solve capped below its pass count gives up, then solves again
        g.set_param("max_rounds", 1)
        with pytest.raises(S.SSSPError) as e:
            g.solve(7, "plain")
        assert e.value.code == S.ERR_NOT_CONVERGED
        g.set_param("max_rounds", 100000)
        d3, _ = g.solve(7, "plain")
        assert np.array_equal(bits(d3), bits(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 2], ids=["edge_list", "blocked"])
def test_plain_kernels_bit_exact(gpu, oracle, kernel):
    """The plain schedule on either pass kernel -- the XCD-partitioned edge
    list, or the destination-blocked list (block minima in LDS, atomicMin
    write-backs for blocks split over several tiles) -- ends on the oracle's
    fixpoint: f32 with vertex ids and with degree-ordered slots, u32, BFS
    levels, and a graph whose hub block is split, with tiles down to 256
    edges."""
    src, dst, w = S.rmat(14, 16, seed=9, threads=64)
    V = 1 << 14
    sources = [0] + list(S.pick_sources(src, V, 109, 3))
    for relabel in (None, True):
        with S.DeviceGraph(src, dst, w, V, relabel=relabel) as g:
            g.set_param("plain_kernel", kernel)
            for s in sources:
                print(f"plain kernel {kernel}: rmat14 relabel {relabel} source {s}", flush=True)
                ref, _ = oracle.bf(src, dst, w, V, s)
                d, st = g.solve(s, "plain")
                assert np.array_equal(bits(d), bits(ref)), (s, st)
                assert st["verify_passes"] == 1
                ok, counts, _ = g.certify(s, parents=False)
                assert ok, counts
    V = 1 << 15
    src, dst, wu = S.er(V, 8, seed=3, weight_mode=1)
    with S.DeviceGraph(src, dst, wu, V) as g:
        g.set_param("plain_kernel", kernel)
        print(f"plain kernel {kernel}: er15 u32", flush=True)
        d, _ = g.solve(0, "plain")
        assert np.array_equal(d, oracle.dijkstra(src, dst, wu, V, 0))
    scale = 17
    src, dst, w = S.rmat(scale, 16, seed=44, threads=16)
    V = 1 << scale
    assert np.bincount(dst >> 14).max() > (1 << 19)  # the hub block is split
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g:
        g.set_param("plain_kernel", kernel)
        for cap in (256, 1 << 17):
            g.set_param("blk_tile_cap", cap)
            for s in (0, int(src[999])):
                print(f"plain kernel {kernel}: rmat17 split blocks, cap {cap}, source {s}", flush=True)
                d, _ = g.solve(s, "plain")
                assert np.array_equal(bits(d), bits(oracle.dijkstra(src, dst, w, V, s))), (cap, s)
                _, lev, _ = g.bfs(s, "plain")
                assert np.array_equal(lev, oracle.bfs(src, dst, V, s)[1]), (cap, s)


@pytest.mark.gpu
def test_blocked_list_wide_chunks(gpu, oracle):
    """The blocked list packs each source as a delta to the previous edge of
    its 256-edge chunk; a chunk with a gap of 2^18 or more keeps its sources
    whole. A sparse graph on 2^20 vertices (few in-edges per block, sources
    far apart) has such chunks next to packed ones; the plain blocked pass,
    the dense rounds and BFS stay exact on both."""
    V = 1 << 20
    rng = np.random.default_rng(5)
    n = 1500
    src = rng.integers(0, V, n).astype(np.int32)
    dst = rng.integers(0, V, n).astype(np.int32)
    hop = np.sort(rng.choice(V, 400, replace=False)).astype(np.int32)  # a long path
    src = np.concatenate([src, hop[:-1], np.full(300, 7, np.int32)])
    dst = np.concatenate([dst, hop[1:], rng.integers(0, 1 << 14, 300).astype(np.int32)])
    w = rng.random(len(src), dtype=np.float32)
    with S.DeviceGraph(src, dst, w, V, relabel=False) as g:
        nw = g.get_param("blk_wide_chunks")
        print(f"wide chunks: {nw} of {g.get_param('blk_tiles')} tiles", flush=True)
        assert nw > 0
        g.set_param("plain_kernel", 2)
        g.set_param("plain_below_edges", 0)
        g.set_param("dense_all_frac", 0.0)  # every all-edge round dense
        for s in (int(hop[0]), 7, int(src[3])):
            ref = oracle.dijkstra(src, dst, w, V, s)
            for sched in ("plain", "all", "heuristic", "learned"):
                d, _ = g.solve(s, sched)
                assert np.array_equal(bits(d), bits(ref)), (s, sched)
            _, lev, _ = g.bfs(s, "all")
            assert np.array_equal(lev, oracle.bfs(src, dst, V, s)[1]), s


@pytest.mark.gpu
def test_blocked_tiles_by_source_range(gpu, oracle):
    """Tiles cut by source range and dispatched range by range (blk_range_bits):
    the default for a graph without hubs from 2^22 vertices (8 ranges), off for
    R-MAT; any range size keeps the dense rounds and the plain blocked pass on
    the oracle's fixpoint, f32 and u32."""
    V = 1 << 20
    src, dst, w = S.er(V, 4, seed=6, weight_mode=0)
    wu = (np.arange(len(src)) % 255 + 1).astype(np.uint32)
    with S.DeviceGraph(src, dst, w, V) as g, S.DeviceGraph(src, dst, wu, V) as gu:
        assert g.get_param("blk_range_bits") == 0  # below 2^22 vertices
        ref = oracle.dijkstra(src, dst, w, V, 5)
        refu = oracle.dijkstra(src, dst, wu, V, 5)
        for h in (g, gu):
            h.set_param("plain_below_edges", 0)
            h.set_param("dense_all_frac", 0.0)  # every all-edge round dense
            h.set_param("plain_kernel", 2)
        for rb in (18, 19, 20, 0):
            print(f"source ranges of 2^{rb}", flush=True)
            for h in (g, gu):
                h.set_param("blk_range_bits", rb)
            assert g.get_param("blk_range_bits") == rb
            for sched in ("all", "plain", "heuristic"):
                d, _ = g.solve(5, sched)
                assert np.array_equal(bits(d), bits(ref)), (rb, sched)
            du, _ = gu.solve(5, "all")
            assert np.array_equal(du, refu), rb
        for bad in (5, 17, 32):
            with pytest.raises(S.SSSPError):
                g.set_param("blk_range_bits", bad)
    V = 1 << 22
    src, dst, w = S.er(V, 2, seed=7, weight_mode=0)
    with S.DeviceGraph(src, dst, w, V) as g:
        assert g.get_param("blk_range_bits") == 19  # 8 ranges of 2^19 vertices
        d, _ = g.solve(0, "plain")
        assert np.array_equal(bits(d), bits(oracle.dijkstra(src, dst, w, V, 0)))
    src, dst, w = S.rmat(16, 16, seed=1, threads=8)
    with S.DeviceGraph(src, dst, w, 1 << 16) as g:
        assert g.get_param("blk_range_bits") == 0


@pytest.mark.gpu
def test_plain_kernel_choice(gpu):
    """plain_kernel 0 picks the blocked pass once the blocked list has 64
    tiles or more (here forced by a small tile cap), else the edge-list
    pass; 1 and 2 force either; other values are rejected."""
    src, dst, w = S.rmat(14, 16, seed=9, threads=64)
    with S.DeviceGraph(src, dst, w, 1 << 14) as g:
        assert g.get_param("plain_kernel") == 0
        assert g.get_param("blk_tiles") < 64 and g.get_param("plain_blocked") == 0
        g.set_param("blk_tile_cap", 2048)
        assert g.get_param("blk_tiles") >= 64 and g.get_param("plain_blocked") == 1
        g.set_param("plain_kernel", 1)
        assert g.get_param("plain_blocked") == 0
        for bad in (-1, 3):
            with pytest.raises(S.SSSPError):
                g.set_param("plain_kernel", bad)


@pytest.mark.gpu
def test_blocked_tiles_any_cap_and_order(gpu, oracle):
    """The dense rounds' tile table rebuilt with other caps (edges per tile,
    down to 256: most blocks split over several workgroups) and dispatch
    orders: every dense round still ends on the oracle's fixpoint."""
    src, dst, w = S.rmat(14, 16, seed=43, threads=16)
    V = 1 << 14
    ref = oracle.dijkstra(src, dst, w, V, 3)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)
        g.set_param("dense_all_frac", 0.0)  # every all-edge round dense
        for cap, order in ((256, 0), (1024, 1), (4096, 2), (1 << 20, 3), (1 << 17, 0)):
            print(f"blocked tiles: cap {cap}, order {order}", flush=True)
            g.set_param("blk_tile_cap", cap)
            g.set_param("blk_tile_order", order)
            assert g.get_param("blk_tile_cap") == cap and g.get_param("blk_tile_order") == order
            for sched in ("all", "heuristic"):
                d, _ = g.solve(3, sched)
                assert np.array_equal(bits(d), bits(ref)), (cap, order, sched)
        for name, bad in (("blk_tile_cap", 1002), ("blk_tile_cap", 128), ("blk_tile_order", 4)):
            with pytest.raises(S.SSSPError):
                g.set_param(name, bad)


@pytest.mark.gpu
def test_batch_u32(gpu, oracle):
    src, dst, w = S.er(1 << 12, 8, seed=4, weight_mode=1)
    V = 1 << 12
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        D, _ = g.solve_batch(np.array([0, 5, 9], np.int32), policy=1)
        for i, s in enumerate((0, 5, 9)):
            assert np.array_equal(D[i], oracle.dijkstra(src, dst, w, V, s))


@pytest.mark.gpu
def test_batch_vertex_major_chunks_and_u32(gpu, oracle):
    """Policy 3 (vertex-major all-edge passes) in chunks of 256 sources: 300
    sources on an R-MAT graph with split blocks (a hub block with more than
    4096 in-edges), repeated sources, and integer weights."""
    src, dst, w = S.rmat(12, 16, seed=23, threads=8)
    V = 1 << 12
    assert np.bincount(dst >> 4).max() > 4096  # at least one split (16-vertex) block
    sources = np.concatenate([S.pick_sources(src, V, 23, 290), np.zeros(10, np.int32)])
    with S.DeviceGraph(src, dst, w, V) as g:
        D, st = g.solve_batch(sources, policy=3)
        assert D.shape == (300, V) and st["t_solve_ms"] > 0
        D1, _ = g.solve_batch(sources, policy=1)
        assert np.array_equal(bits(D), bits(D1))
        for i in (0, 255, 256, 299):
            assert np.array_equal(bits(D[i]), bits(oracle.dijkstra(src, dst, w, V, int(sources[i])))), i
    su, tu, wu = S.er(1 << 12, 8, seed=4, weight_mode=1)
    with S.DeviceGraph(su, tu, wu, 1 << 12) as g:
        D, _ = g.solve_batch(np.array([0, 5, 9], np.int32), policy=3)
        for i, s in enumerate((0, 5, 9)):
            assert np.array_equal(D[i], oracle.dijkstra(su, tu, wu, 1 << 12, s))


def _pack_case(n, seed):
    rng = np.random.default_rng(seed)
    scales = rng.integers(4, 11, n).astype(np.int32)
    s, d, w, offs = S.rmat_many(scales, 8, seed0=seed)
    vcounts = (1 << scales).astype(np.int32)
    srcs = np.array([int(s[offs[k]]) for k in range(n)], np.int32)  # a vertex with out-edges
    return scales, s, d, w, offs, vcounts, srcs


@pytest.mark.gpu
@pytest.mark.parametrize("policy", [1, 2])
def test_pack_matches_per_graph_oracle(gpu, oracle, policy):
    scales, s, d, w, offs, vc, srcs = _pack_case(40, 3)
    p = S.GraphPack(s, d, w, offs, vc)
    try:
        D, st = p.solve(srcs, policy=policy)
    finally:
        p.close()
    voff = np.concatenate([[0], np.cumsum(vc)])
    for k in range(40):
        e0, e1 = offs[k], offs[k + 1]
        ref, _ = oracle.bf(s[e0:e1], d[e0:e1], w[e0:e1], int(vc[k]), int(srcs[k]))
        assert np.array_equal(bits(D[voff[k]:voff[k + 1]]), bits(ref)), k


@pytest.mark.gpu
def test_autotune_batch_and_pack(gpu, tmp_path):
    src, dst, w = S.rmat(12, 8, seed=2, threads=8)
    V = 1 << 12
    sources = S.pick_sources(src, V, 2, 16)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        ref, _ = g.solve_batch(sources, policy=2)
        best = g.autotune_batch(sources, tmp_path / "tune.txt")
        assert best > 0
        text = (tmp_path / "tune.txt").read_text()
        assert text.count("best") == 1 and "workgroup_per_task" in text and "merged_frontier" in text
        assert "vertex_major_all_edge" in text
        D, _ = g.solve_batch(sources, policy=0)
        assert np.array_equal(bits(D), bits(ref))
    # untuned: 32 or more sources on one graph run policy 3, fewer policy 2
    many = S.pick_sources(src, V, 3, 40)
    with S.DeviceGraph(src, dst, w, V) as g:
        D, st = g.solve_batch(many, policy=0)
        assert st["iterations"] == 3
        D2, _ = g.solve_batch(many, policy=1)
        assert np.array_equal(bits(D), bits(D2))
        _, st = g.solve_batch(many[:8], policy=0)
        assert st["iterations"] == 2
    scales, s, d, w2, offs, vc, srcs = _pack_case(24, 5)
    p = S.GraphPack(s, d, w2, offs, vc)
    try:
        a, _ = p.solve(srcs, policy=2)
        assert p.autotune(srcs, tmp_path / "pack.txt") > 0
        b, _ = p.solve(srcs, policy=0)
        assert np.array_equal(bits(a), bits(b))
    finally:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("delta,hub", [(0.01, 0), (0.1, 0), (0.5, 0), (3.0, 0), (0.005, 64),
                                       (0.02, 16)])
def test_far_pile_keeps_the_fixpoint(gpu, oracle, delta, hub):
    """Push rounds defer vertices (d >= min + delta, out-degree >= hub) to the
    far pile; mixed schedules flush it before all-edge and pull rounds."""
    src, dst, w = S.rmat(14, 16, seed=31, threads=16)
    V = 1 << 14
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("delta", delta)
        g.set_param("hub_degree", hub)
        for s in (0, int(src[123])):
            ref = oracle.dijkstra(src, dst, w, V, s)
            for sched in ("push", "heuristic", "learned", "all", "pull"):
                d, st = g.solve(s, sched)
                assert np.array_equal(bits(d), bits(ref)), (delta, hub, s, sched)
                assert st["reached"] == int((d != S.F32_INF).sum())
        ok, counts, _ = g.certify(s)
        assert ok, counts
    su, tu, wu = S.er(1 << 13, 8, seed=9, weight_mode=1)
    with S.DeviceGraph(su, tu, wu, 1 << 13) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("delta", delta * 100)
        g.set_param("hub_degree", hub)
        d, _ = g.solve(0, "push")
        assert np.array_equal(d, oracle.dijkstra(su, tu, wu, 1 << 13, 0))


@pytest.mark.gpu
def test_far_pile_after_failed_solve_and_collection(gpu, oracle):
    """A solve stopped by max_rounds leaves vertices piled and stamps set; the
    next solve and a timing collection must start clean."""
    src, dst, w = S.rmat(13, 16, seed=5, threads=8)
    V = 1 << 13
    ref = oracle.dijkstra(src, dst, w, V, 3)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        g.set_param("delta", 0.01)
        g.set_param("hub_degree", 32)
        g.set_param("max_rounds", 4)
        with pytest.raises(S.SSSPError):
            g.solve(3, "push")
        g.set_param("max_rounds", 100000)
        d, _ = g.solve(3, "push")
        assert np.array_equal(bits(d), bits(ref))
        feats, times = g.collect_timings(3)
        assert len(feats) > 0 and np.all(np.isfinite(times[:, :8]))
        print(f"collection: {len(feats)} rows, {int(g.get_param('collect_rounds'))} rounds run",
              flush=True)
        d, _ = g.solve(3, "heuristic")
        assert np.array_equal(bits(d), bits(ref))
        # a collection past its round budget gives up from the host between
        # rounds (every kernel has ended); the graph stays usable
        g.set_param("collect_budget", 40)
        with pytest.raises(S.SSSPError) as e:
            g.collect_timings(3)
        assert e.value.code == S.ERR_NOT_CONVERGED and "budget" in str(e.value)
        assert g.get_param("collect_rounds") == 41
        d, _ = g.solve(3, "push")
        assert np.array_equal(bits(d), bits(ref))
        g.set_param("collect_budget", 1 << 22)


@pytest.mark.gpu
def test_round_limits_are_bounded(gpu):
    src, dst, w = S.rmat(8, 4, seed=1, threads=1)
    with S.DeviceGraph(src, dst, w, 256) as g:
        for name, bad in (("max_rounds", 0), ("max_rounds", 2.0 ** 28 + 1), ("max_rounds", 2.0 ** 31),
                          ("collect_budget", 0)):
            with pytest.raises(S.SSSPError) as e:
                g.set_param(name, bad)
            assert e.value.code == S.ERR_ARG
        g.set_param("max_rounds", 2 ** 28)
        d, _ = g.solve(0, "push")
        assert g.get_param("max_rounds") == 2 ** 28


# ---- BFS on the same engine (BFS::gpu_bfs, algorithm/bfs/bfs.hpp:213-253) -----

@pytest.mark.gpu
@pytest.mark.parametrize("name,V,src,dst,w", small_graphs(), ids=[g[0] for g in small_graphs()])
def test_bfs_small_graphs_every_schedule(gpu, oracle, name, V, src, dst, w):
    vis_ref, lev_ref, n = oracle.bfs(src, dst, V, 0)
    with S.DeviceGraph(src, dst, w, V) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for sched in SCHEDS:
            vis, lev, st = g.bfs(0, sched)
            assert np.array_equal(vis.view(np.uint8), vis_ref), sched
            assert np.array_equal(lev, lev_ref), sched
            assert st["reached"] == n
        ok, counts, _ = g.certify(0)
        assert ok, counts


@pytest.mark.gpu
def test_bfs_rmat_and_er_every_schedule(gpu, oracle):
    src, dst, w = S.rmat(14, 16, seed=17, threads=16)
    V = 1 << 14
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        for s in [0] + list(S.pick_sources(src, V, 5, 3)):
            vis_ref, lev_ref, _ = oracle.bfs(src, dst, V, int(s))
            for sched in SCHEDS:
                vis, lev, _ = g.bfs(int(s), sched)
                assert np.array_equal(lev, lev_ref), (int(s), sched)
                assert np.array_equal(vis.view(np.uint8), vis_ref)
            ok, counts, par = g.certify(int(s))
            assert ok, counts
            # the tight tree of a BFS: every visited non-source vertex's parent
            # is one level up
            m = (par >= 0)
            assert np.all(lev[m] == lev[par[m]] + 1)
            # BFS levels are not distances: the SSSP result is gone
            with pytest.raises(S.SSSPError):
                g.distances()
        d, _ = g.solve(3, "heuristic")
        assert np.array_equal(bits(d), bits(oracle.dijkstra(src, dst, w, V, 3)))
        ok, counts, _ = g.certify(3)
        assert ok, counts
    su, tu, wu = S.er(1 << 13, 8, seed=11, weight_mode=1)
    with S.DeviceGraph(su, tu, wu, 1 << 13) as g:
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        vis_ref, lev_ref, _ = oracle.bfs(su, tu, 1 << 13, 0)
        for sched in ("plain", "heuristic", "push", "pull"):
            vis, lev, _ = g.bfs(0, sched)
            assert np.array_equal(lev, lev_ref), sched


@pytest.mark.gpu
def test_small_graphs_hand_the_rule_and_network_to_plain_passes(gpu, oracle):
    """Below plain_below_edges a heuristic or learned solve (and BFS) runs the
    plain pass loop -- same fixpoint, no rounds; 0 switches this off."""
    src, dst, w = S.rmat(13, 16, seed=8, threads=8)
    V = 1 << 13
    ref = oracle.dijkstra(src, dst, w, V, 0)
    with S.DeviceGraph(src, dst, w, V) as g:
        assert g.get_param("plain_below_edges") == 2 ** 26
        for sched in ("heuristic", "learned"):
            d, st = g.solve(0, sched)
            assert np.array_equal(bits(d), bits(ref)), sched
            assert st["n_push"] == 0 and st["n_pull"] == 0 and st["mlp_rounds"] == 0
            assert st["n_all"] == st["iterations"] > 0
        vis, lev, st = g.bfs(0, "heuristic")
        assert np.array_equal(lev, oracle.bfs(src, dst, V, 0)[1]) and st["n_push"] == 0
        g.set_param("plain_below_edges", 0)
        d, st = g.solve(0, "heuristic")
        assert np.array_equal(bits(d), bits(ref)) and st["n_push"] > 0


@pytest.mark.gpu
def test_dense_threshold_default_follows_the_graph(gpu, oracle):
    """A graph without hub blocks (uniform random: every block of 2^14
    destinations near the mean in-edge count) runs dense rounds from 0.4 E of
    frontier edges, a skewed one (R-MAT) from 0.15 E; a set value stays."""
    V = 1 << 16
    src, dst, w = S.er(V, 16, seed=3, weight_mode=1)
    ref = oracle.dijkstra(src, dst, w, V, 0)
    with S.DeviceGraph(src, dst, w, V) as g:
        assert g.get_param("dense_all_frac") == 0.4
        g.set_param("plain_below_edges", 0)  # exercise the round engine
        d, _ = g.solve(0, "heuristic")
        assert np.array_equal(bits(d), bits(ref))
        g.set_param("dense_all_frac", 0.2)
        assert g.get_param("dense_all_frac") == 0.2
    src, dst, w = S.rmat(18, 8, seed=3, threads=8)
    with S.DeviceGraph(src, dst, w, 1 << 18) as g:
        assert g.get_param("dense_all_frac") == 0.15


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("frac,unreached", [(0.0, 2.0), (1e9, 0.0), (1e9, 2.0), (0.15, 0.5)])
def test_dense_all_edge_rounds_keep_the_fixpoint(gpu, oracle, frac, unreached, kernel):
    """All-edge rounds run dense when the frontier is large or most of the
    graph is unreached: over the destination-blocked edge list (kernel 0: LDS
    minima, the frontier written by the same kernel) or as the plain pass
    followed by a compaction against a snapshot (kernel 1). Forced on for
    every all-edge round, on only by the unreached rule, and off (the
    fixpoint check stays dense), the distances, levels and frontiers stay
    exact."""
    src, dst, w = S.rmat(14, 16, seed=41, threads=16)
    V = 1 << 14
    su, tu, wu = S.er(1 << 13, 8, seed=13, weight_mode=1)
    with S.DeviceGraph(src, dst, w, V) as g, S.DeviceGraph(su, tu, wu, 1 << 13) as gu:
        for h in (g, gu):
            h.set_param("plain_below_edges", 0)  # exercise the round engine
            h.set_param("dense_all_frac", frac)
            h.set_param("dense_unreached", unreached)
            h.set_param("dense_kernel", kernel)
            assert h.get_param("dense_all_frac") == frac
        for s in (0, int(src[77])):
            ref = oracle.dijkstra(src, dst, w, V, s)
            for sched in ("all", "heuristic", "learned", "push"):
                d, st = g.solve(s, sched)
                assert np.array_equal(bits(d), bits(ref)), (frac, unreached, s, sched)
                assert st["reached"] == int((d != S.F32_INF).sum())
            ok, counts, _ = g.certify(s)
            assert ok, counts
            _, lev, _ = g.bfs(s, "all")
            assert np.array_equal(lev, oracle.bfs(src, dst, V, s)[1])
        for sched in ("all", "heuristic"):
            d, _ = gu.solve(0, sched)
            assert np.array_equal(d, oracle.dijkstra(su, tu, wu, 1 << 13, 0)), sched
        with pytest.raises(S.SSSPError):
            g.set_param("dense_unreached", -1)
        with pytest.raises(S.SSSPError):
            g.set_param("dense_kernel", 2)


@pytest.mark.gpu
def test_dense_rounds_with_split_blocks(gpu, oracle):
    """A destination block with more in-edges than one tile holds is split
    over several tiles, which write back with atomicMin and stamp with an
    exchange; the R-MAT hub block here is split (checked on the host) and
    every dense round stays exact, for f32, u32 and BFS levels."""
    scale = 17
    src, dst, w = S.rmat(scale, 16, seed=44, threads=16)
    V = 1 << scale
    per_block = np.bincount(dst >> 14)
    assert per_block.max() > (1 << 19), per_block.max()  # at least one split block
    wu = (np.arange(len(src)) % 255 + 1).astype(np.uint32)
    with S.DeviceGraph(src, dst, w, V, relabel=True) as g, S.DeviceGraph(src, dst, wu, V) as gu:
        for h in (g, gu):
            h.set_param("plain_below_edges", 0)
            h.set_param("dense_all_frac", 0.0)  # every all-edge round dense
        for s in (0, int(src[999])):
            ref = oracle.dijkstra(src, dst, w, V, s)
            for sched in ("all", "heuristic", "learned"):
                d, st = g.solve(s, sched)
                assert np.array_equal(bits(d), bits(ref)), (s, sched)
            ok, counts, _ = g.certify(s)
            assert ok, counts
            du, _ = gu.solve(s, "all")
            assert np.array_equal(du, oracle.dijkstra(src, dst, wu, V, s)), s
            _, lev, _ = g.bfs(s, "all")
            assert np.array_equal(lev, oracle.bfs(src, dst, V, s)[1]), s


@pytest.mark.gpu
def test_python_mirror_cache_follows_in_place_edits(gpu, oracle):
    """SingleSourceShortestPaths.gpu_bellman_ford caches the device graph; an
    edit of the graph's arrays in place is seen by the next call."""
    g = S.GraphGenerationAPI.R_MAT_parallel(1 << 11, 8, 57, 19, 19, 5, 8, True, 4)
    op = S.SingleSourceShortestPaths()
    d = np.empty(g.vertices_count, np.float32)
    op.gpu_bellman_ford(g, d, 0)
    assert np.array_equal(bits(d), bits(oracle.dijkstra(g.src_ids, g.dst_ids, g.weights,
                                                          g.vertices_count, 0)))
    g.weights[g.src_ids == 0] = 0.0  # in place
    op.gpu_bellman_ford(g, d, 0)
    assert np.array_equal(bits(d), bits(oracle.dijkstra(g.src_ids, g.dst_ids, g.weights,
                                                          g.vertices_count, 0)))
