"""CPU tests that pin the oracle before it is trusted (no GPU needed).

Pins: golden vectors written by the reference's own code (tests/golden/),
Dijkstra bit-equality, brute force over all simple paths on tiny graphs,
closed forms, edge-permutation invariance, and planted mistakes that the
certificate must reject.
"""
import ctypes as C
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

import sssp_amd as S
from conftest import bits, small_graphs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FMAX = np.float32(np.finfo(np.float32).max)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.fixture(scope="module")
def meta():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- generators
def test_lcg_rand_r_matches_libc():
    """The generators restate glibc rand_r; compare with the libc function."""
    libc = C.CDLL("libc.so.6")
    libc.rand_r.argtypes = [C.POINTER(C.c_uint)]
    libc.rand_r.restype = C.c_int
    # the library exposes the stream only through the generator: scale-1
    # R-MAT draws exactly 0 quadrant + 2 coin + 1 weight rand_r per edge
    src, dst, w = S.rmat(1, 1000, seed=7, threads=1)
    seed = C.c_uint(0)
    for e in range(1000):
        c1 = libc.rand_r(C.byref(seed))
        c2 = libc.rand_r(C.byref(seed))
        r = libc.rand_r(C.byref(seed))
        assert src[e] == (1 - (c1 % 2 == 0)) and dst[e] == (1 - (c2 % 2 == 0))
        assert w[e] == np.float32(np.float32(r) / np.float32(2147483648.0))


@pytest.mark.parametrize("V,ef", [(256, 4), (1024, 8)])
def test_rmat_matches_reference_vectors(V, ef):
    g = gold(f"rmat_v{V}_ef{ef}.npz")
    s, d, w = S.rmat(int(np.log2(V)), ef, seed=999, threads=1)  # one thread: seed*0 = 0
    assert np.array_equal(s, g["src"]) and np.array_equal(d, g["dst"])
    assert np.array_equal(bits(w), bits(g["w"]))


def test_rmat_c1_matches_reference_hashes(meta):
    m = meta["rmat_s16_ef16"]
    s, d, w = S.rmat(16, 16, seed=1, threads=1)
    assert sha(s) == m["sha256_src"] and sha(d) == m["sha256_dst"] and sha(w) == m["sha256_w"]
    smp = gold("rmat_s16_ef16_samples.npz")
    ei = smp["edge_idx"]
    assert np.array_equal(s[ei], smp["src"]) and np.array_equal(bits(w[ei]), bits(smp["w"]))


def test_rmat_virtual_threads_are_core_count_independent():
    a = S.rmat(12, 8, seed=5, threads=64)
    b = S.rmat(12, 8, seed=5, threads=64)
    for x, y in zip(a, b):
        assert np.array_equal(bits(x), bits(y))
    c = S.rmat(12, 8, seed=5, threads=3)
    assert not np.array_equal(a[0], c[0])


def test_rmat_statistics():
    s, d, w = S.rmat(16, 8, seed=2, threads=64)
    V = 1 << 16
    assert len(s) == V * 8 and s.min() >= 0 and s.max() < V and d.min() >= 0 and d.max() < V
    assert w.min() >= 0 and w.max() <= 1.0
    # top bit of the source id is 0 with probability a + b = 0.76 (first level)
    assert abs(np.mean(s < V // 2) - 0.76) < 0.01
    assert abs(np.mean(d < V // 2) - 0.76) < 0.01


def test_er_matches_reference(meta):
    g = gold("er_v64_d4_s3.npz")
    s, d, w = S.er(64, 4, seed=3, weight_mode=0)
    assert np.array_equal(s, g["src"]) and np.array_equal(d, g["dst"])
    assert np.array_equal(bits(w), bits(g["w"]))
    m = meta["er_v5000_d16_s3"]
    s, d, w = S.er(5000, 16, seed=3, weight_mode=0)
    assert sha(s) == m["sha256_src"] and sha(d) == m["sha256_dst"] and sha(w) == m["sha256_w"]
    # u32 mode consumes the same draws; only the weight mapping differs
    s2, d2, wu = S.er(5000, 16, seed=3, weight_mode=1)
    assert np.array_equal(s2, s) and wu.min() >= 1 and wu.max() <= 255


# ---------------------------------------------------------------- oracle BF
@pytest.mark.parametrize("V,ef", [(256, 4), (1024, 8)])
def test_oracle_bf_matches_reference_cpu_bellman_ford(oracle, V, ef):
    g = gold(f"rmat_v{V}_ef{ef}.npz")
    for k, src in enumerate(g["sources"]):
        d, _ = oracle.bf(g["src"], g["dst"], g["w"], V, int(src), threads=4)
        assert np.array_equal(bits(d), bits(g["dist"][k]))


def test_oracle_bf_c1_matches_reference(oracle, meta):
    s, d, w = S.rmat(16, 16, seed=1, threads=1)
    dist, it = oracle.bf(s, d, w, 1 << 16, 0)
    assert sha(dist) == meta["rmat_s16_ef16"]["sha256_dist_source0"]
    smp = gold("rmat_s16_ef16_samples.npz")
    assert np.array_equal(bits(dist[smp["vert_idx"]]), bits(smp["dist"]))
    assert int((dist < FMAX).sum()) == meta["rmat_s16_ef16"]["reached_source0"]
    dj = oracle.dijkstra(s, d, w, 1 << 16, 0)
    assert np.array_equal(bits(dist), bits(dj))


def test_oracle_against_compiled_reference(oracle, ref_lib):
    """Direct cross-check with the reference's cpu_bellman_ford when the
    oracle/_ref build is present."""
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    s, d, w = S.rmat(13, 16, seed=4, threads=8)
    V = 1 << 13
    for src in (0, 17, 4000):
        r = np.empty(V, np.float32)
        assert ref_lib.ref_cpu_bellman_ford_f32(P(s), P(d), P(w), V, len(s), src, 4, P(r)) == 0
        o, _ = oracle.bf(s, d, w, V, src)
        assert np.array_equal(bits(o), bits(r))


def _brute(V, src, dst, w, source, dtype):
    """Min over all simple paths of the left-to-right sum in `dtype`."""
    adj = {}
    for a, b, c in zip(src, dst, w):
        adj.setdefault(int(a), []).append((int(b), c))
    inf = {np.float32: FMAX, np.float64: np.finfo(np.float64).max}.get(dtype, np.uint32(0xFFFFFFFF))
    best = np.full(V, inf, dtype)
    best[source] = 0

    def dfs(u, dist, seen):
        for v, c in adj.get(u, []):
            if v in seen:
                continue
            if dtype == np.uint32:
                nd = min(int(dist) + int(c), 0xFFFFFFFF)
            else:
                nd = dtype(dist + c)  # one IEEE add in the weight's own precision
            nd = dtype(nd)
            if nd < best[v]:
                best[v] = nd
            dfs(v, nd, seen | {v})

    dfs(source, dtype(0), {source})
    return best


@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_brute_force_tiny(oracle, seed):
    rng = np.random.default_rng(seed)
    V = 7
    E = 16
    src = rng.integers(0, V, E).astype(np.int32)
    dst = rng.integers(0, V, E).astype(np.int32)
    wf = rng.random(E).astype(np.float32)
    wf[rng.integers(0, E, 2)] = 0  # zero weights
    wf[0] = np.float32(1e-9)       # sub-ulp against larger distances
    for source in range(3):
        d, _ = oracle.bf(src, dst, wf, V, source)
        assert np.array_equal(bits(d), bits(_brute(V, src, dst, wf, source, np.float32)))
    wu = rng.integers(1, 256, E).astype(np.uint32)
    for source in range(3):
        d, _ = oracle.bf(src, dst, wu, V, source)
        assert np.array_equal(d, _brute(V, src, dst, wu, source, np.uint32))


@pytest.mark.parametrize("seed", range(6))
def test_oracle_f64_equals_brute_force_tiny(oracle, seed):
    """Double weights: the minimum over all simple paths of the sum taken in
    double, with a zero weight and a weight far below an ulp of the sums (a
    loop run in float, or with a dropped term, gives a different answer)."""
    rng = np.random.default_rng(100 + seed)
    V, E = 7, 16
    src = rng.integers(0, V, E).astype(np.int32)
    dst = rng.integers(0, V, E).astype(np.int32)
    w = rng.random(E)
    w[rng.integers(0, E, 2)] = 0.0
    w[1] = 1e-300
    for source in range(3):
        d, _ = oracle.bf(src, dst, w, V, source)
        exp = _brute(V, src, dst, w, source, np.float64)
        assert np.array_equal(d.view(np.uint64), exp.view(np.uint64))
        # rounding the weights to float changes the answer: the test resolves
        # double precision
        if source == 0:
            df, _ = oracle.bf(src, dst, w.astype(np.float32), V, source)
            reach = df < FMAX
            assert not np.array_equal(df[reach].astype(np.float64), d[reach]) or reach.sum() == 1


@pytest.mark.parametrize("V,ef", [(256, 4), (1024, 8)])
def test_oracle_bf_f64_matches_reference_vectors(oracle, V, ef):
    """cpu_bellman_ford<int, double> as the reference computes it (compiled
    from its own sources into the golden vectors), bit for bit, DBL_MAX on
    unreached vertices."""
    g = gold(f"rmat_v{V}_ef{ef}_f64.npz")
    for k, src in enumerate(g["sources"]):
        d, _ = oracle.bf(g["src"], g["dst"], g["w"], V, int(src), threads=4)
        assert np.array_equal(d.view(np.uint64), g["dist"][k].view(np.uint64))
        assert (g["dist"][k] == np.finfo(np.float64).max).any()


def test_oracle_f64_against_compiled_reference(oracle, ref_lib):
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    s, d, _ = S.rmat(12, 16, seed=9, threads=8)
    w = np.random.default_rng(3).random(len(s))
    V = 1 << 12
    for src in (0, 33, 2000):
        r = np.empty(V, np.float64)
        assert ref_lib.ref_cpu_bellman_ford_f64(P(s), P(d), P(w), V, len(s), src, 4, P(r)) == 0
        o, _ = oracle.bf(s, d, w, V, src)
        assert np.array_equal(o.view(np.uint64), r.view(np.uint64))


def test_oracle_capped_loop(oracle):
    """The pass-capped loop behind the CPU baseline: with a cap at or above
    the passes the full solve needs it is the full solve (same distances,
    converged); below it, it stops after exactly `cap` passes, reports no
    convergence, and its distances bound the fixpoint from above, each finite
    one set through an in-edge whose source has only come down since
    (d[u] + w <= d[v] for some in-edge u -> v with d[u] finite)."""
    s, d, w = S.rmat(12, 8, seed=12, threads=8)
    V = 1 << 12
    full, passes = oracle.bf(s, d, w, V, 0, threads=1)
    assert passes >= 3
    for cap in (passes, passes + 5):
        got, it, conv = oracle.bf_capped(s, d, w, V, 0, cap, threads=1)
        assert conv == 1 and it == passes and np.array_equal(bits(got), bits(full))
    for cap in (1, passes - 1):
        got, it, conv = oracle.bf_capped(s, d, w, V, 0, cap, threads=1)
        assert conv == 0 and it == cap
        assert np.all(got >= full)
        fin = np.flatnonzero((got < FMAX) & (np.arange(V) != 0))
        below = np.zeros(V, bool)
        ok = got[s] < FMAX
        cand = (got[s[ok]] + w[ok]).astype(np.float32)
        below[d[ok][cand <= got[d[ok]]]] = True
        assert below[fin].all()
    # one pass from the source relaxes at least its out-edges in edge order
    got, _, _ = oracle.bf_capped(s, d, w, V, 0, 1, threads=1)
    assert np.all(got[d[s == 0]] <= w[s == 0])


@pytest.mark.parametrize("name,V,src,dst,w", small_graphs(), ids=[g[0] for g in small_graphs()])
def test_oracle_closed_forms(oracle, name, V, src, dst, w):
    d, _ = oracle.bf(src, dst, w, V, 0)
    if name == "chain":
        exp = np.zeros(V, np.float32)
        for i in range(1, V):
            exp[i] = np.float32(exp[i - 1] + np.float32(0.5))
        assert np.array_equal(d, exp)
    elif name == "star":
        assert np.array_equal(d[1:], w)
    elif name == "disconnected":
        assert np.array_equal(d, np.array([0, 1, 2, FMAX, FMAX, FMAX], np.float32))
    elif name == "loops_dups":
        # 0->1 keeps the lighter duplicate (0.3), 1->2 weighs 0, the 2->3
        # duplicates give 0.3 + 0.25; self loops change nothing
        f = np.float32
        assert np.array_equal(d, np.array([0, f(0.3), f(0.3) + f(0), f(0.3) + f(0.25)], np.float32))
    elif name == "empty":
        assert np.array_equal(d, np.zeros(1, np.float32))


def test_oracle_permutation_invariance(oracle):
    s, d, w = S.rmat(12, 16, seed=8, threads=4)
    p = np.random.default_rng(1).permutation(len(s))
    a, _ = oracle.bf(s, d, w, 1 << 12, 3)
    b, _ = oracle.bf(s[p], d[p], w[p], 1 << 12, 3)
    assert np.array_equal(bits(a), bits(b))


def test_u32_oracle_guards_the_reference_wraparound(oracle):
    """The reference's unsigned loop adds w to 0xFFFFFFFF unguarded
    (sssp.hpp:21,56); the oracle does not reproduce that, and equals Dijkstra."""
    g = gold("u32_wrap_v64.npz")
    o, _ = oracle.bf(g["src"], g["dst"], g["w"], 64, 0)
    dj = oracle.dijkstra(g["src"], g["dst"], g["w"], 64, 0)
    assert np.array_equal(o, dj)
    assert not np.array_equal(g["ref_dist"], dj)  # the documented defect


def test_u32_oracle_matches_dijkstra_random(oracle):
    s, d, w = S.er(1 << 12, 8, seed=5, weight_mode=1)
    a, _ = oracle.bf(s, d, w, 1 << 12, 0)
    assert np.array_equal(a, oracle.dijkstra(s, d, w, 1 << 12, 0))


# ---------------------------------------------------------------- certificate
def test_certificate_accepts_oracle_and_rejects_planted_mistakes(oracle):
    s, d, w = S.rmat(11, 8, seed=6, threads=4)
    V = 1 << 11
    dist, _ = oracle.bf(s, d, w, V, 0)
    rc, c = oracle.certify(s, d, w, V, 0, dist)
    assert rc == 0 and c[5] > 1
    reach = np.flatnonzero((dist < FMAX) & (np.arange(V) != 0))
    v = int(reach[len(reach) // 2])
    # one ulp too high: some in-edge is violated
    bad = dist.copy()
    bad[v] = np.nextafter(bad[v], np.float32(np.inf))
    assert oracle.certify(s, d, w, V, 0, bad)[0] != 0
    # one ulp too low: no tight parent
    bad = dist.copy()
    bad[v] = np.nextafter(bad[v], np.float32(0))
    assert oracle.certify(s, d, w, V, 0, bad)[0] != 0
    # a reachable vertex reported unreachable
    bad = dist.copy()
    bad[v] = FMAX
    assert oracle.certify(s, d, w, V, 0, bad)[0] != 0
    # source not at zero
    bad = dist.copy()
    bad[0] = np.float32(1e-3)
    assert oracle.certify(s, d, w, V, 0, bad)[0] != 0


def test_certificate_checks_parents(oracle):
    src = np.array([0, 1, 0, 2], np.int32)
    dst = np.array([1, 2, 2, 3], np.int32)
    w = np.array([1, 1, 2, 1], np.float32)
    dist, _ = oracle.bf(src, dst, w, 4, 0)
    good = np.array([-1, 0, 1, 2], np.int32)   # 2 via 1 (tight: 1+1 = 2)
    alt = np.array([-1, 0, 0, 2], np.int32)    # 2 via 0 (tight: 0+2 = 2)
    assert oracle.certify(src, dst, w, 4, 0, dist, good)[0] == 0
    assert oracle.certify(src, dst, w, 4, 0, dist, alt)[0] == 0
    bad = np.array([-1, 0, 3, 2], np.int32)    # cycle 2 <-> 3, not an edge 3->2
    assert oracle.certify(src, dst, w, 4, 0, dist, bad)[0] != 0
    bad = np.array([-1, 2, 1, 2], np.int32)    # 1 <- 2 is not an edge
    assert oracle.certify(src, dst, w, 4, 0, dist, bad)[0] != 0


# ---- BFS (BFS::cpu_sequential_bfs, algorithm/bfs/bfs.hpp:3-45) ----------------

@pytest.mark.parametrize("V,ef", [(256, 4), (1024, 8)])
def test_oracle_bfs_matches_reference_vectors(oracle, V, ef):
    g = gold(f"rmat_v{V}_ef{ef}.npz")
    vis = gold("bfs_ref.npz")[f"v{V}_ef{ef}"]
    for i, s in enumerate(g["sources"]):
        v, lev, n = oracle.bfs(g["src"], g["dst"], V, int(s))
        assert np.array_equal(v, vis[i]) and n == int(vis[i].sum())


def test_oracle_bfs_c1_matches_reference_hash(oracle, meta):
    s, d, _ = S.rmat(16, 16, seed=1, threads=1)
    v, lev, n = oracle.bfs(s, d, 1 << 16, 0)
    m = meta["rmat_s16_ef16"]
    assert sha(v) == m["bfs_sha256_visited_source0"]
    assert n == m["bfs_reached_source0"]


def test_oracle_bfs_against_compiled_reference(oracle, ref_lib):
    _p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    for seed in range(4):
        s, d, _ = S.er(300, 3, seed=seed + 40, weight_mode=0)
        for src in (0, 17, 299):
            ref = np.empty(300, np.uint8)
            assert ref_lib.ref_cpu_sequential_bfs(_p(s), _p(d), 300, len(s), src, _p(ref)) == 0
            v, _, _ = oracle.bfs(s, d, 300, src)
            assert np.array_equal(v, ref)


def test_oracle_bfs_levels_are_unit_weight_distances(oracle):
    """A FIFO BFS's discovery level is the hop distance: it equals Dijkstra
    with every weight 1 (an independent algorithm), and visited == finite."""
    for seed in range(3):
        s, d, _ = S.rmat(11, 8, seed=seed + 5, threads=4)
        V = 1 << 11
        one = np.ones(len(s), np.uint32)
        for src in (0, int(s[7])):
            v, lev, n = oracle.bfs(s, d, V, src)
            dj = oracle.dijkstra(s, d, one, V, src)
            assert np.array_equal(lev, dj)
            assert np.array_equal(v.astype(bool), dj != 0xFFFFFFFF) and n == int(v.sum())


@pytest.mark.parametrize("name,V,src,dst,w", small_graphs(), ids=[g[0] for g in small_graphs()])
def test_oracle_bfs_closed_forms(oracle, name, V, src, dst, w):
    v, lev, n = oracle.bfs(src, dst, V, 0)
    assert v[0] == 1 and lev[0] == 0
    if name == "chain":
        assert np.array_equal(lev, np.arange(V, dtype=np.uint32)) and n == V
    if name == "star":
        assert np.all(lev[1:] == 1) and n == V
    if name == "disconnected":
        assert v.tolist() == [1, 1, 1, 0, 0, 0] and lev[2] == 2 and lev[3] == 0xFFFFFFFF
    if name == "loops_dups":
        assert lev.tolist() == [0, 1, 2, 3]
    if name == "empty":
        assert n == 1
