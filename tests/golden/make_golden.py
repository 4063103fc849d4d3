"""Writes the golden vectors in tests/golden/ from the reference's own code.

Runs in the build container only: it loads oracle/_ref/libref_sssp.so, which
`make -C oracle ref` compiles from the reference sources in /root/reference
(GraphGenerationAPI::R_MAT_parallel and random_uniform,
SingleSourceShortestPaths::cpu_bellman_ford for float, double and unsigned
weights, BFS::cpu_sequential_bfs,
GraphStorageAPI edge-list I/O).
The outputs are data only (inputs and expected outputs).

    python tests/golden/make_golden.py
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_sssp.so"))
    p, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.ref_rmat_parallel.restype = i64
    L.ref_rmat_parallel.argtypes = [i32, i32, i32, i32, i32, i32, p, p, p]
    L.ref_random_uniform_f32.restype = i64
    L.ref_random_uniform_f32.argtypes = [i32, i32, C.c_uint32, p, p, p]
    L.ref_cpu_bellman_ford_f32.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_cpu_bellman_ford_u32.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_cpu_bellman_ford_f64.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_save_bin_f32.argtypes = [C.c_char_p, p, p, p, i32, i64]
    L.ref_cpu_sequential_bfs.argtypes = [p, p, i32, i64, i32, p]

    def rmat(V, ef):
        E = V * ef
        s, d, w = np.empty(E, np.int32), np.empty(E, np.int32), np.empty(E, np.float32)
        assert L.ref_rmat_parallel(V, ef, 57, 19, 19, 5, P(s), P(d), P(w)) == E
        return s, d, w

    def bf(s, d, w, V, src, threads=1):
        out = np.empty(V, np.float32)
        assert L.ref_cpu_bellman_ford_f32(P(s), P(d), P(w), V, len(s), src, threads, P(out)) == 0
        return out

    meta = {}
    rng = np.random.default_rng(12345)

    # 1. tiny R-MAT graphs: full vectors (edges + distances from 3 sources)
    for V, ef in [(256, 4), (1024, 8)]:
        s, d, w = rmat(V, ef)
        srcs = np.array([0, 1, int(s[len(s) // 2])], np.int32)
        dists = np.stack([bf(s, d, w, V, int(x)) for x in srcs])
        np.savez(os.path.join(HERE, f"rmat_v{V}_ef{ef}.npz"), src=s, dst=d, w=w, sources=srcs,
                 dist=dists)

    # 1b. double weights (the reference's other instantiated type): the tiny
    #     R-MAT topologies with full-precision uniform doubles, one zero weight
    #     and one weight far below an ulp of the distances it is added to
    for V, ef in [(256, 4), (1024, 8)]:
        s, d, _ = rmat(V, ef)
        w64 = np.random.default_rng(V).random(len(s))
        w64[3] = 0.0
        w64[5] = 1e-300
        srcs = np.array([0, 1, int(s[len(s) // 3])], np.int32)
        dists = []
        for x in srcs:
            out = np.empty(V, np.float64)
            assert L.ref_cpu_bellman_ford_f64(P(s), P(d), P(w64), V, len(s), int(x), 1, P(out)) == 0
            dists.append(out)
        np.savez(os.path.join(HERE, f"rmat_v{V}_ef{ef}_f64.npz"), src=s, dst=d, w=w64,
                 sources=srcs, dist=np.stack(dists))

    # 2. the CPU-runnable configuration: R-MAT scale 16, edge factor 16,
    #    one generator thread (seed 0), source 0 -- hashes plus samples
    V, ef = 1 << 16, 16
    s, d, w = rmat(V, ef)
    s16, d16 = s, d
    dist = bf(s, d, w, V, 0, threads=8)
    ei = np.sort(rng.choice(len(s), 4096, replace=False)).astype(np.int64)
    vi = np.sort(rng.choice(V, 4096, replace=False)).astype(np.int64)
    np.savez(os.path.join(HERE, "rmat_s16_ef16_samples.npz"), edge_idx=ei, src=s[ei], dst=d[ei],
             w=w[ei], vert_idx=vi, dist=dist[vi])
    meta["rmat_s16_ef16"] = {"sha256_src": sha(s), "sha256_dst": sha(d), "sha256_w": sha(w),
                             "sha256_dist_source0": sha(dist),
                             "reached_source0": int((dist < np.finfo(np.float32).max).sum())}

    # 3. random_uniform with srand(seed): full vectors for a small graph,
    #    hashes for a larger one
    for V, deg, seed, full in [(64, 4, 3, True), (5000, 16, 3, False)]:
        E = V * deg
        es, ed, ew = np.empty(E, np.int32), np.empty(E, np.int32), np.empty(E, np.float32)
        assert L.ref_random_uniform_f32(V, deg, seed, P(es), P(ed), P(ew)) == E
        if full:
            np.savez(os.path.join(HERE, f"er_v{V}_d{deg}_s{seed}.npz"), src=es, dst=ed, w=ew)
        else:
            meta[f"er_v{V}_d{deg}_s{seed}"] = {"sha256_src": sha(es), "sha256_dst": sha(ed),
                                               "sha256_w": sha(ew)}

    # 4. the reference's unsigned-weight loop as written: it adds w to
    #    0xFFFFFFFF without a guard (sssp.hpp:21,56); recorded to document
    #    the defect the integer oracle does not reproduce
    V = 64
    E = V * 4
    es, ed, ew = np.empty(E, np.int32), np.empty(E, np.int32), np.empty(E, np.float32)
    L.ref_random_uniform_f32(V, 4, 9, P(es), P(ed), P(ew))
    wu = (1 + (np.arange(E) * 7919) % 255).astype(np.uint32)
    du = np.empty(V, np.uint32)
    L.ref_cpu_bellman_ford_u32(P(es), P(ed), P(wu), V, E, 0, 1, P(du))
    np.savez(os.path.join(HERE, "u32_wrap_v64.npz"), src=es, dst=ed, w=wu, ref_dist=du)

    # 5. binary edge-list file written by GraphStorageAPI
    bs = np.array([0, 1, 2], np.int32)
    bd = np.array([1, 2, 0], np.int32)
    bw = np.array([0.5, 0.25, 1.0], np.float32)
    path = os.path.join(HERE, "three_edges.bin")
    assert L.ref_save_bin_f32(path.encode(), P(bs), P(bd), P(bw), 3, 3) == 0

    # 6. BFS::cpu_sequential_bfs reachability (bfs.hpp:3-45) on the tiny R-MAT
    #    graphs of section 1 (same sources) and on the scale-16 graph
    def bfs(s, d, V, src):
        out = np.empty(V, np.uint8)
        assert L.ref_cpu_sequential_bfs(P(s), P(d), V, len(s), src, P(out)) == 0
        return out

    vis = {}
    for V, ef in [(256, 4), (1024, 8)]:
        s1, d1, _ = rmat(V, ef)
        srcs = np.array([0, 1, int(s1[len(s1) // 2])], np.int32)
        vis[f"v{V}_ef{ef}"] = np.stack([bfs(s1, d1, V, int(x)) for x in srcs])
    np.savez(os.path.join(HERE, "bfs_ref.npz"), **vis)
    V = 1 << 16
    vb = bfs(s16, d16, V, 0)
    meta["rmat_s16_ef16"]["bfs_sha256_visited_source0"] = sha(vb)
    meta["rmat_s16_ef16"]["bfs_reached_source0"] = int(vb.sum())

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
