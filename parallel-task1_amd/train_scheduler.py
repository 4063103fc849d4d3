"""Trains the learned round scheduler on timings the solver itself measures.

For directed R-MAT graphs (edge factors 16 and 32, a/b/c/d = 57/19/19/5,
generated on the device) and uniform random graphs (integer weights 1..255), several
sources each, sssp_collect_timings saves the solve
state before every round and tries each of the 12 choices (push, all-edge,
pull x four chunk widths): the round runs that choice and the rest of the
solve follows the hand-written rule; the device time of the whole remainder is
the choice's cost (a rollout). The solve then advances by the best choice.
Each round becomes one training row: the 5 round features -> the mode with the
lowest remainder time and, for that mode, the best chunk bucket. A network
trained on these labels is one step of policy improvement over the rule;
--rollout-weights makes the rollouts continue with a trained network instead,
the next step of policy iteration. It
(5 -> 64 -> 64 -> 3 + 4) is trained on the GPU by the library's own
forward/backward/Adam kernels, and the weights are written to
models/scheduler_mlp.f32 (raw float32, the library's parameter order) with a
JSON sidecar describing the run.

    python parallel-task1_amd/train_scheduler.py [--scales 19 20 21 22 23 24]

The committed network (models/scheduler_mlp.f32) was trained on the rows of
two collections over the default graphs -- one whose rollouts continued with
the previous network (--rollout-weights), one whose rollouts followed the
rule -- concatenated:
    python parallel-task1_amd/train_scheduler.py --rollout-weights prev.f32 --out A.f32
    python parallel-task1_amd/train_scheduler.py --out B.f32
    python parallel-task1_amd/train_scheduler.py --data A_data.npz B_data.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import sssp_amd as S  # noqa: E402

MODELS = os.path.join(HERE, "models")


KIND = "sssp"  # set from --kind before collecting


def _rollout_with(g, net):
    """Rollouts continue with `net` (one more policy-improvement step) when
    given, else with the hand-written rule; --kind bfs collects BFS rounds."""
    if KIND == "bfs":
        g.set_param("collect_kind", 1)
    if net is not None:
        if KIND == "bfs":
            g.set_bfs_mlp(net)
        else:
            g.set_mlp(net)
        g.set_param("rollout_policy", 1)


def collect(scales, sources_per_scale, edge_factor, seed0, net=None):
    X, T, meta = [], [], []
    for k, scale in enumerate(scales):
        seed = seed0 + k
        g = S.DeviceGraph.rmat(scale, edge_factor, seed=seed, threads=64)
        _rollout_with(g, net)
        deg = g.out_degrees()
        srcs = S.pick_sources(np.repeat(np.arange(g.V, dtype=np.int32), deg), g.V, seed,
                              sources_per_scale)
        for s in srcs:
            t0 = time.perf_counter()
            f, t = g.collect_timings(int(s))
            X.append(f)
            T.append(t)
            meta.append({"scale": scale, "edge_factor": edge_factor, "seed": seed,
                         "source": int(s), "rounds": len(f),
                         "seconds": round(time.perf_counter() - t0, 2)})
            print(f"scale {scale} source {int(s)}: {len(f)} rounds", flush=True)
        g.close()
    return np.concatenate(X), np.concatenate(T), meta


def collect_er(logvs, sources_per_graph, degree, seed0, net=None):
    """Uniform random graphs (random_uniform, integer weights 1..255)."""
    X, T, meta = [], [], []
    for k, logv in enumerate(logvs):
        seed = seed0 + k
        V = 1 << logv
        src, dst, w = S.er(V, degree, seed=seed, weight_mode=1)
        g = S.DeviceGraph(src, dst, w, V)
        _rollout_with(g, net)
        srcs = S.pick_sources(src, V, seed, sources_per_graph)
        del src, dst, w
        for s in srcs:
            t0 = time.perf_counter()
            f, t = g.collect_timings(int(s))
            X.append(f)
            T.append(t)
            meta.append({"er_logv": logv, "seed": seed, "source": int(s), "rounds": len(f),
                         "seconds": round(time.perf_counter() - t0, 2)})
            print(f"ER 2^{logv} source {int(s)}: {len(f)} rounds", flush=True)
        g.close()
    return np.concatenate(X), np.concatenate(T), meta


def labels(T):
    """Fastest mode (over its best chunk) and that mode's fastest chunk."""
    t = T.reshape(len(T), S.MLP_MODES, S.MLP_CHUNKS).astype(np.float64)
    t = np.where(t < 0, np.inf, t)  # unavailable mode
    best_per_mode = t.min(axis=2)
    ym = best_per_mode.argmin(axis=1).astype(np.int32)
    yc = t[np.arange(len(t)), ym].argmin(axis=1).astype(np.int32)
    return ym, yc, t


def stakes_weights(t, min_spread, max_copies):
    """Rows whose choices all end within min_spread of each other carry no
    usable label (timing noise decides) and are dropped (0 copies); the others
    are repeated in proportion to what the choice is worth in milliseconds
    (worst minus best remainder), so the loss weighs rounds by their stakes."""
    flat = t.reshape(len(t), -1)
    fin = np.where(np.isfinite(flat), flat, np.nan)
    best, worst = np.nanmin(fin, axis=1), np.nanmax(fin, axis=1)
    keep = worst > best * min_spread
    stake = worst - best
    med = np.median(stake[keep]) if keep.any() else 1.0
    copies = np.clip(np.ceil(stake / max(med, 1e-9)), 1, max_copies).astype(np.int64)
    return np.where(keep, copies, 0)


def regret(t, pm, pc):
    """Total time of the chosen (mode, chunk) over the total of the best."""
    chosen = t[np.arange(len(t)), pm, pc]
    return float(chosen.sum() / t.reshape(len(t), -1).min(axis=1).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scales", type=int, nargs="+", default=[19, 20, 21, 22, 23, 24])
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--ef32-scales", type=int, nargs="*", default=[19, 20, 21, 22])
    ap.add_argument("--er-logv", type=int, nargs="*", default=[20, 21, 22, 23])
    ap.add_argument("--er-degree", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--kind", choices=("sssp", "bfs"), default="sssp",
                    help="rounds of SSSP solves or of BFS (unit weights)")
    ap.add_argument("--out", help="default: models/scheduler_mlp.f32 (sssp), "
                                  "models/scheduler_bfs_mlp.f32 (bfs)")
    ap.add_argument("--data", nargs="+",
                    help="train on the X/T arrays of earlier runs' _data.npz files (concatenated)")
    ap.add_argument("--rollout-weights",
                    help="weights of a network the rollouts continue with (policy iteration); "
                         "default: the hand-written rule")
    ap.add_argument("--min-spread", type=float, default=1.02)
    ap.add_argument("--max-copies", type=int, default=8)
    args = ap.parse_args()
    global KIND
    KIND = args.kind
    if not args.out:
        args.out = os.path.join(MODELS, "scheduler_mlp.f32" if KIND == "sssp"
                                else "scheduler_bfs_mlp.f32")

    if args.data:
        Xs, Ts = [], []
        for path in args.data:
            with np.load(path) as z:
                Xs.append(z["X"])
                Ts.append(z["T"])
        X, T = np.concatenate(Xs), np.concatenate(Ts)
        meta = [{"data": os.path.basename(p)} for p in args.data]
    else:
        net = None
        if args.rollout_weights:
            net = S.MLP(0)
            net.set_params(np.fromfile(args.rollout_weights, np.float32))
        X, T, meta = collect(args.scales, args.sources, args.edge_factor, args.seed, net)
        if args.ef32_scales:
            Xr, Tr, mr = collect(args.ef32_scales, args.sources, 32, args.seed + 50, net)
            X, T, meta = np.concatenate([X, Xr]), np.concatenate([T, Tr]), meta + mr
        if args.er_logv:
            Xe, Te, me = collect_er(args.er_logv, args.sources, args.er_degree, args.seed + 100,
                                    net)
            X, T, meta = np.concatenate([X, Xe]), np.concatenate([T, Te]), meta + me
    ym, yc, t = labels(T)
    n = len(X)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(n)
    nval = max(1, n // 5)
    val, tr = perm[:nval], perm[nval:]
    copies = stakes_weights(t[tr], args.min_spread, args.max_copies)
    trw = np.repeat(tr, copies)

    mlp = S.MLP(0, seed=args.seed)
    losses = mlp.train(X[trw], ym[trw], yc[trw], epochs=args.epochs,
                       batch=min(args.batch, len(trw)), lr=args.lr, seed=args.seed)
    out = mlp.forward(X)
    pm = out[:, :S.MLP_MODES].argmax(axis=1)
    pc = out[:, S.MLP_MODES:].argmax(axis=1)
    # the heuristic's choice on the same rows, for comparison (fixed chunks)
    report = {
        "kind": args.kind,
        "rows": int(n), "train_rows": int(len(tr)), "val_rows": int(nval),
        "train_rows_kept": int((copies > 0).sum()), "train_rows_weighted": int(len(trw)),
        "min_spread": args.min_spread, "max_copies": args.max_copies,
        "rollout_policy": "network " + os.path.basename(args.rollout_weights)
        if args.rollout_weights else "rule",
        "scales": args.scales, "edge_factor": args.edge_factor, "sources_per_scale": args.sources,
        "ef32_scales": args.ef32_scales, "er_logv": args.er_logv, "er_degree": args.er_degree,
        "label_mode_counts": np.bincount(ym, minlength=S.MLP_MODES).tolist(),
        "final_train_loss": float(losses[-1]) if len(losses) else None,
        "val_mode_accuracy": float((pm[val] == ym[val]).mean()),
        "val_chunk_accuracy": float((pc[val] == yc[val]).mean()),
        "val_time_regret": regret(t[val], pm[val], pc[val]),
        "train_time_regret": regret(t[tr], pm[tr], pc[tr]),
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "collection": meta,
    }
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    mlp.params().astype(np.float32).tofile(args.out)
    np.savez(os.path.splitext(args.out)[0] + "_data.npz", X=X, T=T)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps({k: v for k, v in report.items() if k != "collection"}), flush=True)


if __name__ == "__main__":
    main()
