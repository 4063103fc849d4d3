"""Benchmark: SSSP GTEPS + HBM GB/s, learned schedule vs plain GPU Bellman-Ford.

Workload (defaults): directed R-MAT scale 25, edge factor 16 (V = 33,554,432,
E = 536,870,912), a/b/c/d = 57/19/19/5, f32 weights float(rand_r)/RAND_MAX,
base seed 6 with 64 virtual generator threads, 8 sources drawn among vertices
with out-degree > 0 (std::mt19937_64(6)). Generated on the device
(bit-identical to the host generator). One step = one single-source solve,
cycling over the 8 sources (the default 16 steps solve each twice); value =
steps * E / wall time of the timed steps
(GTEPS, the reference's E / t_solve convention, sssp.hpp:274-276).

Every source is solved with the learned, heuristic and plain schedules outside
the timed region; the three distance arrays must be bit-identical and the
learned one must pass the GPU certificate. The plain schedule is plain
Bellman-Ford with the library's default pass kernel for the graph (at this
size the destination-blocked pass, k_blk_plain); besides it the line reports
plain_edge_list_gteps (the same passes over the XCD-partitioned edge list,
k_el_pass) and all_dense_gteps (the round engine with every round an
all-edge round run dense).

roofline: the learned schedule's dominant kernel, the dense all-edge round
(k_blk_pass), timed live with HIP events around each of its launches in one
learned solve; plain_pass_roofline: the plain pass over one plain solve;
edge_list_pass_roofline: k_el_pass over one plain solve on the edge list.
HBM traffic per launch (traffic): before this process touches the GPU it runs
itself twice under `rocprofv3 --pmc` (FETCH_SIZE, then WRITE_SIZE) in
--pmc-child mode, which copies 1 GiB device to device and runs one plain solve
with each pass kernel and one learned solve; each counter is scaled by its
value for the copy. --traffic off skips this (traffic = null).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--traffic auto|off]
"""
import argparse
import csv
import ctypes as C
import glob
import json
import os
import platform
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-task1_amd"))

import sssp_amd as S  # noqa: E402

METRIC = "SSSP GTEPS + HBM GB/s, learned-schedule vs plain GPU, RMAT-25, certified exact"
HBM_PEAK_GBS = 8000.0
WEIGHTS = os.path.join(ROOT, "parallel-task1_amd", "models", "scheduler_mlp.f32")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_schedule(g, sources, sched, steps, warmup):
    for i in range(warmup):
        g.solve(int(sources[i % len(sources)]), sched, download=False)
    t_dev = []
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        _, st = g.solve(int(sources[i % len(sources)]), sched, download=False)
        t_dev.append(st["t_solve_ms"])
        stats.append(st)
    wall = time.perf_counter() - t0
    return wall, t_dev, stats


def cpu_baseline(src, dst, w, V, source, plain_passes, max_passes):
    """The oracle's OpenMP restatement of cpu_bellman_ford on the same graph,
    on omp_get_max_threads() threads as sample/sssp.cpp:82 sets them (OpenMP
    honours OMP_NUM_THREADS and the process's CPU affinity), stopped after
    `max_passes` passes; GTEPS = E / (time per pass * passes the full solve
    needs)."""
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    p = C.c_void_p
    f = lib.orc_bf_f32_capped
    f.argtypes = [p, p, p, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32, p, p]
    f.restype = C.c_int
    lib.orc_max_threads.restype = C.c_int
    threads = int(lib.orc_max_threads())
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    d = np.empty(V, np.float32)
    conv = C.c_int32()
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    t0 = time.perf_counter()
    passes = f(P(src), P(dst), P(w), V, len(src), int(source), threads, max_passes, P(d),
               C.byref(conv))
    t = time.perf_counter() - t0
    per_pass = t / max(passes, 1)
    head = (f"oracle cpu_bellman_ford restatement (OpenMP, schedule(guided), {threads} threads, "
            f"{cpu_model()}) on the same RMAT-{int(np.log2(V))} graph, source {int(source)}: ")
    if conv.value:
        # the whole solve ran: its in-place passes converge in fewer passes
        # than the GPU's plain passes
        full_passes = passes
        sample = (head + f"a full solve to its fixpoint, {passes} passes in {t:.1f} s "
                  f"({per_pass:.2f} s/pass)")
    else:
        full_passes = max(plain_passes, passes)
        sample = (head + f"{passes} passes timed ({t:.1f} s, {per_pass:.2f} s/pass), GTEPS "
                  f"extrapolated to {full_passes} passes (the plain GPU run's count)")
    return {
        "value": len(src) / (per_pass * full_passes) / 1e9,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": sample,
        "omp_threads": threads,
        "nproc": os.cpu_count(),
        "affinity_cpus": affinity,
        "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
        "cpu_model": cpu_model(),
    }


COPY_BYTES = 1 << 30
EDGE_LIST_PASS = "k_el_pass"
BLOCKED_PASS = "k_blk_plain"
DENSE_KERNEL = "k_blk_pass"


def pass_kernel(g):
    """Name of the kernel the plain schedule runs on graph g."""
    return BLOCKED_PASS if g.get_param("plain_blocked") else EDGE_LIST_PASS


def pmc_child(args):
    """Profiled child: a 1 GiB device-to-device copy (calibration), then one
    plain solve with the default pass kernel, one with the edge-list pass and
    one learned solve of the bench graph from the first bench source; prints
    the pass counts, the default pass kernel and the learned solve's dense
    rounds (k_blk_pass launches)."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(COPY_BYTES)) != 0:
            raise RuntimeError("hipMalloc failed")
    if (hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) != 0 or hip.hipDeviceSynchronize() != 0
            or hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) != 0
            or hip.hipDeviceSynchronize() != 0):
        raise RuntimeError("calibration copy failed")
    hip.hipFree(a)
    hip.hipFree(b)
    g = S.DeviceGraph.rmat(args.scale, args.edge_factor, seed=args.seed, threads=64)
    source = bench_sources(g, args)[0]
    kern = pass_kernel(g)
    _, st = g.solve(int(source), "plain", download=False)
    passes = {kern: int(st["iterations"])}
    if kern != EDGE_LIST_PASS:
        g.set_param("plain_kernel", 1)
        _, st = g.solve(int(source), "plain", download=False)
        g.set_param("plain_kernel", 0)
        passes[EDGE_LIST_PASS] = int(st["iterations"])
    if os.path.exists(WEIGHTS):
        mlp = S.MLP(0)
        mlp.set_params(np.fromfile(WEIGHTS, np.float32))
        g.set_mlp(mlp)
    g.solve(int(source), "learned", download=False)
    dense = int(np.isin(g.trace()["mode"], (S.MODE_VERIFY, S.MODE_DENSE)).sum())
    print(json.dumps({"passes": passes, "pass_kernel": kern, "dense_rounds": dense,
                      "source": int(source)}), flush=True)


def bench_sources(g, args):
    """The bench sources: pick_sources depends only on which vertices have
    out-edges, read here from the device CSR."""
    deg = g.out_degrees()
    return S.pick_sources(np.nonzero(deg)[0].astype(np.int32), g.V, args.seed, args.sources)


def _pmc_run(counter, args, workdir):
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = os.path.join(workdir, counter)
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child", "--scale", str(args.scale),
           "--edge-factor", str(args.edge_factor), "--seed", str(args.seed), "--sources",
           str(args.sources)]
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, start_new_session=True)
    try:
        so, se = p.communicate(timeout=240)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError(f"rocprofv3 --pmc {counter} timed out")
    if p.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {p.returncode}: {se[-400:]}")
    child = None
    for line in so.splitlines():
        if line.startswith("{"):
            child = json.loads(line)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if child is None or not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no output")
    passes, dense = child["passes"], child["dense_rounds"]
    copy_kib, dense_kib = [], []
    pass_kib = {k: [] for k in passes}
    with open(files[0]) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name, v = r["Kernel_Name"], float(r["Counter_Value"])
            if "copyBuffer" in name:
                copy_kib.append(v)
            elif DENSE_KERNEL in name:
                dense_kib.append(v)
            else:
                for k in pass_kib:
                    if k in name:
                        pass_kib[k].append(v)
    if (not copy_kib or len(dense_kib) != dense
            or any(len(pass_kib[k]) < n for k, n in passes.items())):
        raise RuntimeError(f"rocprofv3 --pmc {counter}: dispatches missing")
    # the copy's counter in KiB against the bytes it moves; a plain solve's
    # first `passes` pass dispatches are its real passes (later queued ones
    # return at once); every dense-kernel dispatch is a dense round
    factor = COPY_BYTES / (max(copy_kib) * 1024.0)
    per_pass = {k: float(np.mean(v[:passes[k]])) * 1024.0 * factor for k, v in pass_kib.items()}
    per_dense = float(np.mean(dense_kib)) * 1024.0 * factor if dense else None
    return per_pass, per_dense, factor, passes, dense


def copy_bandwidth(reps=5):
    """Achievable HBM bandwidth on this GPU: a 1 GiB device-to-device copy
    (reads and writes 1 GiB), best of `reps`, timed with HIP events."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    ok = (hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
          and hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
          and hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0)
    best = None
    if ok and hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0:
        for _ in range(reps):
            ms = C.c_float()
            if (hip.hipEventRecord(e0, None) != 0
                    or hip.hipMemcpyAsync(b, a, C.c_size_t(COPY_BYTES), 3, None) != 0
                    or hip.hipEventRecord(e1, None) != 0 or hip.hipEventSynchronize(e1) != 0
                    or hip.hipEventElapsedTime(C.byref(ms), e0, e1) != 0):
                best = None
                break
            best = ms.value if best is None else min(best, ms.value)
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    for e in (e0, e1):
        if e.value:
            hip.hipEventDestroy(e)
    return None if not best else 2 * COPY_BYTES / (best * 1e-3) / 1e9


def measure_traffic(args):
    """HBM bytes per launch (FETCH + WRITE, copy-calibrated) of each plain
    pass kernel and of the dense round kernel."""
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        fetch, dfetch, ffac, passes, dense = _pmc_run("FETCH_SIZE", args, td)
        write, dwrite, wfac, _, _ = _pmc_run("WRITE_SIZE", args, td)
    out = {"passes": {k: {"bytes_per_launch": fetch[k] + write[k], "fetch_bytes": fetch[k],
                          "write_bytes": write[k], "launches": passes[k]} for k in passes},
           "dense_round": None, "fetch_factor": ffac, "write_factor": wfac}
    if dfetch is not None and dwrite is not None:
        out["dense_round"] = {"bytes_per_launch": dfetch + dwrite, "fetch_bytes": dfetch,
                              "write_bytes": dwrite, "launches": dense}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--seed", type=int, default=6)
    ap.add_argument("--cpu-passes", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--traffic", choices=("auto", "off"), default="auto")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): no multi-GPU "
              f"path is built; the independent sources of a batch would be the unit to shard "
              f"across GPUs, each holding a replica of the graph", file=sys.stderr)
        sys.exit(2)

    # host side only so far: the profiled children run before this process
    # initialises the GPU
    traffic = None
    if args.traffic == "auto":
        t0 = time.perf_counter()
        try:
            traffic = measure_traffic(args)
            log(f"traffic: {json.dumps(traffic)} ({time.perf_counter() - t0:.0f} s)")
        except Exception as e:  # the bench line stands without it
            log(f"traffic measurement skipped: {e}")

    info = S.device_info(0)
    log("device", info)
    t0 = time.perf_counter()
    want_host = not args.no_cpu
    res = S.DeviceGraph.rmat(args.scale, args.edge_factor, seed=args.seed, threads=64,
                             download_edges=want_host)
    g, (h_src, h_dst, h_w) = res if want_host else (res, (None, None, None))
    V, E = g.V, g.E
    log(f"graph: scale {args.scale} V={V} E={E} built on device in {time.perf_counter()-t0:.1f} s")
    sources = bench_sources(g, args)
    log("sources", sources.tolist())

    mlp = None
    if os.path.exists(WEIGHTS):
        mlp = S.MLP(0)
        mlp.set_params(np.fromfile(WEIGHTS, np.float32))
        g.set_mlp(mlp)
    log("scheduler network:", "loaded" if mlp else "absent (learned schedule = heuristic)")

    # correctness: learned == heuristic == plain bit for bit, certificate holds
    certified = None
    if not args.no_check:
        certified = True
        for s in sources:
            dl, _ = g.solve(int(s), "learned")
            ok, counts, _ = g.certify(int(s), parents=False)
            dh, _ = g.solve(int(s), "heuristic")
            dp, _ = g.solve(int(s), "plain")
            same = (np.array_equal(dl.view(np.uint32), dp.view(np.uint32))
                    and np.array_equal(dh.view(np.uint32), dp.view(np.uint32)))
            certified &= bool(ok and same)
            log(f"source {int(s)}: certificate {ok} {counts.tolist()} identical {same}")

    # timed runs; "plain_edge_list" (the plain passes over the XCD-partitioned
    # edge list) and "all_dense" (the round engine, every round an all-edge
    # round run dense) beside the three schedules
    results = {}
    for sched in ("learned", "heuristic", "plain", "plain_edge_list", "all_dense"):
        if sched == "all_dense":
            frac = g.get_param("dense_all_frac")
            g.set_param("dense_all_frac", 0.0)
            wall, tdev, stats = time_schedule(g, sources, "all", args.steps, args.warmup)
            g.set_param("dense_all_frac", frac)
        elif sched == "plain_edge_list":
            g.set_param("plain_kernel", 1)
            wall, tdev, stats = time_schedule(g, sources, "plain", args.steps, args.warmup)
            g.set_param("plain_kernel", 0)
        else:
            wall, tdev, stats = time_schedule(g, sources, sched, args.steps, args.warmup)
        results[sched] = {
            "gteps": args.steps * E / wall / 1e9,
            "gteps_device": args.steps * E / (sum(tdev) * 1e-3) / 1e9,
            "ms_per_step": wall * 1e3 / args.steps,
            "rounds": float(np.mean([s["iterations"] for s in stats])),
            "push": float(np.mean([s["n_push"] for s in stats])),
            "all": float(np.mean([s["n_all"] for s in stats])),
            "pull": float(np.mean([s["n_pull"] for s in stats])),
            "mlp_rounds": float(np.mean([s["mlp_rounds"] for s in stats])),
            "edges_relaxed_per_E": float(np.mean([s["edges_relaxed"] for s in stats])) / E,
            # per-source device GTEPS (sample/sssp.cpp:148-151 averages per source)
            "gteps_source_mean": float(np.mean([E / (t * 1e-3) / 1e9 for t in tdev])),
            "gteps_source_hmean": float(len(tdev) / np.sum([(t * 1e-3) * 1e9 / E for t in tdev])),
        }
        log(sched, json.dumps(results[sched]))

    # rooflines, from HIP events recorded around each round on the graph's
    # stream: the plain pass (the default kernel, then k_el_pass) over a plain
    # solve of the first source, and the learned schedule's dominant kernel,
    # the dense all-edge round (k_blk_pass: one launch per dense round and per
    # fixpoint check, nothing else inside its events), over a learned solve of
    # the same source. All stream the edges once and read and write d once:
    # 12 E + 8 V algorithmic bytes per launch.
    algo_bytes = 12 * E + 8 * V
    kern = pass_kernel(g)
    g.set_param("trace_events", 1)
    _, st = g.solve(int(sources[0]), "plain", download=False)
    tr = g.trace()
    g.set_param("plain_kernel", 1)
    g.solve(int(sources[0]), "plain", download=False)
    tre = g.trace()
    g.set_param("plain_kernel", 0)
    g.solve(int(sources[0]), "learned", download=False)
    trl = g.trace()
    g.set_param("trace_events", 0)
    dense_ms_all = trl["ms"][np.isin(trl["mode"], (S.MODE_VERIFY, S.MODE_DENSE))]
    dense_ms = float(np.mean(dense_ms_all))
    copy_gbs = copy_bandwidth()
    log(f"device-to-device copy: {copy_gbs} GB/s")
    dense_t = traffic["dense_round"] if traffic else None
    achieved = algo_bytes / (dense_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm",
                "kernel": "k_blk_pass (dense all-edge round, destination-blocked edge list)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(dense_t["bytes_per_launch"]) if dense_t else None,
                "traffic_detail": dense_t,
                "algorithmic_bytes_per_launch": algo_bytes,
                "avg_launch_ms": round(dense_ms, 4), "launches": int(len(dense_ms_all)),
                "achievable_copy_GBs": round(copy_gbs, 1) if copy_gbs else None}
    if traffic:
        roofline["calibration"] = {"fetch_factor": traffic["fetch_factor"],
                                   "write_factor": traffic["write_factor"]}
    labels = {BLOCKED_PASS: "k_blk_plain (plain pass, destination-blocked edge list)",
              EDGE_LIST_PASS: "k_el_pass (plain pass, XCD-partitioned src-sorted edge list)"}

    def pass_roofline(name, ms):
        pt = traffic["passes"].get(name) if traffic else None
        a = algo_bytes / (float(np.mean(ms)) * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": labels[name], "achieved": round(a, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(a / HBM_PEAK_GBS, 4),
                "traffic": round(pt["bytes_per_launch"]) if pt else None, "traffic_detail": pt,
                "algorithmic_bytes_per_launch": algo_bytes,
                "avg_launch_ms": round(float(np.mean(ms)), 4), "launches": int(len(ms))}

    plain_roofline = pass_roofline(kern, tr["ms"])
    edge_list_roofline = pass_roofline(EDGE_LIST_PASS, tre["ms"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(h_src, h_dst, h_w, V, sources[0], int(st["iterations"]),
                           args.cpu_passes)
        log("cpu", json.dumps(cpu))

    learned = results["learned"]
    line = {
        "metric": METRIC,
        "value": round(learned["gteps"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(learned["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (R-MAT generated on device, seeded)",
        "config": {"workload": f"rmat{args.scale}_ef{args.edge_factor}_{args.sources}src",
                   "scale": args.scale, "edge_factor": args.edge_factor, "vertices": V,
                   "edges": E, "sources": [int(s) for s in sources], "schedule": "learned",
                   "parallelism": "single GPU"},
        "plain_gteps": round(results["plain"]["gteps"], 3),
        "plain_kernel": kern,
        "plain_edge_list_gteps": round(results["plain_edge_list"]["gteps"], 3),
        "all_dense_gteps": round(results["all_dense"]["gteps"], 3),
        "heuristic_gteps": round(results["heuristic"]["gteps"], 3),
        "speedup_vs_plain": round(learned["gteps"] / results["plain"]["gteps"], 3),
        "speedup_vs_plain_edge_list": round(learned["gteps"] / results["plain_edge_list"]["gteps"], 3),
        "learned_vs_heuristic": round(learned["gteps"] / results["heuristic"]["gteps"], 3),
        "certified_identical": certified,
        "scheduler_network": "loaded" if mlp else "absent",
        "schedules": results,
        "roofline": roofline,
        "plain_pass_roofline": plain_roofline,
        "edge_list_pass_roofline": edge_list_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
