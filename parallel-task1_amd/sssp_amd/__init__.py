"""ctypes binding of libsssp_mi355x.so (the C ABI declared in include/sssp_mi355x.h).

Mirrors the reference's SSSP interface for Python callers:
  SingleSourceShortestPaths.gpu_bellman_ford(graph, distances, source) -> MTEPS
    (algorithm/sssp/sssp.h:23-41, sssp.hpp:219-278)
  GraphGenerationAPI.R_MAT_parallel / random_uniform
    (include/test/generation/graph_generation_API.h:9-26)
  GraphStorageAPI.save/load_from_edges_list_bin_file
    (include/storage/graph_storage_API.h)
plus the device-resident DeviceGraph handle the GPU path is built around.

Errors: every non-zero status raises SSSPError carrying sssp_last_error(); the
reference threw `const char*` (include/common/cuda_error_hadling.h:4-24).
The library must exist: nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsssp_mi355x.so")

OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_INTERNAL, ERR_IO, ERR_NOT_CONVERGED = range(7)
W_F32, W_U32, W_F64 = 0, 1, 2
SCHED_PLAIN, SCHED_HEURISTIC, SCHED_LEARNED, SCHED_PUSH, SCHED_ALL, SCHED_PULL = range(6)
MODE_PUSH, MODE_ALL, MODE_PULL, MODE_VERIFY, MODE_DENSE, MODE_TAIL = 0, 1, 2, 3, 4, 5
INPUT_DEVICE, NO_CSC, NO_RELABEL, RELABEL = 1, 2, 4, 8
MLP_IN, MLP_HID, MLP_MODES, MLP_CHUNKS = 5, 64, 3, 4
MLP_OUT = MLP_MODES + MLP_CHUNKS
MLP_NPARAM = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT
F32_INF = np.float32(np.finfo(np.float32).max)
U32_INF = np.uint32(0xFFFFFFFF)
SCHED_NAMES = {"plain": SCHED_PLAIN, "heuristic": SCHED_HEURISTIC, "learned": SCHED_LEARNED,
               "push": SCHED_PUSH, "all": SCHED_ALL, "pull": SCHED_PULL}


class SSSPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sssp error {code}: {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("n_push", C.c_int32), ("n_all", C.c_int32),
                ("n_pull", C.c_int32), ("verify_passes", C.c_int32), ("mlp_rounds", C.c_int32),
                ("reached", C.c_int32), ("reserved", C.c_int32), ("edges_relaxed", C.c_int64),
                ("t_solve_ms", C.c_double), ("gteps", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None

P = C.c_void_p
I32, I64, U32, U64, F32, F64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float, C.c_double
PI32, PI64, PF32 = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_float)

_SIGS = {
    "sssp_last_error": (C.c_char_p, []),
    "sssp_version": (I32, []),
    "sssp_device_count": (I32, [PI32]),
    "sssp_device_info": (I32, [I32, C.c_char_p, I32, PI32, PI64, PI32]),
    "sssp_device_props": (I32, [I32, C.c_char_p, I32, PI64, PI64, PI32, PI32]),
    "sssp_graph_create": (I32, [I32, P, P, P, I32, I64, I32, U32, C.POINTER(P)]),
    "sssp_graph_create_from_bin": (I32, [C.c_char_p, I32, I32, I32, U32, C.POINTER(P)]),
    "sssp_graph_create_rmat": (I32, [I32, I32, I32, I32, I32, I32, U32, I32, I32, U32, P, P, P,
                                     C.POINTER(P)]),
    "sssp_graph_destroy": (I32, [P]),
    "sssp_graph_info": (I32, [P, PI32, PI64, PI32, PI64]),
    "sssp_graph_download_layout": (I32, [P, P, P, P, P, P, P, P]),
    "sssp_graph_vertex_order": (I32, [P, P]),
    "sssp_solve": (I32, [P, I32, I32, P, C.POINTER(Stats)]),
    "sssp_get_distances": (I32, [P, P]),
    "sssp_get_trace": (I32, [P, I32, P, P, P, P, PI32]),
    "sssp_certify": (I32, [P, I32, P, PI64, PI32]),
    "sssp_set_param": (I32, [P, C.c_char_p, F64]),
    "sssp_get_param": (I32, [P, C.c_char_p, C.POINTER(F64)]),
    "sssp_solve_batch": (I32, [P, P, I32, I32, P, C.POINTER(Stats)]),
    "sssp_pack_create": (I32, [P, P, P, P, P, I32, I32, C.POINTER(P)]),
    "sssp_pack_destroy": (I32, [P]),
    "sssp_pack_solve": (I32, [P, P, I32, P, C.POINTER(Stats)]),
    "sssp_autotune_batch": (I32, [P, P, I32, C.c_char_p, C.POINTER(F64)]),
    "sssp_autotune_pack": (I32, [P, P, C.c_char_p, C.POINTER(F64)]),
    "sssp_init_distances_f32": (I32, [P, I32, I32]),
    "sssp_init_distances_f64": (I32, [P, I32, I32]),
    "sssp_init_distances_u32": (I32, [P, I32, I32]),
    "sssp_bellman_ford_f32": (I32, [P, P, P, I32, I64, I32, P, PI32]),
    "sssp_bfs": (I32, [P, I32, I32, P, P, P]),
    "sssp_bfs_edges": (I32, [P, P, I32, I64, P, I32, PI32]),
    "sssp_bellman_ford_f64": (I32, [P, P, P, I32, I64, I32, P, PI32]),
    "sssp_bellman_ford_u32": (I32, [P, P, P, I32, I64, I32, P, PI32]),
    "sssp_bellman_ford_noalloc_f32": (I32, [P, P, P, I32, I64, I32, P, PI32]),
    "sssp_rmat_generate": (I32, [I32, I32, I32, I32, I32, I32, U32, I32, P, P, P]),
    "sssp_er_generate": (I32, [I32, I32, U32, I32, P, P, P]),
    "sssp_rmat_generate_many": (I32, [I32, P, I32, I32, I32, I32, I32, U32, P, P, P, P]),
    "sssp_save_bin": (I32, [C.c_char_p, I32, P, P, P, I32, I32, I64]),
    "sssp_load_bin_header": (I32, [C.c_char_p, PI32, PI32, PI64]),
    "sssp_load_bin": (I32, [C.c_char_p, I32, P, P, P, I64]),
    "sssp_pick_sources": (I32, [P, I64, I32, U64, I32, P]),
    "sssp_mlp_create": (I32, [I32, U64, C.POINTER(P)]),
    "sssp_mlp_destroy": (I32, [P]),
    "sssp_mlp_get_params": (I32, [P, P]),
    "sssp_mlp_set_params": (I32, [P, P]),
    "sssp_mlp_forward": (I32, [P, P, I32, P]),
    "sssp_mlp_loss_grad": (I32, [P, P, P, P, I32, PF32, P]),
    "sssp_mlp_train": (I32, [P, P, P, P, I32, I32, I32, F32, U64, P]),
    "sssp_graph_set_mlp": (I32, [P, P]),
    "sssp_graph_set_bfs_mlp": (I32, [P, P]),
    "sssp_collect_timings": (I32, [P, I32, I32, P, P, PI32]),
    "sssp_features": (I32, [F64, F64, F64, F64, F64, F64, F64, P]),
    "sssp_debug_fail_next_alloc": (I32, []),
}


def lib():
    """Loads the library (once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SSSPError(ERR_INTERNAL, f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def last_error() -> str:
    """Message of the last failed call on this thread (sssp_last_error)."""
    return lib().sssp_last_error().decode(errors="replace")


def _check(rc):
    if rc != OK:
        raise SSSPError(rc, last_error())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    _check(lib().sssp_device_count(C.byref(n)))
    return n.value


def device_info(device: int = 0) -> dict:
    name = C.create_string_buffer(128)
    cu, hbm, l2 = C.c_int32(), C.c_int64(), C.c_int32()
    _check(lib().sssp_device_info(device, name, 128, C.byref(cu), C.byref(hbm), C.byref(l2)))
    return {"arch": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value,
            "l2_bytes": l2.value}


# --------------------------------------------------------------------------
# host services
# --------------------------------------------------------------------------
def rmat(scale: int, edge_factor: int, a=57, b=19, c=19, d=5, seed: int = 1, threads: int = 64):
    """R_MAT_parallel with a fixed base seed and virtual thread count."""
    E = (1 << scale) * edge_factor
    src = np.empty(E, np.int32)
    dst = np.empty(E, np.int32)
    w = np.empty(E, np.float32)
    _check(lib().sssp_rmat_generate(scale, edge_factor, a, b, c, d, seed & 0xFFFFFFFF, threads,
                                    _ptr(src), _ptr(dst), _ptr(w)))
    return src, dst, w


def rmat_many(scales, edge_factor: int, seed0: int, a=57, b=19, c=19, d=5):
    scales = np.asarray(scales, np.int32)
    counts = (np.int64(1) << scales.astype(np.int64)) * edge_factor
    offs = np.zeros(len(scales) + 1, np.int64)
    np.cumsum(counts, out=offs[1:])
    E = int(offs[-1])
    src = np.empty(E, np.int32)
    dst = np.empty(E, np.int32)
    w = np.empty(E, np.float32)
    _check(lib().sssp_rmat_generate_many(len(scales), _ptr(scales), edge_factor, a, b, c, d,
                                         seed0 & 0xFFFFFFFF, _ptr(offs), _ptr(src), _ptr(dst),
                                         _ptr(w)))
    return src, dst, w, offs


def er(vertices: int, avg_degree: int, seed: int, weight_mode: int = 1):
    E = vertices * avg_degree
    src = np.empty(E, np.int32)
    dst = np.empty(E, np.int32)
    w = np.empty(E, np.float32 if weight_mode == 0 else np.uint32)
    _check(lib().sssp_er_generate(vertices, avg_degree, seed & 0xFFFFFFFF, weight_mode,
                                  _ptr(src), _ptr(dst), _ptr(w)))
    return src, dst, w


def pick_sources(src, vertices: int, seed: int, n: int):
    out = np.empty(n, np.int32)
    _check(lib().sssp_pick_sources(_ptr(np.ascontiguousarray(src, np.int32)), len(src), vertices,
                                   seed, n, _ptr(out)))
    return out


def save_bin(path, src, dst, w, vertices, directed=True):
    w = np.ascontiguousarray(w)
    _check(lib().sssp_save_bin(str(path).encode(), 1 if directed else 0,
                               _ptr(np.ascontiguousarray(src, np.int32)),
                               _ptr(np.ascontiguousarray(dst, np.int32)), _ptr(w), w.itemsize,
                               vertices, len(src)))


def load_bin(path, weight_dtype=np.float32):
    d, V, E = C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib().sssp_load_bin_header(str(path).encode(), C.byref(d), C.byref(V), C.byref(E)))
    src = np.empty(E.value, np.int32)
    dst = np.empty(E.value, np.int32)
    w = np.empty(E.value, weight_dtype)
    _check(lib().sssp_load_bin(str(path).encode(), w.itemsize, _ptr(src), _ptr(dst), _ptr(w),
                               E.value))
    return bool(d.value), V.value, src, dst, w


def features(fv, fe, unsettled, rnd, mean_deg, V, E):
    out = np.empty(MLP_IN, np.float32)
    _check(lib().sssp_features(fv, fe, unsettled, rnd, mean_deg, V, E, _ptr(out)))
    return out


# --------------------------------------------------------------------------
# device graph
# --------------------------------------------------------------------------
def _flags(csc, relabel):
    """relabel: None = the library's choice (degree-ordered slots on skewed
    graphs from 2^23 vertices), True = always, False = never."""
    order = 0 if relabel is None else (RELABEL if relabel else NO_RELABEL)
    return (0 if csc else NO_CSC) | order


class DeviceGraph:
    """A graph uploaded once to HBM (edge list sorted by src, CSR, CSC)."""

    @classmethod
    def rmat(cls, scale, edge_factor, a=57, b=19, c=19, d=5, seed=1, threads=64, device=0,
             csc=True, download_edges=False, relabel=None):
        """Generates the R-MAT graph on the device (bit-identical to rmat())
        and builds it there; optionally returns the generated edge list."""
        self = cls.__new__(cls)
        E = (1 << scale) * edge_factor
        arrs = (np.empty(E, np.int32), np.empty(E, np.int32), np.empty(E, np.float32)) \
            if download_edges else (None, None, None)
        h = C.c_void_p()
        _check(lib().sssp_graph_create_rmat(scale, edge_factor, a, b, c, d, seed & 0xFFFFFFFF,
                                            threads, device, _flags(csc, relabel),
                                            _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                            C.byref(h)))
        self._h = h
        self.device = device
        self.wtype, self.wdtype = W_F32, np.float32
        self.V, self.E = 1 << scale, E
        return (self, arrs) if download_edges else self

    @classmethod
    def from_bin(cls, path, directed=True, weight_dtype=np.float32, device=0, csc=True,
                 relabel=None):
        """Device ingest of a binary edge-list file (the reference's format):
        chunks stream through pinned buffers onto the device, where the
        layouts are built; no host copy of the graph is made."""
        self = cls.__new__(cls)
        if np.dtype(weight_dtype) == np.float32:
            self.wtype, self.wdtype = W_F32, np.float32
        elif np.dtype(weight_dtype) == np.uint32:
            self.wtype, self.wdtype = W_U32, np.uint32
        else:
            raise SSSPError(ERR_ARG, f"unsupported weight dtype {weight_dtype}")
        h = C.c_void_p()
        _check(lib().sssp_graph_create_from_bin(str(path).encode(), 1 if directed else 0,
                                                self.wtype, device, _flags(csc, relabel),
                                                C.byref(h)))
        self._h = h
        self.device = device
        V, E = C.c_int32(), C.c_int64()
        _check(lib().sssp_graph_info(h, C.byref(V), C.byref(E), None, None))
        self.V, self.E = V.value, E.value
        return self

    def __init__(self, src, dst, w, vertices: int, device: int = 0, csc: bool = True,
                 relabel=None):
        w = np.ascontiguousarray(w)
        if w.dtype == np.float32:
            self.wtype, self.wdtype = W_F32, np.float32
        elif w.dtype == np.uint32:
            self.wtype, self.wdtype = W_U32, np.uint32
        else:
            raise SSSPError(ERR_ARG, f"unsupported weight dtype {w.dtype}")
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        if not (len(src) == len(dst) == len(w)):
            raise SSSPError(ERR_ARG, "src/dst/w lengths differ")
        self.V, self.E = int(vertices), len(src)
        h = C.c_void_p()
        _check(lib().sssp_graph_create(self.wtype, _ptr(src), _ptr(dst), _ptr(w), self.V, self.E,
                                       device, _flags(csc, relabel), C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().sssp_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_param(self, name: str, value: float):
        _check(lib().sssp_set_param(self._h, name.encode(), float(value)))

    def get_param(self, name: str) -> float:
        v = C.c_double()
        _check(lib().sssp_get_param(self._h, name.encode(), C.byref(v)))
        return v.value

    def solve(self, source: int, schedule="heuristic", download=True):
        sched = SCHED_NAMES[schedule] if isinstance(schedule, str) else int(schedule)
        st = Stats()
        d = np.empty(self.V, self.wdtype) if download else None
        _check(lib().sssp_solve(self._h, int(source), sched, _ptr(d), C.byref(st)))
        return d, st.as_dict()

    def bfs(self, source: int, schedule="heuristic", levels=True):
        """BFS from `source` on the same engine (weights ignored): returns the
        reference's visited flags (bool), the hop levels (uint32, 0xFFFFFFFF
        unreached) when asked, and the solve stats."""
        sched = SCHED_NAMES[schedule] if isinstance(schedule, str) else int(schedule)
        st = Stats()
        vis = np.empty(self.V, np.uint8)
        lev = np.empty(self.V, np.uint32) if levels else None
        _check(lib().sssp_bfs(self._h, int(source), sched, _ptr(vis), _ptr(lev), C.byref(st)))
        return vis.view(np.bool_), lev, st.as_dict()

    def distances(self):
        d = np.empty(self.V, self.wdtype)
        _check(lib().sssp_get_distances(self._h, _ptr(d)))
        return d

    def trace(self, cap: int = 1 << 16):
        modes = np.empty(cap, np.int32)
        fr = np.empty(cap, np.int64)
        fe = np.empty(cap, np.int64)
        ms = np.empty(cap, np.float32)
        n = C.c_int32()
        _check(lib().sssp_get_trace(self._h, cap, _ptr(modes), _ptr(fr), _ptr(fe), _ptr(ms),
                                    C.byref(n)))
        k = min(n.value, cap)
        return {"mode": modes[:k], "frontier": fr[:k], "frontier_edges": fe[:k], "ms": ms[:k]}

    def certify(self, source: int, parents: bool = True):
        counts = np.zeros(8, np.int64)
        par = np.empty(self.V, np.int32) if parents else None
        valid = C.c_int32()
        _check(lib().sssp_certify(self._h, int(source), _ptr(par),
                                  counts.ctypes.data_as(PI64), C.byref(valid)))
        return bool(valid.value), counts, par

    def vertex_order(self):
        """slot[v]: the position of vertex v in the device layouts (the
        identity when the graph keeps its vertex ids)."""
        o = np.empty(self.V, np.int32)
        _check(lib().sssp_graph_vertex_order(self._h, _ptr(o)))
        return o

    def out_degrees(self):
        """Out-degree of every vertex, in the caller's vertex ids."""
        deg = np.diff(self.layout(csc=False)["row_ptr"])
        return deg[self.vertex_order()]

    def layout(self, csc=True):
        """The device layouts; vertices are numbered by slot (vertex_order)."""
        V1, E = self.V + 1, self.E
        rp = np.empty(V1, np.int64)
        s = np.empty(E, np.int32)
        t = np.empty(E, np.int32)
        w = np.empty(E, self.wdtype)
        cp = np.empty(V1, np.int64) if csc else None
        cs = np.empty(E, np.int32) if csc else None
        cw = np.empty(E, self.wdtype) if csc else None
        _check(lib().sssp_graph_download_layout(self._h, _ptr(rp), _ptr(s), _ptr(t), _ptr(w),
                                                _ptr(cp), _ptr(cs), _ptr(cw)))
        return {"row_ptr": rp, "el_src": s, "el_dst": t, "el_w": w, "col_ptr": cp, "csc_src": cs,
                "csc_w": cw}

    def collect_timings(self, source: int, cap: int = 4096):
        nt = MLP_MODES * MLP_CHUNKS
        f = np.empty((cap, MLP_IN), np.float32)
        t = np.empty((cap, nt), np.float32)
        n = C.c_int32()
        _check(lib().sssp_collect_timings(self._h, int(source), cap, _ptr(f), _ptr(t), C.byref(n)))
        return f[:n.value].copy(), t[:n.value].copy()

    def set_mlp(self, mlp):
        _check(lib().sssp_graph_set_mlp(self._h, None if mlp is None else mlp.handle))
        self._mlp = mlp

    def set_bfs_mlp(self, mlp):
        """Scheduler network of learned BFS solves (None: the rule)."""
        self._bfs_mlp = mlp
        _check(lib().sssp_graph_set_bfs_mlp(self._h, mlp.handle if mlp else None))

    def solve_batch(self, sources, policy=0, download=True):
        sources = np.ascontiguousarray(sources, np.int32)
        st = Stats()
        d = np.empty((len(sources), self.V), self.wdtype) if download else None
        _check(lib().sssp_solve_batch(self._h, _ptr(sources), len(sources), policy, _ptr(d),
                                      C.byref(st)))
        return d, st.as_dict()

    def autotune_batch(self, sources, tuning_file=None):
        sources = np.ascontiguousarray(sources, np.int32)
        best = C.c_double()
        _check(lib().sssp_autotune_batch(self._h, _ptr(sources), len(sources),
                                         None if tuning_file is None else str(tuning_file).encode(),
                                         C.byref(best)))
        return best.value


class GraphPack:
    """Many independent small graphs solved together."""

    def __init__(self, src, dst, w, edge_offsets, vertex_counts, device: int = 0):
        self.offs = np.ascontiguousarray(edge_offsets, np.int64)
        self.vcounts = np.ascontiguousarray(vertex_counts, np.int32)
        self.n = len(self.vcounts)
        self.vtotal = int(self.vcounts.sum())
        h = C.c_void_p()
        _check(lib().sssp_pack_create(_ptr(np.ascontiguousarray(src, np.int32)),
                                      _ptr(np.ascontiguousarray(dst, np.int32)),
                                      _ptr(np.ascontiguousarray(w, np.float32)), _ptr(self.offs),
                                      _ptr(self.vcounts), self.n, device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sssp_pack_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, sources, policy=0, download=True):
        sources = np.ascontiguousarray(sources, np.int32)
        st = Stats()
        d = np.empty(self.vtotal, np.float32) if download else None
        _check(lib().sssp_pack_solve(self._h, _ptr(sources), policy, _ptr(d), C.byref(st)))
        return d, st.as_dict()

    def autotune(self, sources, tuning_file=None):
        sources = np.ascontiguousarray(sources, np.int32)
        best = C.c_double()
        _check(lib().sssp_autotune_pack(self._h, _ptr(sources),
                                        None if tuning_file is None else str(tuning_file).encode(),
                                        C.byref(best)))
        return best.value


class MLP:
    """The round scheduler network, resident on the device."""

    def __init__(self, device: int = 0, seed: int = 0):
        h = C.c_void_p()
        _check(lib().sssp_mlp_create(device, seed, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().sssp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        _check(lib().sssp_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_NPARAM
        _check(lib().sssp_mlp_set_params(self._h, _ptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        out = np.empty((len(x), MLP_OUT), np.float32)
        _check(lib().sssp_mlp_forward(self._h, _ptr(x), len(x), _ptr(out)))
        return out

    def loss_grad(self, x, y_mode, y_chunk):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        ym = np.ascontiguousarray(y_mode, np.int32)
        yc = np.ascontiguousarray(y_chunk, np.int32)
        loss = C.c_float()
        g = np.empty(MLP_NPARAM, np.float32)
        _check(lib().sssp_mlp_loss_grad(self._h, _ptr(x), _ptr(ym), _ptr(yc), len(x),
                                        C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, x, y_mode, y_chunk, epochs=200, batch=128, lr=3e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        ym = np.ascontiguousarray(y_mode, np.int32)
        yc = np.ascontiguousarray(y_chunk, np.int32)
        losses = np.empty(epochs, np.float32)
        _check(lib().sssp_mlp_train(self._h, _ptr(x), _ptr(ym), _ptr(yc), len(x), epochs, batch,
                                    lr, seed, _ptr(losses)))
        return losses


# --------------------------------------------------------------------------
# reference-shaped interface
# --------------------------------------------------------------------------
@dataclass
class EdgesListGraph:
    """The parts of Graph<int, W> in EDGES_LIST form that the SSSP path reads
    (get_vertices_count, get_edges_count, get_graph_data: graph.h:110-165)."""
    vertices_count: int
    src_ids: np.ndarray
    dst_ids: np.ndarray
    weights: np.ndarray
    directed: bool = True

    def get_vertices_count(self):
        return self.vertices_count

    def get_edges_count(self):
        return len(self.src_ids)


def edge_fingerprint(src, dst, w, chunk=1 << 22) -> int:
    """Order-independent 64-bit fingerprint of an edge list: the sum over
    edges of a mix of (src, dst, weight bits). Lists with the same multiset of
    edges agree; an edited id or weight changes it (up to ~2^-64)."""
    def mix(x):
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xff51afd7ed558ccd)
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xc4ceb9fe1a85ec53)
        return x ^ (x >> np.uint64(33))

    wv = np.ascontiguousarray(w)
    wv = wv.view(np.uint32 if wv.itemsize == 4 else np.uint64).astype(np.uint64)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for i in range(0, len(src), chunk):
            k = (np.asarray(src[i:i + chunk]).astype(np.uint32).astype(np.uint64) << np.uint64(32)) \
                | np.asarray(dst[i:i + chunk]).astype(np.uint32).astype(np.uint64)
            h = mix(k ^ mix(wv[i:i + chunk] + np.uint64(0x9e3779b97f4a7c15)))
            total = total + np.add.reduce(h, dtype=np.uint64)
    return int(total)


class GraphGenerationAPI:
    @staticmethod
    def R_MAT_parallel(vertices_count, average_connections, a, b, c, d, omp_threads,
                       directed=True, seed=1):
        if not directed:
            raise SSSPError(ERR_ARG, "only directed R-MAT graphs are generated "
                                     "(the reference's undirected branch overwrites edges)")
        scale = int(vertices_count).bit_length() - 1
        if (1 << scale) != vertices_count:
            raise SSSPError(ERR_ARG, "vertices_count must be a power of two")
        s, t, w = rmat(scale, average_connections, a, b, c, d, seed=seed, threads=omp_threads)
        return EdgesListGraph(vertices_count, s, t, w)

    @staticmethod
    def random_uniform(vertices_count, average_degree, seed, weight_mode=0):
        s, t, w = er(vertices_count, average_degree, seed, weight_mode)
        return EdgesListGraph(vertices_count, s, t, w)


class GraphStorageAPI:
    @staticmethod
    def save_to_edges_list_bin_file(graph: EdgesListGraph, file_name):
        save_bin(file_name, graph.src_ids, graph.dst_ids, graph.weights, graph.vertices_count,
                 graph.directed)
        return True

    @staticmethod
    def load_from_edges_list_bin_file(file_name, directed=True, weight_dtype=np.float32):
        d, V, s, t, w = load_bin(file_name, weight_dtype)
        if d != directed:
            raise SSSPError(ERR_ARG, "loading directed graph to non-directed container"
                                     "(or undirected to directed)")
        return EdgesListGraph(V, s, t, w, d)


class BFS:
    """Drop-in for BFS<int, W>::gpu_bfs (algorithm/bfs/bfs.hpp:213-253): fills
    the bool `result` with reachability from the source and returns MTEPS
    over the solve window. The device graph is cached per graph object."""

    def __init__(self, omp_threads=1, device=0, schedule="heuristic"):
        self.omp_threads = omp_threads
        self._sssp = SingleSourceShortestPaths(omp_threads, device, schedule)

    def set_omp_threads(self, n):
        self.omp_threads = n

    def gpu_bfs(self, graph: EdgesListGraph, result: np.ndarray, source_vertex: int):
        g = self._sssp._device_graph(graph)
        vis, _, st = g.bfs(source_vertex, self._sssp.schedule, levels=False)
        result[:] = vis
        return graph.get_edges_count() / (st["t_solve_ms"] * 1e-3 * 1e6)


class SingleSourceShortestPaths:
    """Drop-in for SingleSourceShortestPaths<int, W>::gpu_bellman_ford
    (sssp.hpp:219-278): fills `distances` and returns MTEPS over the solve
    window (uploads and downloads excluded). The device graph is cached per
    graph object so repeated sources do not re-upload."""

    def __init__(self, omp_threads=1, device=0, schedule="heuristic"):
        self.omp_threads = omp_threads
        self.device = device
        self.schedule = schedule
        self._cache = {}

    def set_omp_threads(self, n):
        self.omp_threads = n

    def _device_graph(self, graph: EdgesListGraph):
        # reused only while the arrays and their contents are unchanged: an
        # edit in place changes the fingerprint and the graph is uploaded again
        # (the reference uploaded on every call, sssp.hpp:245-253)
        key = (id(graph.src_ids), id(graph.dst_ids), id(graph.weights), graph.vertices_count,
               graph.get_edges_count(),
               edge_fingerprint(graph.src_ids, graph.dst_ids, graph.weights))
        g = self._cache.get(key)
        if g is None:
            self._cache.clear()
            g = DeviceGraph(graph.src_ids, graph.dst_ids, graph.weights, graph.vertices_count,
                            self.device)
            self._cache[key] = g
        return g

    def gpu_bellman_ford(self, graph: EdgesListGraph, distances: np.ndarray, source_vertex: int):
        g = self._device_graph(graph)
        d, st = g.solve(source_vertex, self.schedule)
        distances[:] = d
        return graph.get_edges_count() / (st["t_solve_ms"] * 1e-3 * 1e6)
