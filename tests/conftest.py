"""Shared fixtures: the library binding, the CPU oracle (ctypes) and helpers.

The oracle (oracle/liboracle.so) is the checker only; the product path under
test is libsssp_mi355x.so.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parallel-task1_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import sssp_amd as S  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """ctypes view of oracle/liboracle.so."""

    def __init__(self, path):
        L = C.CDLL(path)
        i32, i64, p = C.c_int32, C.c_int64, C.c_void_p
        for n in ("orc_bf_f32", "orc_bf_u32", "orc_bf_f64"):
            getattr(L, n).argtypes = [p, p, p, i32, i64, i32, i32, p]
            getattr(L, n).restype = C.c_int
        for n in ("orc_dijkstra_f32", "orc_dijkstra_u32"):
            getattr(L, n).argtypes = [p, p, p, i32, i64, i32, p]
            getattr(L, n).restype = C.c_int
        for n in ("orc_certify_f32", "orc_certify_u32"):
            getattr(L, n).argtypes = [p, p, p, i32, i64, i32, p, p, p]
            getattr(L, n).restype = C.c_int
        L.orc_bf_f32_capped.argtypes = [p, p, p, i32, i64, i32, i32, i32, p, p]
        L.orc_bf_f32_capped.restype = C.c_int
        L.orc_max_threads.restype = C.c_int
        L.orc_wtime.restype = C.c_double
        L.orc_bfs.argtypes = [p, p, i32, i64, i32, p, p]
        L.orc_bfs.restype = C.c_int64
        self.L = L

    def bfs(self, src, dst, V, source):
        """(visited uint8, level uint32, reached count)"""
        vis = np.empty(V, np.uint8)
        lev = np.empty(V, np.uint32)
        n = self.L.orc_bfs(_ptr(np.ascontiguousarray(src, np.int32)),
                           _ptr(np.ascontiguousarray(dst, np.int32)), V, len(src), source,
                           _ptr(vis), _ptr(lev))
        assert n >= 0
        return vis, lev, n

    def bf(self, src, dst, w, V, source, threads=0):
        w = np.ascontiguousarray(w)
        kind = {np.dtype(np.float32): "f32", np.dtype(np.uint32): "u32",
                np.dtype(np.float64): "f64"}[w.dtype]
        d = np.empty(V, w.dtype)
        it = getattr(self.L, "orc_bf_" + kind)(_ptr(src), _ptr(dst), _ptr(w), V, len(src), source,
                                               threads, _ptr(d))
        return d, it

    def bf_capped(self, src, dst, w, V, source, max_passes, threads=0):
        """(distances, passes run, converged) of the pass-capped f32 loop."""
        d = np.empty(V, np.float32)
        conv = C.c_int32(-1)
        it = self.L.orc_bf_f32_capped(_ptr(src), _ptr(dst), _ptr(np.ascontiguousarray(w)), V,
                                      len(src), source, threads, max_passes, _ptr(d),
                                      C.byref(conv))
        return d, it, conv.value

    def dijkstra(self, src, dst, w, V, source):
        w = np.ascontiguousarray(w)
        kind = "f32" if w.dtype == np.float32 else "u32"
        d = np.empty(V, w.dtype)
        rc = getattr(self.L, "orc_dijkstra_" + kind)(_ptr(src), _ptr(dst), _ptr(w), V, len(src),
                                                     source, _ptr(d))
        assert rc == 0
        return d

    def certify(self, src, dst, w, V, source, dist, parent=None):
        w = np.ascontiguousarray(w)
        kind = "f32" if w.dtype == np.float32 else "u32"
        counts = np.zeros(8, np.int64)
        rc = getattr(self.L, "orc_certify_" + kind)(_ptr(src), _ptr(dst), _ptr(w), V, len(src),
                                                    source, _ptr(dist), _ptr(parent), _ptr(counts))
        return rc, counts


@pytest.fixture(scope="session")
def oracle():
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        pytest.fail("oracle/liboracle.so not built (run __graft_entry__.build())")
    return Oracle(path)


@pytest.fixture(scope="session")
def ref_lib():
    """The reference's own CPU code compiled into oracle/_ref (absent -> skip)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libref_sssp.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference tree at build time)")
    L = C.CDLL(path)
    p, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.ref_rmat_parallel.restype = i64
    L.ref_rmat_parallel.argtypes = [i32, i32, i32, i32, i32, i32, p, p, p]
    L.ref_random_uniform_f32.restype = i64
    L.ref_random_uniform_f32.argtypes = [i32, i32, C.c_uint32, p, p, p]
    L.ref_cpu_bellman_ford_f32.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_cpu_bellman_ford_u32.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_cpu_bellman_ford_f64.argtypes = [p, p, p, i32, i64, i32, i32, p]
    L.ref_save_bin_f32.argtypes = [C.c_char_p, p, p, p, i32, i64]
    L.ref_load_bin_f32.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i64), p, p, p]
    L.ref_cpu_sequential_bfs.argtypes = [p, p, i32, i64, i32, p]
    return L


@pytest.fixture(scope="session")
def lib():
    return S.lib()


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; used by every @pytest.mark.gpu test."""
    n = S.device_count()
    if n < 1:
        pytest.skip("no GPU visible")
    return 0


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def small_graphs():
    """Hand-built graphs with known answers (closed forms / edge cases)."""
    gs = []
    # chain 0->1->...->9 with weight 0.5
    n = 10
    gs.append(("chain", n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32),
               np.full(n - 1, 0.5, np.float32)))
    # star from 0
    n = 65
    gs.append(("star", n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32),
               np.linspace(0.1, 1.0, n - 1).astype(np.float32)))
    # disconnected: two chains, source in the first
    gs.append(("disconnected", 6, np.array([0, 1, 3, 4], np.int32), np.array([1, 2, 4, 5], np.int32),
               np.array([1, 1, 1, 1], np.float32)))
    # self loops, duplicates, zero weights
    gs.append(("loops_dups", 4, np.array([0, 0, 0, 1, 1, 2, 2], np.int32),
               np.array([0, 1, 1, 2, 1, 3, 3], np.int32),
               np.array([0.0, 0.7, 0.3, 0.0, 0.2, 0.5, 0.25], np.float32)))
    # single vertex, no edges
    gs.append(("empty", 1, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)))
    return gs


class HipBuffer:
    """Device memory for the tests of the reference-shaped wrappers, which take
    caller-owned device pointers (hipMalloc / hipMemcpy through ctypes)."""

    _hip = None

    def __init__(self, host: np.ndarray):
        if HipBuffer._hip is None:
            HipBuffer._hip = C.CDLL("libamdhip64.so")
        self.dtype, self.n = host.dtype, host.size
        self.nbytes = max(host.nbytes, 1)
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        if host.nbytes:
            assert self._hip.hipMemcpy(self.ptr, _ptr(np.ascontiguousarray(host)),
                                       C.c_size_t(host.nbytes), 1) == 0  # host to device

    def get(self):
        out = np.empty(self.n, self.dtype)
        if out.nbytes:
            assert self._hip.hipMemcpy(_ptr(out), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr.value:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
