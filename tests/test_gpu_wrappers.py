"""GPU parity of the reference-shaped wrappers over caller-owned device arrays,
for the three weight types: init_distances_wrapper<W> and
bellman_ford_wrapper<W> (algorithm/sssp/gpu/bellman_ford.cu:38-144,
instantiated for float and double at :150-156; unsigned is the integer-weight
configuration). Results are compared bit for bit with the oracle's
cpu_bellman_ford restatement (sssp.hpp:3-71), itself pinned to the reference's
compiled code for float and double; unreached vertices must hold FLT_MAX,
DBL_MAX or 0xFFFFFFFF. The arrays are checked before any pass: bad ids and
negative or NaN weights give SSSP_ERR_ARG.
"""
import ctypes as C

import numpy as np
import pytest

import sssp_amd as S
from conftest import HipBuffer

DBL_MAX = np.finfo(np.float64).max
INF = {np.float32: S.F32_INF, np.float64: DBL_MAX, np.uint32: S.U32_INF}
SUFFIX = {np.float32: "f32", np.float64: "f64", np.uint32: "u32"}


def _graph(dtype):
    s, d, w32 = S.rmat(12, 16, seed=21, threads=64)
    rng = np.random.default_rng(5)
    if dtype == np.float32:
        w = w32
    elif dtype == np.float64:
        w = rng.random(len(s))  # full 53-bit doubles
        w[::97] = 0.0
    else:
        w = rng.integers(1, 256, len(s)).astype(np.uint32)
    return s, d, w, 1 << 12


def _bits(a):
    return a.view({4: np.uint32, 8: np.uint64}[a.itemsize])


def _run(lib, dtype, s, d, w, V, source):
    with HipBuffer(s) as ds, HipBuffer(d) as dd, HipBuffer(w) as dw, \
            HipBuffer(np.zeros(V, dtype)) as dist:
        k = SUFFIX[dtype]
        rc = getattr(lib, "sssp_init_distances_" + k)(dist.ptr, V, source)
        assert rc == S.OK, S.last_error()
        init = dist.get()
        passes = C.c_int32(0)
        rc = getattr(lib, "sssp_bellman_ford_" + k)(ds.ptr, dd.ptr, dw.ptr, V, len(s), source,
                                                     dist.ptr, C.byref(passes))
        assert rc == S.OK, S.last_error()
        return init, dist.get(), passes.value


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32], ids=["f32", "f64", "u32"])
def test_wrappers_match_oracle(gpu, lib, oracle, dtype):
    s, d, w, V = _graph(dtype)
    for source in (0, int(s[len(s) // 2])):
        print(f"wrapper {SUFFIX[dtype]} RMAT-12 ef16 source {source}", flush=True)
        init, got, passes = _run(lib, dtype, s, d, w, V, source)
        # init_distances_wrapper: the type's max everywhere but the source
        exp_init = np.full(V, INF[dtype], dtype)
        exp_init[source] = 0
        assert np.array_equal(_bits(init), _bits(exp_init))
        ref, ref_passes = oracle.bf(s, d, w, V, source)
        assert np.array_equal(_bits(got), _bits(ref))
        unreached = ref == INF[dtype]
        assert unreached.any() and (~unreached).sum() > 1
        assert np.all(_bits(got[unreached]) == _bits(np.array([INF[dtype]], dtype))[0])
        assert 1 <= passes <= V


@pytest.mark.gpu
def test_wrappers_reject_bad_arrays(gpu, lib):
    s, d, w, V = _graph(np.float32)
    bad_w = w.copy()
    bad_w[100] = -0.5
    nan_w = np.random.default_rng(1).random(len(s))
    nan_w[7] = np.nan
    bad_id = d.copy()
    bad_id[3] = V
    cases = [("f32 negative weight", np.float32, s, d, bad_w),
             ("f64 NaN weight", np.float64, s, d, nan_w),
             ("u32 id out of range", np.uint32, s, bad_id, np.ones(len(s), np.uint32))]
    for name, dtype, ss, dd_, ww in cases:
        print("wrapper rejects:", name, flush=True)
        with HipBuffer(ss) as a, HipBuffer(dd_) as b, HipBuffer(ww.astype(dtype)) as c, \
                HipBuffer(np.zeros(V, dtype)) as dist:
            k = SUFFIX[dtype]
            assert getattr(lib, "sssp_init_distances_" + k)(dist.ptr, V, 0) == S.OK
            rc = getattr(lib, "sssp_bellman_ford_" + k)(a.ptr, b.ptr, c.ptr, V, len(ss), 0,
                                                         dist.ptr, None)
            assert rc == S.ERR_ARG, name
            msg = S.last_error()
            assert ("weight" in msg) if "weight" in name else ("outside" in msg), msg
    with HipBuffer(s) as a, HipBuffer(bad_id) as b, HipBuffer(np.zeros(V, np.uint8)) as vis:
        assert lib.sssp_bfs_edges(a.ptr, b.ptr, V, len(s), vis.ptr, 0, None) == S.ERR_ARG
