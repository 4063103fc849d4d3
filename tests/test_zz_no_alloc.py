"""The allocation-free f32 wrapper, sssp_bellman_ford_noalloc_f32, against the
allocating one it varies, sssp_bellman_ford_f32 (bellman_ford_wrapper<float>,
algorithm/sssp/gpu/bellman_ford.cu:103-144), through the C ABI on caller-owned
device arrays: the same distance bits on two calls in a row,
the second on a smaller graph, and on the oracle's fixpoint. Its scratch (a
file-scope device array shared by every call on the device) keeps the last
call's contents, so the checks also run after a call that the validation
rejected, and after a call on the larger graph."""
import ctypes as C

import numpy as np
import pytest

import sssp_amd as S
from conftest import HipBuffer, bits


def _wrapper(lib, name, s, d, w, V, source):
    with HipBuffer(s) as ds, HipBuffer(d) as dd, HipBuffer(w) as dw, \
            HipBuffer(np.zeros(V, np.float32)) as dist:
        assert lib.sssp_init_distances_f32(dist.ptr, V, source) == S.OK, S.last_error()
        passes = C.c_int32(-1)
        rc = getattr(lib, name)(ds.ptr, dd.ptr, dw.ptr, V, len(s), source, dist.ptr,
                                C.byref(passes))
        return rc, dist.get(), passes.value


@pytest.mark.gpu
def test_noalloc_wrapper_matches_the_allocating_one(gpu, lib, oracle):
    big = S.rmat(14, 16, seed=31, threads=64) + (1 << 14,)
    small = S.rmat(8, 4, seed=32, threads=1) + (1 << 8,)
    for label, (s, d, w, V) in (("RMAT-14 ef16", big), ("RMAT-8 ef4", small)):
        source = int(s[len(s) // 3])
        print(f"noalloc wrapper: {label} source {source}", flush=True)
        rc, got, p_got = _wrapper(lib, "sssp_bellman_ford_noalloc_f32", s, d, w, V, source)
        assert rc == S.OK, S.last_error()
        rc, exp, p_exp = _wrapper(lib, "sssp_bellman_ford_f32", s, d, w, V, source)
        assert rc == S.OK, S.last_error()
        # the pass count is not compared: within a pass, how far updates
        # propagate depends on the order the workgroups run in
        assert np.array_equal(bits(got), bits(exp))
        assert 1 <= p_got <= V + 4 and 1 <= p_exp <= V + 4
        assert np.array_equal(bits(got), bits(oracle.bf(s, d, w, V, source)[0]))


@pytest.mark.gpu
def test_noalloc_wrapper_checks_and_clears_its_scratch(gpu, lib, oracle):
    s, d, w = S.rmat(10, 8, seed=33, threads=8)
    V = 1 << 10
    bad = w.copy()
    bad[5] = -1.0
    print("noalloc wrapper: negative weight rejected", flush=True)
    rc, _, _ = _wrapper(lib, "sssp_bellman_ford_noalloc_f32", s, d, bad, V, 0)
    assert rc == S.ERR_ARG and "weight" in S.last_error()
    # the rejected call left its counts in the scratch; the next call clears them
    print("noalloc wrapper: valid call after the rejected one", flush=True)
    rc, got, _ = _wrapper(lib, "sssp_bellman_ford_noalloc_f32", s, d, w, V, 0)
    assert rc == S.OK, S.last_error()
    assert np.array_equal(bits(got), bits(oracle.bf(s, d, w, V, 0)[0]))
    # no device allocation: with the next allocation set to fail, the
    # allocation-free call succeeds and leaves the failure armed for the
    # allocating wrapper, which then fails
    print("noalloc wrapper: no allocation made", flush=True)
    with HipBuffer(s) as ds, HipBuffer(d) as dd, HipBuffer(w) as dw, \
            HipBuffer(np.zeros(V, np.float32)) as dist:
        assert lib.sssp_init_distances_f32(dist.ptr, V, 0) == S.OK
        assert lib.sssp_debug_fail_next_alloc() == S.OK
        assert lib.sssp_bellman_ford_noalloc_f32(ds.ptr, dd.ptr, dw.ptr, V, len(s), 0, dist.ptr,
                                                 None) == S.OK, S.last_error()
        assert np.array_equal(bits(dist.get()), bits(got))
        assert lib.sssp_bellman_ford_f32(ds.ptr, dd.ptr, dw.ptr, V, len(s), 0, dist.ptr,
                                         None) == S.ERR_OOM
