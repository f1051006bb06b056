"""bfg_matrix_device_noalloc: the allocation-free matrix call against
bfg_matrix_device on the same inputs, through the C ABI, bit-exact; two calls
in a row, the second smaller; bounded and unbounded; requests beyond its
limits are refused with BFG_NOT_SUPPORTED; the library's own count of device
allocations does not change across the calls (the device's free memory may
still move with the HIP runtime's internal pools, so it is only bounded)."""
import ctypes as C

import numpy as np
import pytest

from conftest import grid_graph
from oracle.oracle import INF, Oracle

pytestmark = pytest.mark.gpu

BFG_NOT_SUPPORTED = 4


def say(*a):
    print("[gpu-test]", *a, flush=True)


def run_both(e, hip, src, tgt, thr):
    from bfgpu.engine import _check
    S, T = src.size, tgt.size
    d_src, d_tgt = hip.DeviceArray.from_host(src), hip.DeviceArray.from_host(tgt)
    d_a, d_b = hip.DeviceArray(S * T, np.uint32), hip.DeviceArray(S * T, np.uint32)
    _check(e._lib.bfg_matrix_device(e._h, d_src.ptr, S, d_tgt.ptr, T, thr, d_a.ptr))
    hip.sync()
    free0, _ = hip.mem_info()
    n0, n1 = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    _check(e._lib.bfg_device_allocations(n0.ctypes.data_as(C.POINTER(C.c_uint64))))
    rc = e._lib.bfg_matrix_device_noalloc(e._h, d_src.ptr, S, d_tgt.ptr, T, thr, d_b.ptr)
    hip.sync()
    _check(e._lib.bfg_device_allocations(n1.ctypes.data_as(C.POINTER(C.c_uint64))))
    free1, _ = hip.mem_info()
    _check(rc)
    assert n1[0] == n0[0], "the allocation-free call made a device allocation"
    return d_a.to_host(), d_b.to_host(), free0, free1


@pytest.mark.parametrize("rows, chords", [(224, 0.0), (64, 0.02)])
def test_noalloc_matrix_equals_device_matrix(rows, chords):
    from bfgpu import BatchedPhastEngine, hip
    g = grid_graph(rows, chord_frac=chords)
    rng = np.random.default_rng(rows)
    o = Oracle(g)
    say("noalloc", rows, "n", g.n)
    with BatchedPhastEngine(g) as e:
        # two calls in a row, the second with fewer sources and targets; then bounded
        cases = [(64, 700, INF), (9, 40, INF), (50, 333, 1200), (1, 1, INF)]
        for S, T, thr in cases:
            src = rng.integers(0, g.n, S).astype(np.uint32)
            tgt = rng.integers(0, g.n, T).astype(np.uint32)
            want, got, free0, free1 = run_both(e, hip, src, tgt, thr)
            assert np.array_equal(got, want), (S, T, thr)
            exp = o.compute_matrix_flat_mt(src, tgt)
            if thr != INF:
                exp = np.where(exp > thr, INF, exp)
            assert np.array_equal(got, exp)
            # the runtime's own pools (kernel arguments, staging) may grow by a chunk
            assert free1 >= free0 - (16 << 20), "the allocation-free call took device memory"


def test_noalloc_refuses_requests_beyond_its_limits():
    from bfgpu import BatchedPhastEngine, hip
    g = grid_graph(24)
    with BatchedPhastEngine(g) as e:
        src = hip.DeviceArray.from_host(np.zeros(65, dtype=np.uint32))
        tgt = hip.DeviceArray.from_host(np.zeros(4, dtype=np.uint32))
        out = hip.DeviceArray(65 * 4, np.uint32)
        rc = e._lib.bfg_matrix_device_noalloc(e._h, src.ptr, 65, tgt.ptr, 4, INF, out.ptr)
        assert rc == BFG_NOT_SUPPORTED
        assert b"64 sources" in e._lib.bfg_last_error()
        # a valid call right after the refused one still works
        rc = e._lib.bfg_matrix_device_noalloc(e._h, src.ptr, 3, tgt.ptr, 4, INF, out.ptr)
        assert rc == 0
        assert np.array_equal(out.to_host(0, 12), np.zeros(12, dtype=np.uint32))
