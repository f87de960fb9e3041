"""jet_edge_cut_noalloc (no allocation at all) against jet_edge_cut on the
same device arrays: equal cuts on two calls in a row, the second on a smaller
graph, plus its rejection of invalid graphs. Device memory for the test's own
arrays comes from the HIP runtime through ctypes."""
import ctypes
import sys

import numpy as np
import pytest

import graphs
import jet_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.fixture
def hip():
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    h = Hip()
    yield h
    h.free()


def cuts(hip, g, part, unit=False):
    n = len(g["rowptr"]) - 1
    rp = hip.put(g["rowptr"].astype(np.int64))
    col = hip.put(g["col"].astype(np.int32))
    ew = None if unit else hip.put(g["ew"].astype(np.int32))
    pp = hip.put(part.astype(np.int32))
    a = jet_amd.edge_cut2_noalloc(n, rp, col, ew, pp)
    b = jet_amd.edge_cut2_device(n, rp, col, ew, pp)
    return a, b


def test_noalloc_cut_equals_edge_cut_twice_second_smaller(hip):
    rs = np.random.default_rng(3)
    big = graphs.stencil3d(40, 40, 30, seed=1)
    small = graphs.with_weights(graphs.rmat(10, 6, seed=2), seed=3, wmax=9)
    for name, g, unit in (("stencil 40x40x30", big, False), ("rmat-10", small, False), ("rmat-10 unit", small, True)):
        n = len(g["rowptr"]) - 1
        part = rs.integers(0, 17, n).astype(np.int32)
        say("noalloc vs jet_edge_cut on", name, "n", n)
        a, b = cuts(hip, g, part, unit)
        ref = g if not unit else dict(g, ew=np.ones(len(g["col"]), dtype=np.int32))
        assert a == b == jet_amd.edge_cut2(ref, part)


def test_noalloc_rejects_invalid_graphs(hip):
    g = graphs.grid2d(8, 8)
    n = 64
    part = np.zeros(n, dtype=np.int32)
    bad_col = dict(g, col=g["col"].copy())
    bad_col["col"][5] = 64
    bad_w = dict(g, ew=g["ew"].copy())
    bad_w["ew"][3] = 0
    for name, gg in (("column out of range", bad_col), ("zero weight", bad_w)):
        say("noalloc rejects", name)
        rp = hip.put(gg["rowptr"].astype(np.int64))
        col = hip.put(gg["col"].astype(np.int32))
        ew = hip.put(gg["ew"].astype(np.int32))
        pp = hip.put(part)
        with pytest.raises(jet_amd.JetError) as e:
            jet_amd.edge_cut2_noalloc(n, rp, col, ew, pp)
        assert e.value.code == jet_amd.JET_EINVAL
    # a valid call afterwards starts from cleared accumulators
    a, b = cuts(hip, g, (np.arange(n) % 3).astype(np.int32))
    assert a == b
